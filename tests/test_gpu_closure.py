"""GPU checks of the learned closure and its tools, through the C ABI, against the
oracle: the MFMA network, the fused learned right-hand side, the training gradient
and Adam step, the FP64 velocity gradient, the 2:1 filter, the subgrid samples and
the equispaced interpolation."""
import math
import os

import numpy as np
import pytest

from cmdg import abi, atmos, closure
from cmdg.grid import uniform_brick
from oracle import mlp_oracle as MO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _net(seed=5, y0=-2.5, shift=None, scale=None):
    rng = np.random.default_rng(seed)
    mlp = closure.MLPParams.random(seed=seed, y0=y0, shift=shift, scale=scale)
    mlp.flat[closure.OFF_B1:closure.OFF_W2] = rng.normal(0, 0.3, 32)
    mlp.flat[closure.OFF_B2:closure.OFF_W3] = rng.normal(0, 0.3, 32)
    return mlp


def _tgv_case(n=3, seed=0):
    prm = atmos.ideal_gas_parameters()
    L = 1.0
    grid = uniform_brick((n, n, n), (0, 0, 0), (2 * np.pi * L,) * 3, N=4)
    return prm, grid


def _learned_model(grid, prm, mlp, C_smag=0.21, turbulence=abi.LEARNED_CLOSURE, precision=abi.MIXED):
    from cmdg.dg import DGModel
    cfg = atmos.AtmosConfig(prm, turbulence=turbulence, C_smag=C_smag)
    aux = atmos.init_aux(cfg, grid)
    dg = DGModel(cfg, grid, aux=aux, precision=precision)
    if turbulence == abi.LEARNED_CLOSURE:
        dg.set_closure(mlp)
    return cfg, aux, dg


def test_mfma_network_matches_oracle():
    import torch
    print("[closure] network", flush=True)
    prm, grid = _tgv_case(2)
    rng = np.random.default_rng(11)
    mlp = _net(shift=rng.normal(0, 0.1, 10), scale=rng.uniform(0.5, 2.0, 10))
    cfg, aux, dg = _learned_model(grid, prm, mlp)
    orc = O.Oracle(cfg, grid, aux)
    orc.set_mlp(mlp)
    for n in (1, 77, 128, 1000, 4099):
        raw = (rng.standard_normal((n, 10)) * 2).astype(np.float32)
        y = dg.closure_forward(torch.from_numpy(raw).cuda()).cpu().numpy().astype(np.float64)
        ref = orc.mlp_forward(raw.astype(np.float64))
        err = np.abs(y - ref)
        print(f"  n={n}: max |dy| {err.max():.2e}", flush=True)
        # FP32 accumulation order and exp approximation, plus rare BF16 rounding
        # flips of a hidden unit (each moves y by about 2^-8 |w2 w3 h1|)
        # measured: max 2.5e-6 over 5300 rows
        assert err.max() < 1e-3 and np.median(err) < 1e-6, (n, err.max(), np.median(err))
    dg.close()


def _tgv_state(prm, grid, seed=0):
    Q = atmos.taylor_green(prm, grid, mach=0.1)
    rng = np.random.default_rng(seed)
    Q[:, 1:4, :] *= 1 + 0.05 * rng.standard_normal(Q[:, 1:4, :].shape)
    return Q


def test_learned_rhs_matches_oracle():
    import torch
    from cmdg.dg import stress_form_layout, to_device, to_host
    print("[closure] learned rhs", flush=True)
    prm, grid = _tgv_case(3)
    Q0 = _tgv_state(prm, grid)
    # inputs of order one: velocity gradients of this flow are O(100) 1/s; outputs
    # inside the clamp and nu_0 positive (the network varies by a few 1e-2 about 0.03)
    mlp = _net(seed=7, y0=0.03, scale=np.r_[np.full(9, 0.01), 10.0])
    mlp.flat[closure.OFF_W3:closure.OFF_B3] *= 0.05
    cfg, aux, dg = _learned_model(grid, prm, mlp)
    orc = O.Oracle(cfg, grid, aux)
    orc.set_mlp(mlp)
    ref = orc.tendency(Q0)
    Q = to_device(Q0)
    dQ = torch.full_like(Q, float("nan"))
    dg(dQ, Q, None, 0.0)
    got = to_host(dQ)
    errs = [rel_l2(got[:, s], ref[:, s]) for s in range(5)]
    print("  rel L2 per component: " + " ".join(f"{r:.2e}" for r in errs), flush=True)
    gfl = to_host(dg.state_gradient_flux_f32)
    nu_gpu = gfl[:, stress_form_layout(gfl.shape[1])["nu"][0]]
    nu_ref = orc.gradient_flux()[:, 19]
    assert (nu_ref > 0).all()
    nerr = np.abs(nu_gpu - nu_ref) / nu_ref
    print(f"  nu_0: median rel {np.median(nerr):.2e}, max rel {nerr.max():.2e}", flush=True)
    # measured: nu_0 median 1.9e-7, max 1.3e-6; tendency <= 3.0e-6 (energy)
    assert np.median(nerr) < 1e-5 and nerr.max() < 1e-3
    for r in errs:
        assert r < 2e-5
    dg.close()


def test_constant_network_matches_mixed_smagorinsky():
    import torch
    from cmdg.dg import to_device, to_host
    print("[closure] constant network = Smagorinsky", flush=True)
    prm, grid = _tgv_case(3)
    Q0 = _tgv_state(prm, grid, seed=1)
    Cs = 0.21
    _, _, dl = _learned_model(grid, prm, closure.MLPParams.constant(Cs ** 2), C_smag=Cs)
    _, _, ds = _learned_model(grid, prm, None, C_smag=Cs, turbulence=abi.SMAGORINSKY_LILLY)
    Q = to_device(Q0)
    a, b = torch.zeros_like(Q), torch.zeros_like(Q)
    dl(a, Q, None, 0.0)
    ds(b, Q, None, 0.0)
    A, B = to_host(a), to_host(b)
    errs = [rel_l2(A[:, s], B[:, s]) for s in range(5)]
    print("  rel L2 per component: " + " ".join(f"{r:.2e}" for r in errs), flush=True)
    for r in errs:
        assert r < 1e-6
    dl.close()
    ds.close()


def test_training_gradient_and_adam_match_oracle():
    import torch
    print("[closure] training gradient", flush=True)
    rng = np.random.default_rng(3)
    n, ld = 1000, 1500
    table = np.zeros((13, ld), np.float32)
    table[:10, :n] = rng.standard_normal((10, n)) * 3
    P = rng.uniform(0.1, 2.0, n)
    Qv = -rng.uniform(0, 1, n) * P
    table[10, :n], table[11, :n], table[12, :n] = P, Qv, Qv * Qv / P + rng.uniform(0, 0.5, n)
    mlp = _net(seed=9, y0=0.02, shift=rng.normal(0, 0.1, 10), scale=rng.uniform(0.3, 1.0, 10))
    mlp.flat[closure.OFF_W3:closure.OFF_B3] *= 0.1   # outputs inside the clamp: non-zero gradients
    tr = closure.ClosureTrainer(mlp, max_batch=4096)
    T = torch.from_numpy(table).cuda()
    idx = torch.from_numpy(rng.permutation(n)[:700].astype(np.int64)).cuda()
    for rows, batch, ix in ((np.arange(n), n, None), (idx.cpu().numpy(), 700, idx)):
        g = tr.loss_grad(T, batch, ix, 1.0 / batch).cpu().numpy().astype(np.float64)
        x = MO.normalise(table, mlp.shift, mlp.scale, rows)
        L, gref = MO.loss_grad(mlp.flat, x, table[10, rows], table[11, rows], table[12, rows], 1.0 / batch)
        gerr = np.linalg.norm(g[:1441] - gref) / np.linalg.norm(gref)
        print(f"  batch {batch}: loss {g[1441]:.6e} vs {L:.6e}, grad rel L2 {gerr:.2e}", flush=True)
        assert abs(g[1441] - L) < 1e-5 * abs(L)
        assert gerr < 1e-4
    # one Adam step against the textbook update
    p0 = tr.params.cpu().numpy().astype(np.float64)
    gq = tr.grad[:1441].cpu().numpy().astype(np.float64)
    tr.adam(1e-3)
    p1 = tr.params.cpu().numpy().astype(np.float64)
    m = 0.1 * gq
    v = 0.001 * gq * gq
    ref = p0 - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(p1, ref, rtol=1e-5, atol=1e-8)


def test_training_recovers_a_known_viscosity():
    """Samples whose optimal coefficient is 0.01 + 0.03 tanh(x0), negative (backscatter)
    for about a third of them: the trained network must come close to it."""
    import torch
    print("[closure] training convergence", flush=True)
    rng = np.random.default_rng(8)
    n = 200000
    table = np.zeros((13, n), np.float32)
    table[:10] = rng.standard_normal((10, n))
    target = 0.01 + 0.03 * np.tanh(table[0])
    P = rng.uniform(0.5, 2.0, n)
    table[10], table[11], table[12] = P, -target * P, target * target * P
    tr = closure.ClosureTrainer(closure.MLPParams.random(seed=1, y0=0.01), max_batch=8192)
    T = torch.from_numpy(table).cuda()
    hist = tr.fit(T, n, steps=1500, batch=8192, lr=3e-3, lr_final=1e-4, inv_norm=1.0 / float(np.mean(table[12])))
    print(f"  loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}", flush=True)
    m = tr.mlp()
    x = (table[:10, :5000].T - m.shift) * m.scale
    y = MO.forward(m.flat, x.astype(np.float64))
    err = np.abs(y - target[:5000])
    clear = np.abs(target[:5000]) > 0.01
    signs = float(np.mean(np.sign(y[clear]) == np.sign(target[:5000][clear])))
    print(f"  y vs target: median abs {np.median(err):.2e}, max {err.max():.2e}; sign right where "
          f"|target| > 0.01: {signs:.4f}", flush=True)
    # measured: median 2.4e-3 against a target spread (std) of 0.019
    assert hist[-1][1] < 0.05 * hist[0][1]
    assert np.median(err) < 4e-3
    assert signs > 0.97


def test_velocity_gradient():
    import torch
    from cmdg.dg import DGModel, to_device, to_host
    print("[closure] velocity gradient", flush=True)
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((3, 3, 2), (0, 0, 0), (300.0, 300.0, 200.0), N=4, periodicity=(True, True, False),
                         boundary=((0, 0), (0, 0), (1, 2)))
    cfg = atmos.AtmosConfig(prm, orientation=abi.FLAT_ORIENTATION, ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0),
                            gravity_source=True, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.23)
    aux = atmos.init_aux(cfg, grid)
    Q0 = atmos.rising_bubble(prm, grid, radius=100.0, centre=(150.0, 150.0, 100.0))
    Q0[:, 1:4, :] += np.random.default_rng(2).uniform(-1, 1, Q0[:, 1:4, :].shape)
    dg = DGModel(cfg, grid, aux=aux)
    Q = to_device(Q0)
    dQ = torch.zeros_like(Q)
    dg(dQ, Q, None, 0.0)
    gf = to_host(dg.state_gradient_flux)
    G = to_host(dg.velocity_gradient(Q))
    S = [G[:, 0], 0.5 * (G[:, 1] + G[:, 3]), 0.5 * (G[:, 2] + G[:, 6]), G[:, 4], 0.5 * (G[:, 5] + G[:, 7]), G[:, 8]]
    scale = np.abs(gf[:, 3:9]).max()
    for a in range(6):
        err = np.abs(S[a] - gf[:, 3 + a]).max() / scale
        assert err < 1e-13, (a, err)
    # a linear velocity field on a periodic-free brick: the DG gradient is exact
    grid2 = uniform_brick((2, 2, 2), (0, 0, 0), (1.0, 1.0, 1.0), N=4, periodicity=(False,) * 3,
                          boundary=((1, 1), (1, 1), (1, 1)))
    cfg2 = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.2)
    A = np.random.default_rng(3).standard_normal((3, 3))
    x = np.stack(grid2.coords, axis=1)            # (Np, 3, E)
    u = np.einsum("ab,nbe->nae", A, x)
    Q2 = np.zeros((grid2.Np, 5, grid2.nelem), order="F")
    Q2[:, 0] = 1.2
    Q2[:, 1:4] = 1.2 * u
    Q2[:, 4] = 2.5e5
    dg2 = DGModel(cfg2, grid2)
    G2 = to_host(dg2.velocity_gradient(to_device(Q2)))
    interior = np.ones(grid2.Np, bool)         # walls remove the normal velocity: compare off the boundary
    ii, jj, kk = grid2.node_ijk
    interior &= (ii > 0) & (ii < 4) & (jj > 0) & (jj < 4) & (kk > 0) & (kk < 4)
    for c in range(3):
        for d in range(3):
            assert np.allclose(G2[interior, 3 * c + d], A[c, d], atol=1e-10), (c, d)
    dg.close()
    dg2.close()


def _filter_numpy(Qf, fine, coarse):
    P = closure.filter_projection(fine.N)
    ch = closure.children_map(coarse, fine).reshape(-1, 8)
    Nq = fine.Nq
    r, m = Qf[:, 0], Qf[:, 1:4]
    mom = np.concatenate([Qf[:, :5], np.stack([m[:, a] * m[:, b] / r for a, b in
                                               ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))], 1)], 1)
    out = np.zeros((coarse.Np, 11, coarse.nelem))
    for e in range(coarse.nelem):
        for h in range(8):
            hx, hy, hz = h & 1, (h >> 1) & 1, h >> 2
            f = mom[:, :, ch[e, h]].reshape(Nq, Nq, Nq, 11)      # [k][j][i][field]
            g = np.einsum("ia,jb,kc,cbaf->kjif", P[hx], P[hy], P[hz], f)
            out[:, :, e] += g.reshape(coarse.Np, 11)
    return out


def test_box_filter_and_samples():
    import torch
    from cmdg.dg import DGModel, to_device, to_host
    print("[closure] filter and samples", flush=True)
    prm = atmos.ideal_gas_parameters()
    fine = uniform_brick((4, 4, 4), (0, 0, 0), (2 * np.pi,) * 3, N=4)
    coarse = uniform_brick((2, 2, 2), (0, 0, 0), (2 * np.pi,) * 3, N=4)
    Qf = _tgv_state(prm, fine, seed=4)
    Qd = to_device(Qf)
    M = closure.box_filter(Qd, fine, coarse)
    ref = _filter_numpy(Qf, fine, coarse)
    got = to_host(M)
    for f in range(11):      # (rho w vanishes identically in this flow)
        assert np.abs(got[:, f] - ref[:, f]).max() <= 1e-13 * max(1.0, np.abs(ref[:, f]).max()), f
    # coarse-polynomial fields pass unchanged: rho = 1, rho u linear per coarse element
    # is not continuous across elements, so use a global quadratic in x
    xf, xc = fine.coords[0], coarse.coords[0]
    Qp = np.zeros_like(Qf)
    Qp[:, 0] = 1.0
    Qp[:, 1] = 0.3 * xf ** 2 - xf
    Qp[:, 4] = 2.5e5
    Mp = to_host(closure.box_filter(to_device(Qp), fine, coarse))
    assert np.allclose(Mp[:, 1], 0.3 * xc ** 2 - xc, atol=1e-12)
    assert np.allclose(Mp[:, 5], (0.3 * xc ** 2 - xc) ** 2, atol=1e-11)
    # samples against a numpy restatement of the formula
    cfg = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.2)
    dg = DGModel(cfg, coarse)
    Qc = M[:, :5].contiguous()
    Gu = dg.velocity_gradient(Qc)
    Delta, nu = 0.3, 0.02
    n = coarse.nelem * coarse.Np
    table = torch.zeros((13, n + 7), dtype=torch.float32, device="cuda")
    closure.closure_samples(M, Gu, coarse, Delta, nu, table, 7)
    T = table.cpu().numpy()[:, 7:]
    Mh, Gh = got, to_host(Gu)
    r = Mh[:, 0].T.ravel()
    u = np.stack([Mh[:, a].T.ravel() / r for a in (1, 2, 3)])
    idx = {(0, 0): 5, (0, 1): 6, (0, 2): 7, (1, 1): 8, (1, 2): 9, (2, 2): 10}
    tau = np.zeros((3, 3, n))
    G = np.zeros((3, 3, n))
    for a in range(3):
        for b in range(3):
            tau[a, b] = Mh[:, idx[(min(a, b), max(a, b))]].T.ravel() / r - u[a] * u[b]
            G[a, b] = Gh[:, 3 * a + b].T.ravel()
    S = 0.5 * (G + G.transpose(1, 0, 2))
    I = np.eye(3)[:, :, None]
    Sd = S - np.trace(S)[None, None] / 3 * I
    Td = tau - np.trace(tau)[None, None] / 3 * I - 2 * (nu - 1e-5) * Sd
    c = Delta ** 2 * np.sqrt(2 * np.sum(S * S, axis=(0, 1)))
    a = 2 * c * Sd
    for row, val in ((10, np.sum(a * a, axis=(0, 1))), (11, np.sum(a * Td, axis=(0, 1))),
                     (12, np.sum(Td * Td, axis=(0, 1)))):
        assert rel_l2(T[row], val) < 1e-6, row
    assert np.allclose(T[:9], G.reshape(9, n), rtol=1e-6, atol=1e-6 * np.abs(G).max())
    assert np.all(T[9] == np.float32(Delta))
    # element-averaged targets: P, Q, R replaced by their element means
    table2 = torch.zeros((13, n), dtype=torch.float32, device="cuda")
    closure.closure_samples(M, Gu, coarse, Delta, nu, table2, 0, element_average=True)
    T2 = table2.cpu().numpy()
    assert np.array_equal(T2[:10], T[:10])
    for row in (10, 11, 12):
        mean = T[row].astype(np.float64).reshape(coarse.nelem, coarse.Np).mean(1)
        assert np.allclose(T2[row].reshape(coarse.nelem, coarse.Np), mean[:, None], rtol=1e-6,
                           atol=1e-6 * np.abs(mean).max()), row
    dg.close()


def test_equispaced_velocity():
    from cmdg.dg import to_device
    print("[closure] equispaced interpolation", flush=True)
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((2, 3, 2), (0, 0, 0), (2 * np.pi,) * 3, N=4)
    Q = _tgv_state(prm, grid, seed=5)
    p = 4
    out = closure.equispaced_velocity(to_device(Q), grid, p).cpu().numpy()
    L = closure.interpolation_matrix(grid.N, p)
    Nq = grid.Nq
    for e in range(grid.nelem):
        u = (Q[:, 1:4, e] / Q[:, 0:1, e]).reshape(Nq, Nq, Nq, 3)     # [k][j][i][c]
        v = np.einsum("xi,yj,zk,kjic->czyx", L, L, L, u)
        ix, iy, iz = grid.ex[e], grid.ey[e], grid.ez[e]
        blk = out[:, iz * p:(iz + 1) * p, iy * p:(iy + 1) * p, ix * p:(ix + 1) * p]
        assert np.allclose(blk, v, rtol=0, atol=1e-12 * np.abs(v).max())


def test_learned_closure_a_posteriori_configuration_4(tmp_path):
    """Configuration 4 at its own size: coarse 32^3-element runs (Mach 0.1, Re 1600, N = 4,
    to t* = 8, 8652 LSRK54 steps) from the filtered initial state against the filtered
    64^3-element plain FP64 run (molecular viscosity), whose kinetic energy and spectra at
    t* = 1, ..., 8 are stored in tests/golden/tgv_filtered_reference_64_32.json (written by
    the eval stage of cmdg.tgv_closure with --save-reference, 17,304 fine steps).  Models:
    the chosen form (mixed precision + the committed network, the molecular viscosity as
    its floor), the reference's Smagorinsky-Lilly (C_s = 0.21, mixed) and the molecular
    viscosity alone (FP64, the method without the network).
    The closure's filtered-KE error must be below Smagorinsky-Lilly's at t* = 1..7 and in
    the time average, and below the molecular-only run's at t* = 7, 8 (where the coarse mesh
    has lost the small scales and its upwind flux over-dissipates) and in the time average.
    Measured (relative KE error): learned 5.4e-5 .. 2.9e-3 through t* = 6, 1.6e-3 at 7, 9.4e-3
    at 8, mean 2.05e-3; Smagorinsky-Lilly 3.1e-3 .. 2.2e-2, 7.9e-3 at 8, mean 1.37e-2;
    molecular only 1.8e-6 .. 1.7e-3 through t* = 6, 6.0e-3 at 7, 1.33e-2 at 8, mean 2.73e-3.
    Not met: at t* = 8 Smagorinsky-Lilly's error (its excess energy, shrinking since t* = 6)
    is lower, and early the molecular-only run, nearly exact there, beats every model."""
    import json
    from cmdg import tgv_closure
    print("[closure] a-posteriori, configuration 4 (64^3 -> 32^3, t* <= 8)", flush=True)
    weights = os.path.join(os.path.dirname(closure.__file__), "data", "tgv_closure.npz")
    ref = os.path.join(os.path.dirname(__file__), "golden", "tgv_filtered_reference_64_32.json")
    tgv_closure.main(["--stage", "eval", "--fine", "64", "--coarse", "32", "--tend", "8",
                      "--eval-times", "1,2,3,4,5,6,7,8", "--weights", weights, "--models", "nets,smag,none",
                      "--reference", ref, "--out", str(tmp_path)])
    ev = json.load(open(tmp_path / "eval.json"))
    m = ev["models"]
    assert all(r["finite"] for r in m.values())
    le_all, no_all, sm_all = [], [], []
    for t in [f"{x:.1f}" for x in range(1, 9)]:
        le, sm, no = m["learned_tgv_closure"]["times"][t], m["smag"]["times"][t], m["none"]["times"][t]
        print(f"  t*={t}: KE err learned {le['ke_rel_err']:.3e}, smagorinsky {sm['ke_rel_err']:.3e}, "
              f"molecular only {no['ke_rel_err']:.3e}; spectrum learned {le['spectrum_rel_l2']:.3e}, "
              f"smagorinsky {sm['spectrum_rel_l2']:.3e}, molecular only {no['spectrum_rel_l2']:.3e}", flush=True)
        if float(t) <= 7:
            assert le["ke_rel_err"] < sm["ke_rel_err"], t
        if float(t) >= 7:
            assert le["ke_rel_err"] < no["ke_rel_err"], t
        le_all.append(le["ke_rel_err"])
        no_all.append(no["ke_rel_err"])
        sm_all.append(sm["ke_rel_err"])
    print(f"  time-averaged KE err: learned {np.mean(le_all):.3e}, smagorinsky {np.mean(sm_all):.3e}, "
          f"molecular only {np.mean(no_all):.3e}", flush=True)
    assert np.mean(le_all) < np.mean(no_all)
    assert np.mean(le_all) < np.mean(sm_all)

"""CPU checks of the learned-closure oracle and of the closure's host-side operators:
BF16 rounding, the training-loss gradient (finite differences), the C++ oracle's
network against an independent numpy evaluation, the constant network against
Smagorinsky-Lilly, the 2:1 projection filter, the equispaced interpolation and the
kinetic-energy spectrum binning of power_spectrum_les.jl."""
import math

import numpy as np
import pytest
import torch

from cmdg import abi, atmos, closure
from cmdg.grid import lglpoints, uniform_brick
from oracle import mlp_oracle as MO
from oracle import oracle as O


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10 ** rng.uniform(-6, 6, 20000),
                        # exact ties: bit 15 set, low 15 bits zero, both parities of bit 16
                        (np.arange(1, 2001, dtype=np.uint32) << 16 | 0x8000).view(np.float32)]).astype(np.float32)
    ours = closure.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)
    assert np.array_equal(MO.bf16(x).astype(np.float32), ref)


def _samples(rng, n):
    x = rng.standard_normal((n, 10))
    P = rng.uniform(0.1, 2.0, n)
    Q = -rng.uniform(0.0, 1.0, n) * P
    R = Q * Q / P + rng.uniform(0.0, 0.5, n)
    return x, P, Q, R


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = closure.MLPParams.random(seed=3, y0=0.02).flat.astype(np.float64)
    p[closure.OFF_W3:closure.OFF_B3] *= 0.2          # outputs inside the clamp (-e^-1, e^-1)
    x, P, Q, R = _samples(rng, 64)
    L, g = MO.loss_grad(p, x, P, Q, R, 0.5)
    y = MO.forward(p, x)
    assert np.all(np.abs(y) < 0.9 * MO.CMAX), "the finite differences must see the unclamped loss"
    assert abs(L - MO.loss(p, x, P, Q, R, 0.5)) < 1e-12 * abs(L)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(0, 320), 12, replace=False), np.arange(320, 324),
                          rng.choice(np.arange(352, 1376), 12, replace=False), np.arange(1376, 1380),
                          np.arange(1408, 1412), [1440]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (MO.loss(p + e, x, P, Q, R, 0.5) - MO.loss(p - e, x, P, Q, R, 0.5)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), (i, fd, g[i])


def test_loss_gradient_clamped_output_is_flat():
    rng = np.random.default_rng(2)
    x, P, Q, R = _samples(rng, 8)
    for y0 in (0.5, -0.5):          # beyond e^-1 on either side: clamped
        p = closure.MLPParams.constant(y0).flat.astype(np.float64)
        _, g = MO.loss_grad(p, x, P, Q, R)
        assert np.all(g == 0.0)


def _small_grid():
    return uniform_brick((2, 2, 2), (0, 0, 0), (1.0, 1.0, 1.0), N=4)


def test_oracle_network_matches_numpy():
    prm = atmos.ideal_gas_parameters()
    grid = _small_grid()
    cfg = atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, C_smag=0.21)
    orc = O.Oracle(cfg, grid, atmos.init_aux(cfg, grid))
    rng = np.random.default_rng(4)
    mlp = closure.MLPParams.random(seed=5, y0=-2.0, shift=rng.normal(0, 0.1, 10), scale=rng.uniform(0.5, 2, 10))
    mlp.flat[closure.OFF_B1:closure.OFF_W2] = rng.normal(0, 0.3, 32)
    mlp.flat[closure.OFF_B2:closure.OFF_W3] = rng.normal(0, 0.3, 32)
    orc.set_mlp(mlp)
    raw = rng.standard_normal((500, 10)) * 3
    y_c = orc.mlp_forward(raw)
    y_np = MO.forward_bf16(mlp.flat, mlp.shift, mlp.scale, raw)
    assert np.max(np.abs(y_c - y_np)) < 1e-12
    # and the BF16 emulation stays close to the exact network
    x = (raw.astype(np.float32) - mlp.shift) * mlp.scale
    assert np.max(np.abs(y_np - MO.forward(mlp.flat, x))) < 0.05


def test_constant_network_is_smagorinsky():
    """nu_0 = 1e-5 + C_s^2 Delta^2 |S| = |S| (C_s Delta)^2 + 1e-5."""
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((2, 2, 2), (0, 0, 0), (300.0, 300.0, 300.0), N=4, periodicity=(True, True, False),
                         boundary=((0, 0), (0, 0), (1, 2)))
    kw = dict(orientation=abi.FLAT_ORIENTATION, ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0),
              gravity_source=True, C_smag=0.23)
    cs = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, **kw)
    cl = atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, **kw)
    aux = atmos.init_aux(cs, grid)
    assert np.array_equal(aux, atmos.init_aux(cl, grid))
    Q = atmos.rising_bubble(prm, grid, radius=100.0, centre=(150.0, 150.0, 100.0))
    Q[:, 1:4, :] += np.random.default_rng(6).uniform(-1, 1, Q[:, 1:4, :].shape)
    ref = O.Oracle(cs, grid, aux).tendency(Q)
    ol = O.Oracle(cl, grid, aux)
    ol.set_mlp(closure.MLPParams.constant(0.23 ** 2))
    got = ol.tendency(Q)
    # the only difference is y = C_s^2 held in FP32 (relative 1e-7 in nu_0), and
    # the viscous part is a small fraction of the tendency
    for s in range(5):
        rel = np.linalg.norm(got[:, s] - ref[:, s]) / np.linalg.norm(ref[:, s])
        assert rel < 1e-8, (s, rel)
    gl, gs = ol.gradient_flux(), O.Oracle(cs, grid, aux)
    gs.tendency(Q)
    gs = gs.gradient_flux()
    assert np.allclose(gl[:, :10], gs, rtol=0, atol=1e-12 * np.abs(gs).max())
    S = gs[:, 3:9]
    normS = np.sqrt(2 * (S[:, 0] ** 2 + S[:, 3] ** 2 + S[:, 5] ** 2 + 2 * (S[:, 1] ** 2 + S[:, 2] ** 2 + S[:, 4] ** 2)))
    Delta = aux[:, cs.layout["Delta"], :]
    nu0 = normS * (0.23 * Delta) ** 2 + 1e-5
    assert np.allclose(gl[:, 19], nu0, rtol=3e-7, atol=0)
    # with a kinematic viscosity in the descriptor the closure's floor is that viscosity
    cv = atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, viscosity=0.3, **kw)
    ov = O.Oracle(cv, grid, aux)
    ov.set_mlp(closure.MLPParams.constant(0.23 ** 2))
    ov.tendency(Q)
    assert np.allclose(ov.gradient_flux()[:, 19], normS * (0.23 * Delta) ** 2 + np.float32(0.3), rtol=3e-7, atol=0)
    # the velocity gradient the network sees: its symmetric part is S
    G = gl[:, 10:19]
    assert np.allclose(0.5 * (G[:, 1] + G[:, 3]), S[:, 1], rtol=0, atol=1e-12 * np.abs(S).max())


@pytest.mark.parametrize("N", [3, 4])
def test_filter_projection_reproduces_coarse_polynomials(N):
    xi, _ = lglpoints(N)
    P = closure.filter_projection(N)
    rng = np.random.default_rng(N)
    c = rng.standard_normal(N + 1)
    f = lambda s: np.polynomial.polynomial.polyval(s, c)
    # fine nodes of the two halves in coarse coordinates
    got = P[0] @ f((xi - 1) / 2) + P[1] @ f((xi + 1) / 2)
    assert np.allclose(got, f(xi), rtol=0, atol=1e-13 * np.abs(f(xi)).max())
    # a function the coarse space cannot hold: its projection preserves the mean
    g = lambda s: np.abs(s) ** 3 + np.sin(3 * s)
    xg, wg = np.polynomial.legendre.leggauss(20)
    L = closure.lagrange_matrix(xi, xg)
    coarse = P[0] @ g((xi - 1) / 2) + P[1] @ g((xi + 1) / 2)
    mean_c = np.sum(wg * (L @ coarse))
    mean_f = sum(np.sum(wg / 2 * (L @ g((xi + s) / 2))) for s in (-1, 1))
    assert abs(mean_c - mean_f) < 1e-13


def test_children_map_covers_fine_brick():
    fine = uniform_brick((4, 6, 2), (0, 0, 0), (1, 1, 1))
    coarse = uniform_brick((2, 3, 1), (0, 0, 0), (1, 1, 1))
    ch = closure.children_map(coarse, fine).reshape(-1, 8)
    assert sorted(ch.ravel().tolist()) == list(range(fine.nelem))
    xf, xc = fine.coords[0], coarse.coords[0]
    for e in range(coarse.nelem):
        for h in range(8):
            hx = h & 1
            lo = xf[:, ch[e, h]].min()
            mid = (xc[:, e].min() + xc[:, e].max()) / 2
            assert (lo >= mid - 1e-12) if hx else (lo < mid)


def test_interpolation_matrix():
    for N in (3, 4):
        for p in (1, 3, 4, 5):
            L = closure.interpolation_matrix(N, p)
            assert np.allclose(L.sum(1), 1.0, atol=1e-14)
            xi, _ = lglpoints(N)
            xm = -1 + (2 * np.arange(p) + 1) / p
            assert np.allclose(L @ xi ** N, xm ** N, atol=1e-13)


def _spectrum_loops(u, v, w, L, dim, nor):
    """power_spectrum_les.jl:13-75 restated loop for loop (1-based indices)."""
    mu = np.abs(np.fft.fftn(u)) / u.shape[0] ** 3
    mv = np.abs(np.fft.fftn(v)) / v.shape[0] ** 3
    mw = np.abs(np.fft.fftn(w)) / w.shape[0] ** 3
    if dim % 2 == 0:
        rx = np.arange(dim) - dim / 2 + 1
        R = np.roll(rx, dim // 2 + 1)
        k_nyq = dim / 2
    else:
        rx = np.arange(dim) - (dim - 1) / 2
        R = np.roll(rx, (dim + 1) // 2)
        k_nyq = (dim - 1) / 2
    r = np.sqrt(R[:, None, None] ** 2 + R[None, :, None] ** 2 + R[None, None, :] ** 2)
    dx = 2 * math.pi / L
    k = np.arange(1, int(k_nyq) + 1) * dx
    endk = len(k)
    spec = np.zeros(endk)
    cnt = np.zeros(endk)
    e = mu ** 2 + mv ** 2 + mw ** 2
    for N in range(2, int(k_nyq - 1) + 1):
        for i in range(dim):
            for j in range(dim):
                for l in range(dim):
                    if r[i, j, l] * dx <= (k[N] + k[N - 1]) / 2 and r[i, j, l] * dx > (k[N - 1] + k[N - 2]) / 2:
                        spec[N - 1] += e[i, j, l]
                        cnt[N - 1] += 1
    for i in range(dim):
        for j in range(dim):
            for l in range(dim):
                if r[i, j, l] * dx <= (k[1] + k[0]) / 2:
                    spec[0] += e[i, j, l]
                    cnt[0] += 1
                if r[i, j, l] * dx <= k[endk - 1] and r[i, j, l] * dx > (k[endk - 1] + k[endk - 2]) / 2:
                    spec[endk - 1] += e[i, j, l]
                    cnt[endk - 1] += 1
    return spec * 2 * math.pi * k ** 2 / (cnt * dx ** 3) / nor, k


@pytest.mark.parametrize("dim", [8, 9])
def test_power_spectrum_matches_reference_loops(dim):
    rng = np.random.default_rng(dim)
    u, v, w = (rng.standard_normal((dim, dim, dim)) for _ in range(3))
    L = 2 * math.pi
    s1, k1 = closure.power_spectrum_3d(u, v, w, L, dim, 1.0)
    s2, k2 = _spectrum_loops(u, v, w, L, dim, 1.0)
    assert np.allclose(k1, k2)
    assert np.allclose(s1, s2, rtol=1e-12, atol=0)


def test_power_spectrum_single_mode():
    """u = cos(3x) on a 2 pi box puts all energy in shell 3."""
    dim = 16
    x = 2 * math.pi * np.arange(dim) / dim
    X = x[:, None, None] + 0 * x[None, :, None] + 0 * x[None, None, :]
    u = np.cos(3 * X)
    z = np.zeros_like(u)
    s, k = closure.power_spectrum_3d(u, z, z, 2 * math.pi, dim, 1.0)
    assert np.argmax(s) == 2 and np.count_nonzero(s > 1e-20) == 1

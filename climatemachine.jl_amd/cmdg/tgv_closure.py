"""Taylor-Green closure experiment: train the learned closure on a filtered fine run
and evaluate it by filtered kinetic-energy spectra.

Stage ``train``: a plain FP64 run on the fine brick (constant kinematic viscosity
nu = U0 L / Re, seeded phase shift of the initial vortex) is filtered 2:1 onto the
coarse brick at `--snapshots` times; every coarse node gives one sample of the
subgrid stress; the network is trained with the library's loss-gradient and Adam
kernels and saved.

Stage ``eval``: a plain fine run of the unshifted vortex gives the reference
(its filtered state); the coarse brick is run from the filtered initial state with
(a) the learned closure (mixed precision), (b) Smagorinsky-Lilly (mixed and FP64)
and (c) no closure (FP64, the molecular viscosity only) and (d) optionally the constant network
of the coefficient fitted to the training data (--fitted-cs); kinetic-energy spectra
(power_spectrum_3d on p points per element) and total kinetic energy are compared
with the filtered reference at the evaluation times.

    python -m cmdg.tgv_closure --stage train --fine 64 --coarse 32 --out DIR
    python -m cmdg.tgv_closure --stage eval --fine 64 --coarse 32 --weights DIR/closure.npz --out DIR
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np
import torch

from . import abi, atmos, closure
from .dg import DGModel, LSRK54CarpenterKennedy, to_device
from .grid import uniform_brick


def log(msg: str):
    print(msg, flush=True)


def setup(n: int, mach: float, re: float, L: float = 1.0):
    """The periodic (2 pi L)^3 brick of n^3 elements (lean: the library builds its element
    records from the break points, and nothing here reads per-node geometry)."""
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((n, n, n), (0, 0, 0), (2 * math.pi * L,) * 3, N=4, lean=True)
    c0 = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    U0 = mach * c0
    return prm, grid, U0, U0 * L / re


def model(prm, grid, kind: str, nu: float, mlp=None, C_smag: float = 0.21):
    """kind: fine / none (constant kinematic viscosity nu, FP64; fine_mixed the same in mixed
    precision), smag / smag64 (reference Smagorinsky-Lilly, mixed / FP64), net (the learned
    closure, mixed precision, with nu as its viscosity floor, running network `mlp`)."""
    if kind in ("fine", "none", "fine_mixed"):
        cfg = atmos.AtmosConfig(prm, turbulence=abi.CONSTANT_KINEMATIC_VISCOSITY, viscosity=nu,
                                with_divergence=True)
        return DGModel(cfg, grid, precision=abi.MIXED if kind == "fine_mixed" else abi.FP64)
    if kind in ("smag", "smag64"):
        cfg = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=C_smag)
        return DGModel(cfg, grid, precision=abi.MIXED if kind == "smag" else abi.FP64)
    if kind == "net":
        cfg = atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, C_smag=C_smag, viscosity=nu)
        dg = DGModel(cfg, grid, precision=abi.MIXED)
        dg.set_closure(mlp)
        return dg
    raise ValueError(kind)


def time_step(grid, U0: float, courant: float, prm) -> float:
    c0 = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    return courant * grid.min_node_distance() / (c0 + U0)


def run(dg, Q, dt: float, nsteps: int, callback=None, every: int = 0, label: str = ""):
    """nsteps fused LSRK54 steps; callback(step, t) every `every` steps and at the end."""
    solver = LSRK54CarpenterKennedy(dg, Q, dt=dt)
    t0 = time.time()
    last = t0
    for s in range(1, nsteps + 1):
        solver.dostep(Q, None, solver.t)
        solver.t += dt
        if callback and every and (s % every == 0 or s == nsteps):
            callback(s, solver.t)
        if time.time() - last > 30:
            torch.cuda.synchronize()
            last = time.time()
            log(f"[{label}] step {s}/{nsteps} t={solver.t:.4e} ({last - t0:.0f} s)")
    torch.cuda.synchronize()
    el = time.time() - t0
    dof = grid_dof(dg) * 5 * nsteps
    log(f"[{label}] {nsteps} steps in {el:.1f} s ({dof / el / 1e9:.2f} GDOF/s per stage)")
    return el


def grid_dof(dg) -> int:
    return dg.nelem * dg.Np * 5


def spectrum(Q, grid, p: int, L: float):
    u = closure.equispaced_velocity(Q, grid, p)
    dim = u.shape[1]
    E, k = closure.power_spectrum_3d(u[0], u[1], u[2], 2 * math.pi * L, dim, 1.0)
    ke = float(0.5 * (u * u).sum(0).mean().item())
    return E, k, ke


def stage_train(a):
    prm, fine, U0, nu = setup(a.fine, a.mach, a.re)
    _, coarse, _, _ = setup(a.coarse, a.mach, a.re)
    rng = np.random.default_rng(a.seed)
    phase = tuple(rng.uniform(0, 2 * math.pi, 3)) if a.phase else (0.0, 0.0, 0.0)
    log(f"[train] fine {a.fine}^3, coarse {a.coarse}^3, U0 {U0:.3f} m/s, nu {nu:.4e}, phase {phase}")
    dg = model(prm, fine, "fine", nu)
    dgc = model(prm, coarse, "smag64", nu)          # for the coarse velocity gradient only
    Q = to_device(atmos.taylor_green(prm, fine, mach=a.mach, phase=phase))
    dt = time_step(fine, U0, a.courant, prm)
    tend = a.tend / U0
    nsteps = int(math.ceil(tend / dt))
    dt = tend / nsteps
    every = max(1, nsteps // a.snapshots)
    nper = coarse.nelem * coarse.Np
    nsnap = nsteps // every + (1 if nsteps % every else 0)
    ld = nsnap * nper
    # two tables from the same run: pointwise and element-averaged targets; the
    # closure's floor is the molecular viscosity, so the network supplies the
    # subgrid stress alone
    tables = {v: torch.zeros((closure.SAMPLE_W, ld), dtype=torch.float32, device="cuda")
              for v in ("pointwise", "average")}
    Delta = float(atmos.lengthscale(coarse).ravel()[0])
    state = {"off": 0}

    def snap(step, t):
        M = closure.box_filter(Q, fine, coarse)
        Gu = dgc.velocity_gradient(M[:, :5].contiguous())
        off = state["off"]
        for v, tab in tables.items():
            state["off"] = closure.closure_samples(M, Gu, coarse, Delta, nu, tab, off,
                                                   element_average=(v == "average"), nu_floor=nu)

    log(f"[train] dt {dt:.4e} s, {nsteps} steps to t* = {a.tend}, snapshots every {every} steps, {ld} samples")
    el = run(dg, Q, dt, nsteps, snap, every, "fine")
    dg.close()
    n = state["off"]
    res = dict(stage="train", fine=a.fine, coarse=a.coarse, mach=a.mach, re=a.re, U0=U0, nu=nu, phase=phase,
               dt=dt, steps=nsteps, fine_seconds=el, samples=n, Delta=Delta, train_steps=a.train_steps,
               batch=a.batch, nu_floor=nu, models={})
    os.makedirs(a.out, exist_ok=True)
    for v, table in tables.items():
        tab = table[:, :n]
        rms = float(torch.sqrt((tab[:9].double() ** 2).mean()).item())
        mlp0 = closure.MLPParams.random(seed=a.seed, y0=0.0,
                                        shift=np.zeros(10), scale=np.r_[np.full(9, 1.0 / rms), 1.0 / Delta])
        P, Qv, R = (tab[10].double(), tab[11].double(), tab[12].double())
        c2 = float((-(Qv.sum()) / P.sum()).item())
        log(f"[train] {v}: {n} samples; grad-u rms {rms:.3e} 1/s; Delta {Delta:.4e} m; best constant exp(y) "
            f"{c2:.4e} (C_s {math.sqrt(max(c2, 0)):.3f})")
        tr = closure.ClosureTrainer(mlp0, max_batch=a.batch)
        inv_norm = 1.0 / float(R.mean().item())
        t0 = time.time()
        hist = tr.fit(table, n, steps=a.train_steps, batch=a.batch, lr=a.lr, lr_final=a.lr / 30,
                      inv_norm=inv_norm, seed=a.seed, log=log, log_every=max(1, a.train_steps // 10))
        torch.cuda.synchronize()
        t_train = time.time() - t0

        def const_loss(e):
            return float(((e * e * P + 2 * e * Qv + R).mean() / R.mean()).item())

        # loss of the trained network over the whole table, in batches
        tot = 0.0
        for off in range(0, n, a.batch):
            b = min(a.batch, n - off)
            idx = torch.arange(off, off + b, device="cuda", dtype=torch.int64)
            tot += float(tr.loss_grad(table, b, idx, inv_norm / n)[closure.NPARAM].item())
        mlp = tr.mlp()
        mlp.save(os.path.join(a.out, f"closure_{v}.npz"))
        res["models"][v] = dict(grad_rms=rms, train_seconds=t_train, history=hist, loss_trained=tot,
                                loss_const_best=const_loss(max(c2, 0.0)), loss_no_network=const_loss(0.0),
                                loss_smag_021=const_loss(0.21 ** 2), best_const_coeff=c2)
        log(f"[train] {v} done: loss {tot:.4e} (best constant {const_loss(max(c2, 0.0)):.4e}, floor only "
            f"{const_loss(0.0):.4e}, C_s=0.21 {const_loss(0.21 ** 2):.4e}), {t_train:.1f} s")
    with open(os.path.join(a.out, "train.json"), "w") as f:
        json.dump(res, f, indent=1)


def stage_eval(a):
    prm, fine, U0, nu = setup(a.fine, a.mach, a.re)
    _, coarse, _, _ = setup(a.coarse, a.mach, a.re)
    nets = {}
    for path in (a.weights.split(",") if a.weights else []):
        nets["learned_" + os.path.splitext(os.path.basename(path))[0].replace("closure_", "")] = \
            closure.MLPParams.load(path)
    for name in list(nets):
        for alpha in ([float(x) for x in a.scales.split(",")] if a.scales else []):
            # the network's output times alpha (its last layer scaled): a candidate of the
            # a-posteriori calibration against the filtered fine run
            m = closure.MLPParams(nets[name].flat, nets[name].shift, nets[name].scale)
            m.flat[closure.OFF_W3:closure.OFF_B3 + 1] *= alpha
            nets[f"{name}_x{alpha:g}"] = m
    if a.fitted_cs > 0:
        nets["fitted_const"] = closure.MLPParams.constant(a.fitted_cs ** 2)
    if "smag_floor" in a.models.split(","):
        nets["smag021_floor"] = closure.MLPParams.constant(0.21 ** 2)
    dt_f = time_step(fine, U0, a.courant, prm)
    if a.coarse_steps > 0:
        # a run of a.coarse_steps coarse steps (twice as many fine steps at the fine mesh's
        # Courant step); --eval-times are coarse step counts, reported in t*
        nc = a.coarse_steps
        nf = 2 * nc
        tend = nf * dt_f
        a.tend = tend * U0
        dt_c = tend / nc
        times = [int(x) * dt_c * U0 for x in a.eval_times.split(",")]
    else:
        tend = a.tend / U0
        nf = int(math.ceil(tend / dt_f))
        dt_f = tend / nf
        nc = nf // 2
        dt_c = tend / nc
        times = [float(x) for x in a.eval_times.split(",")]
    marks_f = {int(round(t / a.tend * nf)): t for t in times}
    marks_c = {int(round(t / a.tend * nc)): t for t in times}
    p = a.points
    log(f"[eval] fine {a.fine}^3 ({nf} steps), coarse {a.coarse}^3 ({nc} steps), t* {times}")
    Q0f = atmos.taylor_green(prm, fine, mach=a.mach)
    Q = to_device(Q0f)
    del Q0f
    Qc0 = closure.box_filter(Q, fine, coarse)[:, :5].contiguous()
    ref = {}
    if a.reference:
        # the filtered fine run's figures from an earlier --save-reference of the same setup
        with open(a.reference) as f:
            stored = json.load(f)
        for key in ("fine", "coarse", "mach", "re", "tend", "courant", "points", "fine_precision"):
            saved = stored.get(key, "fp64" if key == "fine_precision" else None)
            if saved != getattr(a, key):
                raise SystemExit(f"--reference was saved with {key} = {saved}, not {getattr(a, key)}")
        for t in times:
            r = stored["times"][str(t)]
            ref[t] = (np.asarray(r["spectrum"]), r["ke"])
        el_f = stored["fine_seconds"]
    else:
        dg = model(prm, fine, "fine_mixed" if a.fine_precision == "mixed" else "fine", nu)

        def snap_f(step, t):
            if step in marks_f:
                Qc = closure.box_filter(Q, fine, coarse)[:, :5].contiguous()
                E, k, ke = spectrum(Qc, coarse, p, 1.0)
                ref[marks_f[step]] = (E, ke)
                log(f"[eval] reference t*={marks_f[step]}: KE {ke:.5e}")
        gcd = math.gcd(*[s for s in marks_f]) if marks_f else 1
        el_f = run(dg, Q, dt_f, nf, snap_f, gcd, "fine")
        dg.close()
        if a.save_reference:
            with open(a.save_reference, "w") as f:
                json.dump(dict(fine=a.fine, coarse=a.coarse, mach=a.mach, re=a.re, tend=a.tend, courant=a.courant,
                               points=p, fine_precision=a.fine_precision, steps_fine=nf, fine_seconds=el_f,
                               times={str(t): dict(ke=ref[t][1], spectrum=[float(x) for x in ref[t][0]])
                                      for t in times}), f, indent=1)
    del Q
    torch.cuda.empty_cache()
    out = dict(stage="eval", fine=a.fine, coarse=a.coarse, U0=U0, nu=nu, tend=a.tend, dt_fine=dt_f,
               steps_fine=nf, steps_coarse=nc, fine_seconds=el_f, points_per_element=p, models={})
    kinds = [k for k in a.models.split(",") if k and k not in ("nets", "smag_floor")] + \
        ([k for k in nets if k != "smag021_floor"] if "nets" in a.models.split(",") else []) + \
        (["smag021_floor"] if "smag021_floor" in nets else [])
    for kind in kinds:
        dgc = model(prm, coarse, "net", nu, nets[kind]) if kind in nets else model(prm, coarse, kind, nu)
        Qc = Qc0.clone()
        res = {}

        def snap_c(step, t):
            if step in marks_c:
                E, k, ke = spectrum(Qc, coarse, p, 1.0)
                Er, ker = ref[marks_c[step]]
                kmax = len(Er)
                err = closure.spectrum_error(E, Er, kmax)
                res[str(marks_c[step])] = dict(
                    spectrum_rel_l2=err["rel_l2"], spectrum_log_rms=err["log_rms"], shells=err["shells"],
                    ke=ke, ke_ref=ker, ke_rel_err=abs(ke - ker) / ker,
                    spectrum=[float(x) for x in E], spectrum_ref=[float(x) for x in Er])
                log(f"[eval] {kind} t*={marks_c[step]}: spectrum rel L2 {err['rel_l2']:.4e}, log-rms "
                    f"{err['log_rms']:.4f} ({err['shells']} shells), KE rel err {abs(ke - ker) / ker:.3e}")
        gcd = math.gcd(*[s for s in marks_c])
        el = run(dgc, Qc, dt_c, nc, snap_c, gcd, kind)
        finite = bool(torch.isfinite(Qc).all().item())
        out["models"][kind] = dict(seconds=el, gdofs=coarse.nelem * coarse.Np * 5 * 5 * nc / el / 1e9,
                                   finite=finite, times=res)
        dgc.close()
    os.makedirs(a.out, exist_ok=True)
    with open(os.path.join(a.out, "eval.json"), "w") as f:
        json.dump(out, f, indent=1)
    for kind, r in out["models"].items():
        last = r["times"].get(str(times[-1]), {})
        log(f"[eval] {kind}: {r['gdofs']:.2f} GDOF/s, final spectrum rel L2 {last.get('spectrum_rel_l2')}, "
            f"log-rms {last.get('spectrum_log_rms')}, KE rel err {last.get('ke_rel_err')}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--stage", choices=("train", "eval"), required=True)
    ap.add_argument("--fine", type=int, default=64)
    ap.add_argument("--coarse", type=int, default=32)
    ap.add_argument("--mach", type=float, default=0.1)
    ap.add_argument("--re", type=float, default=1600.0)
    ap.add_argument("--tend", type=float, default=8.0, help="end time in units of L / U0")
    ap.add_argument("--courant", type=float, default=0.3,
                    help="acoustic Courant number on the smallest node gap (the LSRK54 limit here is between 0.3 and 0.5)")
    ap.add_argument("--snapshots", type=int, default=24)
    ap.add_argument("--phase", type=int, default=1, help="seeded phase shift of the training vortex")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--fitted-cs", type=float, default=0.0,
                    help="also evaluate the constant network of this coefficient (model 'fitted')")
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--models", default="nets,smag,smag64,none",
                    help="'nets' = every network given (--weights, --fitted-cs); 'smag_floor' = the constant "
                         "network of Smagorinsky C_s = 0.21 on top of the molecular floor")
    ap.add_argument("--eval-times", default="4,6,8")
    ap.add_argument("--points", type=int, default=4)
    ap.add_argument("--scales", default=None,
                    help="eval: also run every network with its output multiplied by each of these factors")
    ap.add_argument("--coarse-steps", type=int, default=0,
                    help="eval: run this many coarse steps (2x as many fine ones) instead of --tend; "
                         "--eval-times are then coarse step counts")
    ap.add_argument("--fine-precision", choices=("fp64", "mixed"), default="fp64",
                    help="eval: precision of the fine reference run (constant kinematic viscosity)")
    ap.add_argument("--save-reference", default=None,
                    help="eval: write the filtered fine run's kinetic energy and spectra at the evaluation times here")
    ap.add_argument("--reference", default=None,
                    help="eval: take the filtered fine run's figures from this file (written by --save-reference "
                         "with the same setup) instead of running it")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if not torch.cuda.is_available():
        raise SystemExit("the Taylor-Green experiment needs a GPU")
    if a.fine != 2 * a.coarse:
        raise SystemExit("--fine must be twice --coarse")
    (stage_train if a.stage == "train" else stage_eval)(a)


if __name__ == "__main__":
    main()

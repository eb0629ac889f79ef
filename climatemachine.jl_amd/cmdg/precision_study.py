"""Error of the mixed-precision path against the all-FP64 path on the GPU.

Runs the dry rising thermal bubble LES (gravity, hydrostatic reference state,
Smagorinsky-Lilly C_s = 0.23, walls in z, Rusanov) on an n^3-element brick from the
same seeded state with CMDG_FP64 and with CMDG_MIXED (optionally with the learned
closure) for K LSRK54 steps, and reports the relative L2 error of rho, rho u and
rho e (and of their change over the run) against the FP64 run, with the time per step
of each.

    python -m cmdg.precision_study --elements 64 --steps 200 --out DIR
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


def log(msg):
    print(msg, flush=True)


def workload(ne: int, learned: bool = False):
    prm = atmos.ideal_gas_parameters()
    L = 2500.0 * ne / 25.0          # the 25^3 bubble's element size, scaled out
    grid = uniform_brick((ne, ne, ne), (0.0, 0.0, 0.0), (L, L, L), N=4,
                         periodicity=(True, True, False), boundary=((0, 0), (0, 0), (1, 2)))
    kw = dict(orientation=abi.FLAT_ORIENTATION, ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0),
              gravity_source=True, C_smag=0.23, numerical_flux=abi.RUSANOV)
    cfg = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, **kw)
    cfgl = atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, **kw)
    aux = atmos.init_aux(cfg, grid)
    Q0 = atmos.rising_bubble(prm, grid, theta_ref=300.0, dtheta=2.0, radius=250.0,
                             centre=(L / 2, L / 2, 350.0), noise=1e-3, seed=1)
    c = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    dt = 0.3 * grid.min_node_distance() / c
    return grid, cfg, cfgl, aux, Q0, dt


def run(dg, Q, dt, steps, label):
    solver = LSRK54CarpenterKennedy(dg, Q, dt=dt)
    torch.cuda.synchronize()
    t0 = time.time()
    last = t0
    for s in range(steps):
        solver.dostep(Q, None, solver.t)
        solver.t += dt
        if time.time() - last > 20:
            torch.cuda.synchronize()
            last = time.time()
            log(f"[{label}] step {s + 1}/{steps}")
    torch.cuda.synchronize()
    el = time.time() - t0
    log(f"[{label}] {steps} steps, {el / steps * 1e3:.3f} ms/step")
    return el / steps


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--elements", type=int, default=64)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--learned", action="store_true", help="also run mixed + learned closure (constant network)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    t0 = time.time()
    grid, cfg, cfgl, aux, Q0, dt = workload(a.elements, a.learned)
    log(f"[study] {a.elements}^3 elements ({grid.nelem * grid.Np} nodes), dt {dt:.4g} s, host build "
        f"{time.time() - t0:.1f} s")
    res = dict(elements=a.elements, nodes=grid.nelem * grid.Np, steps=a.steps, dt=dt)
    dg = DGModel(cfg, grid, aux=aux, precision=abi.FP64)
    Q = to_device(Q0)
    res["fp64_ms_per_step"] = run(dg, Q, dt, a.steps, "fp64") * 1e3
    ref = Q.cpu()
    dg.close()
    del Q
    torch.cuda.empty_cache()
    q0 = to_device(Q0).cpu()
    legs = [("mixed", cfg, None)]
    if a.learned:
        legs.append(("mixed_learned_const", cfgl, closure.MLPParams.constant(0.23 ** 2)))
    for name, c, mlp in legs:
        dg = DGModel(c, grid, aux=aux, precision=abi.MIXED)
        if mlp is not None:
            dg.set_closure(mlp)
        Q = to_device(Q0)
        ms = run(dg, Q, dt, a.steps, name) * 1e3
        got = Q.cpu()
        dg.close()
        del Q
        torch.cuda.empty_cache()
        errs = {}
        for key, sl in (("rho", [0]), ("rho_u", [1, 2, 3]), ("rho_e", [4])):
            r, g, z = ref[:, sl].double(), got[:, sl].double(), q0[:, sl].double()
            errs[key] = dict(state=float(torch.linalg.norm(g - r) / torch.linalg.norm(r)),
                             change=float(torch.linalg.norm(g - r) / torch.linalg.norm(r - z)))
        res[name] = dict(ms_per_step=ms, rel_l2=errs, finite=bool(torch.isfinite(got).all()))
        log(f"[study] {name}: " + ", ".join(f"{k} {v['state']:.3e} (change {v['change']:.3e})"
                                          for k, v in errs.items()))
    os.makedirs(a.out, exist_ok=True)
    with open(os.path.join(a.out, f"precision_les{a.elements}.json"), "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()

"""Benchmark: GDOF/s per RK stage of the DG RHS + LSRK54 hot path on one MI355X.

Workload: the dry rising thermal bubble LES on a 25 x 25 x 25 element brick,
polynomial order 4 (1,953,125 nodes x 5 states = 9,765,625 DOF), FP64:
gravity, hydrostatic reference state (theta = 300 K), Smagorinsky-Lilly
C_s = 0.23, periodic x/y, free-slip insulating walls in z, Rusanov flux, one
LSRK54CarpenterKennedy step = 5 stages.  Synthetic seeded initial state.

Timed region: K steps, state already resident in HBM, bracketed by device
synchronisation.  GDOF/s per stage = DOF / (t_step / 5) / 1e9.

The dominant kernel's roofline uses per-kernel HIP-event durations recorded
on the library's stream during a second pass of the same K steps.
The CPU baseline (rank 0, N = 1) times the C++ oracle -- a restatement of
the reference algorithm -- on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "climatemachine.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# algorithmic HBM bytes per node for one launch of the LSRK-fused tendency kernel
# (FP64): Q read + dQ read + dQ write + Q write (4 x 5 x 8 = 160), gradient flux
# read (10 x 8 = 80), geometry xi (9) + MI (1) (80), aux rho_ref, p_ref, Phi (24)
TENDENCY_BYTES_PER_NODE = 160 + 80 + 80 + 24
# algorithmic bytes per node per full RK stage (both kernels), FP64 LES
STAGE_BYTES_PER_NODE = 424


def build_workload(ne=25):
    import numpy as np
    from cmdg import abi, atmos
    from cmdg.grid import uniform_brick
    prm = atmos.ideal_gas_parameters()
    L = 2500.0
    grid = uniform_brick((ne, ne, ne), (0.0, 0.0, 0.0), (L, L, L), N=4,
                         periodicity=(True, True, False), boundary=((0, 0), (0, 0), (1, 2)))
    cfg = atmos.AtmosConfig(prm, orientation=abi.FLAT_ORIENTATION,
                            ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0), gravity_source=True,
                            turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.23, numerical_flux=abi.RUSANOV)
    aux = atmos.init_aux(cfg, grid)
    Q0 = atmos.rising_bubble(prm, grid, theta_ref=300.0, dtheta=2.0, radius=250.0,
                             centre=(L / 2, L / 2, 350.0), noise=1e-3, seed=1)
    c = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    # Courant number 0.3, the reference's default for explicit LSRK54
    # (src/Driver/solver_configs.jl:219-228); 0.8 diverges at N = 4
    dt = 0.3 * grid.min_node_distance() / c
    return prm, grid, cfg, aux, Q0, dt


def cpu_baseline(grid, cfg, aux, Q0, dt, steps):
    """Oracle (reference algorithm restated in C++, OpenMP) on `steps` LSRK54 steps."""
    import numpy as np
    from oracle import oracle as O
    orc = O.Oracle(cfg, grid, aux)
    Q = np.asfortranarray(Q0.copy())
    dQ = np.zeros_like(Q, order="F")
    orc.lsrk54_step(Q, dQ, 0.0, dt)  # warm-up
    t0 = time.perf_counter()
    for k in range(steps):
        orc.lsrk54_step(Q, dQ, (k + 1) * dt, dt)
    el = time.perf_counter() - t0
    dof = grid.Np * 5 * grid.nelem
    return {"value": dof / (el / steps / 5) / 1e9, "unit": "GDOF/s", "cores": int(O.load().oracle_num_threads()),
            "kind": "port",
            "sample": f"{steps} LSRK54 steps (5 RHS stages each) of the same 25^3-element FP64 LES "
                      f"workload after 1 warm-up step, {el:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--elements", type=int, default=25)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by element "
              "blocks with one halo exchange per pass, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cmdg.dg import DGModel, LSRK54CarpenterKennedy, solve, to_device

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    prm, grid, cfg, aux, Q0, dt = build_workload(args.elements)
    from cmdg import abi
    dg = DGModel(cfg, grid, aux=aux, precision=abi.MIXED if args.precision == "mixed" else abi.FP64)
    Q = to_device(Q0)
    solver = LSRK54CarpenterKennedy(dg, Q, dt=dt, t0=0.0)
    print(f"[bench] {grid.nelem} elements, {grid.Np * grid.nelem} nodes, dt {dt:.4g} s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        solver.dostep(Q, None, solver.t)
        solver.t += dt
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.dostep(Q, None, solver.t)
        solver.t += dt
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not torch.isfinite(Q).all():
        print("[bench] non-finite state after the timed steps", file=sys.stderr)
        sys.exit(1)
    ms_step = el / args.steps * 1e3
    nodes = grid.Np * grid.nelem
    dof = nodes * 5
    gdofs = dof / (el / args.steps / 5) / 1e9
    print(f"[bench] {ms_step:.3f} ms/step, {gdofs:.2f} GDOF/s per stage", file=sys.stderr, flush=True)

    # per-kernel durations (HIP events on the library stream) over the same number of steps
    dg.set_kernel_timing(True)
    for _ in range(args.steps):
        solver.dostep(Q, None, solver.t)
        solver.t += dt
    kt = dg.kernel_timing()
    dg.set_kernel_timing(False)
    tend_ms, tend_n = kt["tendency"]
    grad_ms, grad_n = kt["gradient"]
    tend_avg = tend_ms / max(tend_n, 1)
    grad_avg = grad_ms / max(grad_n, 1)
    achieved = TENDENCY_BYTES_PER_NODE * nodes / (tend_avg * 1e-3) / 1e9
    print(f"[bench] k_tendency avg {tend_avg * 1e3:.1f} us ({tend_n} launches), "
          f"k_gradient avg {grad_avg * 1e3:.1f} us ({grad_n} launches)", file=sys.stderr, flush=True)

    result = {
        "metric": "GDOF/s per RK stage at N=4 (plain FP64 path)",
        "value": round(gdofs, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"dry rising thermal bubble LES, {args.elements}^3 elements, N=4, FP64, "
                        "Smagorinsky C_s=0.23, gravity + hydrostatic reference, walls in z, "
                        "Rusanov, LSRK54CarpenterKennedy",
            "elements": grid.nelem, "nodes": nodes, "dof": dof, "dt_s": dt,
        },
        "roofline": {
            "kernel": "k_tendency (LSRK-fused)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_node": TENDENCY_BYTES_PER_NODE,
            "avg_launch_us": round(tend_avg * 1e3, 2),
            "stage_achieved_GBs": round(STAGE_BYTES_PER_NODE * nodes / (el / args.steps / 5) / 1e9, 1),
            "gradient_avg_launch_us": round(grad_avg * 1e3, 2),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and args.cpu_steps > 0:
        print(f"[bench] CPU baseline: oracle, {args.cpu_steps} steps", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(grid, cfg, aux, Q0, dt, args.cpu_steps)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

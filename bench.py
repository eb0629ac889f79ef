"""Benchmark: GDOF/s per RK stage of the DG RHS + LSRK54 hot path on one MI355X,
plain FP64 against the mixed-precision path with the learned closure.

Workload (default): the dry rising thermal bubble LES on a 25 x 25 x 25 element
brick, polynomial order 4 (1,953,125 nodes x 5 states = 9,765,625 DOF): gravity,
hydrostatic reference state (theta = 300 K), periodic x/y, free-slip insulating
walls in z, Rusanov flux, one LSRK54CarpenterKennedy step = 5 stages, seeded
synthetic initial state.  Three legs on the same initial state and steps:
  * plain: FP64, Smagorinsky-Lilly C_s = 0.23 (the reference's method);
  * chosen form (the reported value): FP32 fluxes/derivatives/face fluxes with
    FP64 state, thermodynamics and reference subtraction, and the learned closure
    (BF16-MFMA network fused into the gradient kernel) in place of nu_0;
  * error statement: the same mixed-precision path with Smagorinsky-Lilly, whose
    state change after the timed steps is compared with the plain run's.
--workload tgv96 runs the decaying Taylor-Green vortex on 96^3 elements
(110.6 M nodes) instead: FP64 Smagorinsky-Lilly against mixed + learned.

Timed region: K steps as one solve!(Q, solver; numberofsteps = K) without
callbacks, state already resident in HBM, bracketed by device synchronisation.  GDOF/s per stage = DOF / (t_step / 5) / 1e9.  The roofline
of the chosen form's dominant kernel uses per-kernel HIP-event durations recorded
on the library's stream during a second pass of the same K steps; its HBM
traffic per launch comes from two rocprofv3 counter passes in child processes
(FETCH_SIZE, WRITE_SIZE, scaled by the counters of a 1 GiB copy).  The CPU
baseline (rank 0, N = 1) times the C++ oracle -- a restatement of the
reference algorithm -- on a bounded sample of the plain workload.
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
# Algorithmic HBM bytes per node of one launch:
# FP64 LSRK-fused tendency: Q read + dQ read + dQ write + Q write (4 x 5 x 8 = 160),
# gradient flux read (10 x 8 = 80), geometry xi (9) + MI (1) (80), aux rho_ref,
# p_ref, Phi (24).
TENDENCY_BYTES_FP64 = 160 + 80 + 80 + 24
STAGE_BYTES_FP64 = 424
# mixed LSRK-fused tendency: state 160 (FP64), gradient flux read 11 x 4 = 44,
# geometry 10 x 4 = 40, aux rho_ref, p_ref, Phi (FP64) 24; the gradient kernel:
# Q read 40, aux 24, gradient flux write 44 (FP32).
TENDENCY_BYTES_MIXED = 160 + 44 + 40 + 24
GRADIENT_BYTES_MIXED = 40 + 24 + 44
# the Taylor-Green workload has no Phi / reference state in aux
TENDENCY_BYTES_MIXED_TGV = 160 + 44 + 40
GRADIENT_BYTES_MIXED_TGV = 40 + 44
# Per node per RK stage, both kernels (state 160, geometry 40 in FP32, aux 24, the
# FP32 gradient flux of 11 words written and read 88); the TGV has no aux fields.
STAGE_BYTES_MIXED = 160 + 40 + 24 + 88
STAGE_BYTES_MIXED_TGV = 160 + 40 + 88
WEIGHTS = os.path.join(ROOT, "climatemachine.jl_amd", "cmdg", "data", "tgv_closure.npz")


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


def build_tgv(ne=96):
    from cmdg import abi, atmos
    from cmdg.grid import uniform_brick
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((ne, ne, ne), (0.0, 0.0, 0.0), (2 * math.pi,) * 3, N=4)
    cfgs = {k: atmos.AtmosConfig(prm, turbulence=t, C_smag=0.21)
            for k, t in (("smag", abi.SMAGORINSKY_LILLY), ("learned", abi.LEARNED_CLOSURE))}
    aux = atmos.init_aux(cfgs["smag"], grid)
    Q0 = atmos.taylor_green(prm, grid, mach=0.1)
    c = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    dt = 0.3 * grid.min_node_distance() / (1.1 * c)
    return prm, grid, cfgs, aux, Q0, dt


def closure_weights(les: bool):
    """The network the chosen form runs.  On the LES bubble: the network whose output
    is 2 log C_s everywhere (Smagorinsky-Lilly's nu_0 through the fused MLP, so that
    the state difference to the plain run measures the arithmetic of the chosen form
    alone; the MLP's cost does not depend on its weights).  On the Taylor-Green
    vortex: the network trained on the filtered 64^3 run, when present."""
    from cmdg import closure
    if not les and os.path.exists(WEIGHTS):
        return closure.MLPParams.load(WEIGHTS), "trained on the filtered 64^3 Taylor-Green run"
    return closure.MLPParams.constant(2 * math.log(0.23 if les else 0.21)), \
        "network with constant output 2 log C_s (Smagorinsky-Lilly's nu_0 through the fused MLP)"


def cpu_baseline(grid, cfg, aux, Q0, dt, steps, label):
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
    return {"value": round(dof / (el / steps / 5) / 1e9, 5), "unit": "GDOF/s",
            "cores": int(O.load().oracle_num_threads()), "kind": "port",
            "sample": f"{steps} LSRK54 steps (5 RHS stages each) of {label} after 1 warm-up step, {el:.1f} s wall"}


def timed(dg, Q, dt, steps, warmup):
    """warmup, then `steps` timed steps, each run as solve!(Q, solver; numberofsteps)
    without callbacks."""
    import torch
    from cmdg.dg import LSRK54CarpenterKennedy, solve
    solver = LSRK54CarpenterKennedy(dg, Q, dt=dt, t0=0.0)
    if warmup > 0:
        solve(Q, solver, numberofsteps=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solve(Q, solver, numberofsteps=steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not torch.isfinite(Q).all():
        print("[bench] non-finite state after the timed steps", file=sys.stderr)
        sys.exit(1)
    return el, solver


def kernel_pass(dg, Q, solver, steps):
    from cmdg.dg import solve
    dg.set_kernel_timing(True)
    solve(Q, solver, numberofsteps=steps)
    kt = dg.kernel_timing()
    dg.set_kernel_timing(False)
    return {k: (ms / max(n, 1), n) for k, (ms, n) in kt.items()}


def hbm_ceiling(gib=4, reps=10):
    """Practical HBM bandwidth of this GPU in GB/s: a 4 GiB FP64 read (sum), write
    (fill_) and copy (copy_), each timed with events over `reps` calls after one."""
    import torch
    n = (gib << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)

    def rate(f, nbytes):
        f()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(reps):
            f()
        t1.record()
        torch.cuda.synchronize()
        return nbytes / (t0.elapsed_time(t1) / reps * 1e-3) / 1e9

    out = {"read": round(rate(lambda: a.sum(), n * 8), 1), "write": round(rate(lambda: b.fill_(2.0), n * 8), 1),
           "copy": round(rate(lambda: b.copy_(a), 2 * n * 8), 1),
           "method": f"{gib} GiB FP64 torch sum / fill_ / copy_, {reps} timed calls each"}
    del a, b
    torch.cuda.empty_cache()
    return out


def hbm_traffic(kernel_prefix, workload, ne, timeout_s=150):
    """HBM bytes per launch of the kernels whose name starts with `kernel_prefix`,
    from two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch)
    over a short run of the chosen-form leg in a child process.  Each counter is
    scaled by what it reads for a plain 1 GiB device-to-device copy run first in the
    same child (1 GiB read and 1 GiB written).  Returns (bytes, detail) or
    (None, reason)."""
    import glob
    import shutil
    import signal
    import sqlite3
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    cmd_tail = ["--", sys.executable, os.path.abspath(__file__), "--only", "learned", "--steps", "2",
                "--warmup", "1", "--workload", workload, "--elements", str(ne),
                "--copy-calibration-mib", "1024", "--traffic", "off", "--no-cpu-baseline"]
    env = dict(os.environ, TMPDIR="/tmp")
    per = {}
    with tempfile.TemporaryDirectory(prefix="cmdg_pmc_", dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "-d", out, "-o", counter] + cmd_tail
            p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                 start_new_session=True)
            try:
                _, err = p.communicate(timeout=timeout_s)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return None, f"rocprofv3 --pmc {counter} did not finish in {timeout_s} s"
            dbs = glob.glob(os.path.join(out, "**", "*.db"), recursive=True)
            if p.returncode != 0 and not dbs:
                tail = err.decode(errors="replace").strip().splitlines()[-1:] if err else []
                return None, f"rocprofv3 --pmc {counter} exited {p.returncode}: {' '.join(tail)}"
            kern, calib = [], []
            for db in dbs:
                con = sqlite3.connect(db)
                rows = con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection "
                                   "order by dispatch_id").fetchall()
                con.close()
                for did, name, cname, value in rows:
                    if cname != counter:
                        continue
                    bare = name.replace("void ", "").replace("cmdg::", "")
                    if bare.startswith(kernel_prefix):
                        kern.append(float(value))
                    elif did <= 2:  # the calibration: a 1 GiB fill, then the 1 GiB copy
                        calib.append(float(value))
            if not kern or not calib or max(calib) <= 0:
                return None, f"no {counter} records for {kernel_prefix} or for the calibration copy"
            scale = (1 << 20) / max(calib)  # KiB counted for the 1 GiB copy -> true bytes
            per[counter] = (sum(kern) / len(kern), scale, len(kern))
    fetch = per["FETCH_SIZE"][0] * per["FETCH_SIZE"][1] * 1024
    write = per["WRITE_SIZE"][0] * per["WRITE_SIZE"][1] * 1024
    detail = {"fetch_bytes": round(fetch), "write_bytes": round(write),
              "fetch_scale": round(per["FETCH_SIZE"][1], 4), "write_scale": round(per["WRITE_SIZE"][1], 4),
              "dispatches": per["FETCH_SIZE"][2],
              "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch), averaged over the kernel's "
                        "dispatches in 3 steps of the chosen form, each scaled by the same counter on a 1 GiB "
                        "device copy"}
    return fetch + write, detail


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # at least 100 timed steps after 10 warm-up steps, the measurement the GDOF/s figure is
    # defined on
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=("les25", "tgv96"), default="les25")
    ap.add_argument("--elements", type=int, default=0, help="elements per direction (default 25 / 96)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=("all", "fp64", "mixed", "learned"), default="all",
                    help="profile one leg alone")
    ap.add_argument("--traffic", choices=("on", "off"), default="on",
                    help="measure the dominant kernel's HBM bytes per launch with two rocprofv3 counter passes "
                         "(child processes; use off when the bench itself runs under a profiler)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by element "
              "blocks with one halo exchange per pass, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cmdg import abi
    from cmdg.dg import DGModel, to_device, to_host

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    les = args.workload == "les25"
    ne = args.elements or (25 if les else 96)
    t_build = time.perf_counter()
    if les:
        prm, grid, cfg, aux, Q0, dt = build_workload(ne)
        from cmdg import atmos
        cfg_plain = cfg
        cfg_learned = atmos.AtmosConfig(prm, orientation=abi.FLAT_ORIENTATION,
                                        ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0), gravity_source=True,
                                        turbulence=abi.LEARNED_CLOSURE, C_smag=0.23, numerical_flux=abi.RUSANOV)
        tend_bytes, grad_bytes, stage_bytes = TENDENCY_BYTES_MIXED, GRADIENT_BYTES_MIXED, STAGE_BYTES_MIXED
    else:
        prm, grid, cfgs, aux, Q0, dt = build_tgv(ne)
        cfg_plain, cfg_learned = cfgs["smag"], cfgs["learned"]
        tend_bytes, grad_bytes, stage_bytes = TENDENCY_BYTES_MIXED_TGV, GRADIENT_BYTES_MIXED_TGV, STAGE_BYTES_MIXED_TGV
    nodes = grid.Np * grid.nelem
    dof = nodes * 5
    print(f"[bench] {args.workload}: {grid.nelem} elements, {nodes} nodes, dt {dt:.4g} s "
          f"(host build {time.perf_counter() - t_build:.1f} s)", file=sys.stderr, flush=True)
    gd = lambda el: dof / (el / args.steps / 5) / 1e9
    out = {}

    # plain FP64 leg (and the reference state change for the error statement)
    if args.only in ("all", "fp64"):
        dg = DGModel(cfg_plain, grid, aux=aux, precision=abi.FP64)
        Q = to_device(Q0)
        el, _ = timed(dg, Q, dt, args.steps, args.warmup)
        out["fp64"] = gd(el)
        out["fp64_ms"] = el / args.steps * 1e3
        dQ_fp64 = (Q - to_device(Q0)).cpu().numpy()
        dg.close()
        del Q
        torch.cuda.empty_cache()
        print(f"[bench] plain FP64: {out['fp64_ms']:.3f} ms/step, {out['fp64']:.2f} GDOF/s per stage",
              file=sys.stderr, flush=True)
    # mixed precision with Smagorinsky-Lilly: the error of the precision change
    if args.only in ("all", "mixed"):
        dg = DGModel(cfg_plain, grid, aux=aux, precision=abi.MIXED)
        Q = to_device(Q0)
        el, _ = timed(dg, Q, dt, args.steps, args.warmup)
        out["mixed"] = gd(el)
        if args.only == "all":
            dQ_mix = (Q - to_device(Q0)).cpu().numpy()
            out["mixed_err"] = [float(np.linalg.norm(dQ_mix[:, s] - dQ_fp64[:, s]) / np.linalg.norm(dQ_fp64[:, s]))
                                for s in range(5)]
        dg.close()
        del Q
        torch.cuda.empty_cache()
        print(f"[bench] mixed + Smagorinsky: {out['mixed']:.2f} GDOF/s per stage", file=sys.stderr, flush=True)
    # chosen form: mixed precision + learned closure (and its state change vs plain FP64)
    mlp, mlp_note = closure_weights(les)
    if args.only in ("all", "learned"):
        dg = DGModel(cfg_learned, grid, aux=aux, precision=abi.MIXED)
        dg.set_closure(mlp)
        Q = to_device(Q0)
        el, solver = timed(dg, Q, dt, args.steps, args.warmup)
        out["learned"] = gd(el)
        out["learned_ms"] = el / args.steps * 1e3
        if args.only == "all":
            dQ_l = (Q - to_device(Q0)).cpu().numpy()
            out["learned_err"] = [float(np.linalg.norm(dQ_l[:, s] - dQ_fp64[:, s]) / np.linalg.norm(dQ_fp64[:, s]))
                                  for s in range(5)]
        kt = kernel_pass(dg, Q, solver, args.steps)
        dg.close()
        print(f"[bench] mixed + learned closure: {out['learned_ms']:.3f} ms/step, {out['learned']:.2f} GDOF/s "
              f"per stage; k_tendency {kt['tendency'][0] * 1e3:.1f} us, k_gradient {kt['gradient'][0] * 1e3:.1f} us",
              file=sys.stderr, flush=True)
    if args.only != "all":
        print(json.dumps({k: v for k, v in out.items()}), flush=True)
        return

    ceiling = hbm_ceiling()
    print(f"[bench] HBM ceiling on this GPU: {ceiling}", file=sys.stderr, flush=True)
    t_avg, g_avg = kt["tendency"][0], kt["gradient"][0]
    dom = ("k_tendency_mix (LSRK-fused, mixed)", t_avg, tend_bytes, "k_tendency_mix") if t_avg >= g_avg else \
          ("k_gradient_mix (learned closure fused)", g_avg, grad_bytes, "k_gradient_mix")
    achieved = dom[2] * nodes / (dom[1] * 1e-3) / 1e9
    traffic, traffic_detail = None, "not measured (--traffic off)"
    if args.traffic == "on":
        print(f"[bench] HBM traffic of {dom[3]}: two rocprofv3 counter passes", file=sys.stderr, flush=True)
        traffic, traffic_detail = hbm_traffic(dom[3], args.workload, ne)
        if traffic is not None:
            traffic_detail["algorithmic_bytes"] = dom[2] * nodes
            traffic_detail["ratio_to_algorithmic"] = round(traffic / (dom[2] * nodes), 3)
        print(f"[bench] traffic: {traffic} ({traffic_detail})", file=sys.stderr, flush=True)
    workload = (f"dry rising thermal bubble LES, {ne}^3 elements, N=4, gravity + hydrostatic reference, walls in z, "
                "Rusanov, LSRK54CarpenterKennedy" if les else
                f"Taylor-Green vortex, Mach 0.1, {ne}^3 elements, N=4, periodic, Rusanov, LSRK54CarpenterKennedy")
    result = {
        "metric": "GDOF/s per RK stage at N=4: plain FP64 vs FP32+learned closure, at stated error",
        "value": round(out["learned"], 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(out["learned_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fluxes / f64 state (bf16 MFMA closure)",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "elements": grid.nelem, "nodes": nodes, "dof": dof, "dt_s": dt,
            "plain": "FP64, Smagorinsky-Lilly",
            "chosen_form": "mixed precision + learned closure (" + mlp_note + ")",
        },
        "plain_fp64_value": round(out["fp64"], 4),
        "plain_fp64_ms_per_step": round(out["fp64_ms"], 4),
        "speedup_vs_plain_fp64": round(out["learned"] / out["fp64"], 3),
        "mixed_smagorinsky_value": round(out["mixed"], 4),
        "stated_error": {
            "what": f"relative L2 per state (rho, rho u1..3, rho e) of the state change after {args.warmup + args.steps} "
                    "steps against the plain FP64 run from the same state",
            "chosen_form": [float(f"{e:.3e}") for e in out["learned_err"]],
            "mixed_smagorinsky": [float(f"{e:.3e}") for e in out["mixed_err"]],
        },
        "roofline": {
            "kernel": dom[0],
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "measured_ceiling": ceiling,
            "frac_of_measured_read_ceiling": round(achieved / ceiling["read"], 4),
            "traffic": None if traffic is None else int(traffic),
            "traffic_rate_GBs": None if traffic is None else round(traffic / (dom[1] * 1e-3) / 1e9, 1),
            "traffic_detail": traffic_detail,
            "algorithmic_bytes_per_node": dom[2],
            "avg_launch_us": round(dom[1] * 1e3, 2),
            "k_tendency_avg_us": round(t_avg * 1e3, 2),
            "k_gradient_avg_us": round(g_avg * 1e3, 2),
            "stage": {
                "what": "algorithmic bytes of a whole RK stage (both kernels) over the stage time, ms_per_step / 5",
                "bytes_per_node": stage_bytes,
                "achieved": round(stage_bytes * nodes / (out["learned_ms"] / 5 * 1e-3) / 1e9, 1),
                "frac": round(stage_bytes * nodes / (out["learned_ms"] / 5 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            },
        },
        "cpu_baseline": None,
    }
    if les and not args.no_cpu_baseline and args.cpu_steps > 0:
        print(f"[bench] CPU baseline: oracle, {args.cpu_steps} steps", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(grid, cfg_plain, aux, Q0, dt, args.cpu_steps,
                                              f"the same {ne}^3-element FP64 LES workload")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

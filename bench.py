"""Benchmark: GDOF/s per RK stage of the DG RHS + LSRK54 hot path on one MI355X,
plain FP64 against the mixed-precision path with the learned closure.

Workload (default, --workload tgv96): the full chosen form at scale, the decaying
Taylor-Green vortex on a 96 x 96 x 96 element periodic 2 pi box, polynomial order 4
(110,592,000 nodes x 5 states = 552,960,000 DOF), Mach 0.1, Rusanov flux, one
LSRK54CarpenterKennedy step = 5 stages.  Three legs from the same initial state:
  * plain: FP64, Smagorinsky-Lilly C_s = 0.21 (the reference's method);
  * chosen form (the reported value): FP32 fluxes/derivatives/face fluxes with FP64
    state, thermodynamics and LSRK accumulation, and the learned closure (the network
    trained on the box-filtered 64^3-element Re 1600 run, BF16-MFMA inference fused
    into the gradient kernel) in place of nu_0, its viscosity floor the Re 1600
    molecular viscosity it was trained with;
  * mixed precision with Smagorinsky-Lilly: the error of the precision change alone.
--workload les25 runs the dry rising-bubble LES on 25^3 elements instead (gravity,
hydrostatic reference, walls in z; the network there has the constant output
C_s^2, Smagorinsky-Lilly's nu_0 through the fused MLP).

Timed region: K steps as one solve!(Q, solver; numberofsteps = K) without callbacks,
state already resident in HBM, bracketed by device synchronisation.  GDOF/s per
stage = DOF / (t_step / 5) / 1e9.  The roofline of the chosen form's kernels uses
per-kernel HIP-event durations recorded on the library's stream during a second pass
of the same K steps; HBM traffic and VALU / MFMA instruction counts per launch come
from two rocprofv3 counter passes in child processes (FETCH_SIZE with the FP32 and
MFMA counters, WRITE_SIZE with the FP64 counters; the byte counters scaled by what
they read for a 1 GiB copy).  The CPU baseline (rank 0, N = 1) times the C++ oracle --
a restatement of the reference algorithm -- on a bounded sample of the plain method.
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
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16_mfma": 2500.0}  # MI355X dense peaks
# Algorithmic HBM bytes per node of one launch of the chosen form's kernels:
# LSRK-fused tendency: Q read + dQ read + dQ write + Q write (4 x 5 x 8 = 160), FP32
# gradient flux in stress form read (the stress factor and the enthalpy diffusion: 9 x 4 =
# 36 with an isotropic viscosity, the Taylor-Green vortex; 12 x 4 = 48 with the buoyancy
# factor, the LES bubble), FP32 geometry xi (9) + MI (1) (40), aux rho_ref, p_ref, Phi
# (FP64, 24; absent on the Taylor-Green vortex); gradient kernel: Q read 40, aux 24 (LES),
# stress form written (10 x 4 = 40, or 15 x 4 = 60: with the viscosity planes).
# Whole stage: Q, dQ 160, geometry 40, aux 24 (LES), stress form written and read.
BYTES = {
    "les25": {"tendency": 160 + 48 + 40 + 24, "gradient": 40 + 24 + 60, "stage": 160 + 40 + 24 + 60 + 48},
    "tgv96": {"tendency": 160 + 36 + 40, "gradient": 40 + 40, "stage": 160 + 40 + 40 + 36},
}
# solve calls of the child counter runs: --warmup 1, then --steps 2 timed and 2 timed per kernel
COUNTER_CALLS = (1, 2, 2)


def lsrk_bytes(full, calls):
    """Mean algorithmic bytes per node of the LSRK-fused tendency launch (or of a whole
    stage) over the stages of solve calls of the given step counts: between two steps of
    one call the last stage leaves the zeros of dQ *= 0 unstored and the next step's
    first stage does not read them, 40 B a node fewer each."""
    stages = 5 * sum(calls)
    return full - 40.0 * sum(2 * (n - 1) for n in calls) / stages
WEIGHTS = os.path.join(ROOT, "climatemachine.jl_amd", "cmdg", "data", "tgv_closure.npz")
TRAIN_META = os.path.join(ROOT, "climatemachine.jl_amd", "cmdg", "data", "tgv_closure_train.json")
COUNTS = os.path.join(ROOT, "climatemachine.jl_amd", "cmdg", "data", "algorithmic_counts.json")
# counters of the two passes (rocprofv3 collects at most 8 SQ and 4 TCC counters a run)
# (the SQ_INSTS_VALU_FLOPS_* counters are not used: they count ADD + MUL + TRANS + 2 FMA per
# wave instruction, 1/64 of the lanes' flops, which the instruction counters already give)
PASSES = (("FETCH_SIZE", "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_FMA_F32",
           "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_WAVES"),
          ("WRITE_SIZE", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
           "SQ_INSTS_VALU_TRANS_F64", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "SQ_INSTS_VALU"))


def build_les(ne=25):
    from cmdg import abi, atmos
    from cmdg.grid import uniform_brick
    prm = atmos.ideal_gas_parameters()
    L = 2500.0
    grid = uniform_brick((ne, ne, ne), (0.0, 0.0, 0.0), (L, L, L), N=4,
                         periodicity=(True, True, False), boundary=((0, 0), (0, 0), (1, 2)))
    kw = dict(orientation=abi.FLAT_ORIENTATION, ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0),
              gravity_source=True, C_smag=0.23, numerical_flux=abi.RUSANOV)
    cfgs = {"plain": atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, **kw),
            "learned": atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, **kw)}
    aux = atmos.init_aux(cfgs["plain"], grid)
    Q0 = atmos.rising_bubble(prm, grid, theta_ref=300.0, dtheta=2.0, radius=250.0,
                             centre=(L / 2, L / 2, 350.0), noise=1e-3, seed=1)
    c = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    # Courant number 0.3, the reference's default for explicit LSRK54
    # (src/Driver/solver_configs.jl:219-228); 0.8 diverges at N = 4
    dt = 0.3 * grid.min_node_distance() / c
    return grid, cfgs, aux, Q0, dt


def tgv_viscosity():
    """Molecular viscosity of the Re 1600 Taylor-Green run the network was trained on."""
    with open(TRAIN_META) as f:
        return float(json.load(f)["nu"])


def build_tgv(ne=96, lean=True):
    """The Taylor-Green workload; lean: the brick without per-node geometry arrays (the
    library builds its element records from the break points, cmdg_create_brick)."""
    from cmdg import abi, atmos
    from cmdg.grid import uniform_brick
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((ne, ne, ne), (0.0, 0.0, 0.0), (2 * math.pi,) * 3, N=4, lean=lean)
    cfgs = {"plain": atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.21),
            "learned": atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, C_smag=0.21,
                                         viscosity=tgv_viscosity())}
    aux = atmos.init_aux(cfgs["plain"], grid)
    Q0 = atmos.taylor_green(prm, grid, mach=0.1)
    c = math.sqrt(prm.cp_d / prm.cv_d * prm.R_d * 300.0)
    dt = 0.3 * grid.min_node_distance() / (1.1 * c)
    return grid, cfgs, aux, Q0, dt


def closure_weights(workload: str):
    """The network the chosen form runs: on the Taylor-Green vortex the one trained on
    the filtered 64^3 run; on the LES bubble the network whose output is C_s^2
    everywhere (Smagorinsky-Lilly's nu_0 through the fused MLP: the state difference to
    the plain run then measures the arithmetic of the chosen form alone; the MLP's cost
    does not depend on its weights)."""
    from cmdg import closure
    if workload == "tgv96":
        return closure.MLPParams.load(WEIGHTS), "trained on the box-filtered 64^3 Re 1600 Taylor-Green run"
    return closure.MLPParams.constant(0.23 ** 2), \
        "network with constant output C_s^2 (Smagorinsky-Lilly's nu_0 through the fused MLP)"


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, steps):
    """The oracle (reference algorithm restated in C++, OpenMP, FP64) running the plain
    method for `steps` LSRK54 steps.  LES: the full 25^3 workload.  Taylor-Green: a
    16^3-element sample of the same physics (same element size ratio, same state
    formula); GDOF/s is a per-DOF rate, reported as the rate of the 96^3 workload
    (extrapolated: the oracle's cost per DOF does not depend on the element count)."""
    import numpy as np
    from oracle import oracle as O
    if workload == "tgv96":
        grid, cfgs, aux, Q0, dt = build_tgv(16, lean=False)   # the oracle reads per-node geometry
        label = (f"oracle, plain FP64 Smagorinsky-Lilly, {steps} LSRK54 steps (5 RHS stages each) of a "
                 f"16^3-element sample of the Taylor-Green workload ({grid.Np * grid.nelem:,} nodes) after 1 "
                 "warm-up step; per-DOF rate extrapolated to the 96^3 workload")
    else:
        grid, cfgs, aux, Q0, dt = build_les(25)
        label = (f"oracle, plain FP64 Smagorinsky-Lilly, {steps} LSRK54 steps (5 RHS stages each) of the same "
                 "25^3-element LES workload after 1 warm-up step")
    orc = O.Oracle(cfgs["plain"], grid, aux)
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
            "cpu_model": cpu_model(), "extrapolated": workload == "tgv96",
            "sample": f"{label}, {el:.1f} s wall"}


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


def counter_passes(kernels, workload, ne, timeout_s=240):
    """Counters per launch of the kernels whose names start with the given prefixes,
    from the two rocprofv3 --pmc passes of PASSES over a short run of the chosen-form
    leg in child processes.  FETCH_SIZE / WRITE_SIZE (KiB) are scaled to bytes by what
    they read for a plain 1 GiB device copy run first in the same child.  Returns
    ({prefix: {counter: mean per dispatch}}, detail) or (None, reason)."""
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
                "--copy-calibration-mib", "1024", "--counters", "off", "--no-cpu-baseline"]
    env = dict(os.environ, TMPDIR="/tmp")
    out = {k: {} for k in kernels}
    detail = {}
    with tempfile.TemporaryDirectory(prefix="cmdg_pmc_", dir="/tmp") as tmp:
        for i, counters in enumerate(PASSES):
            d = os.path.join(tmp, f"pass{i}")
            cmd = [prof, "--pmc", *counters, "-d", d, "-o", f"pass{i}"] + cmd_tail
            print(f"[bench] counter pass {i + 1}/{len(PASSES)}: {' '.join(counters)}", file=sys.stderr, flush=True)
            p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                 start_new_session=True)
            try:
                _, err = p.communicate(timeout=timeout_s)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return None, f"rocprofv3 pass {i + 1} did not finish in {timeout_s} s"
            dbs = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)
            if p.returncode != 0 and not dbs:
                tail = err.decode(errors="replace").strip().splitlines()[-1:] if err else []
                return None, f"rocprofv3 pass {i + 1} exited {p.returncode}: {' '.join(tail)}"
            vals = {k: {} for k in kernels}
            calib = {}
            for db in dbs:
                con = sqlite3.connect(db)
                rows = con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection "
                                   "order by dispatch_id").fetchall()
                con.close()
                for did, name, cname, value in rows:
                    bare = name.replace("void ", "").replace("cmdg::", "")
                    hit = [k for k in kernels if bare.startswith(k)]
                    if hit:
                        vals[hit[0]].setdefault(cname, {}).setdefault(did, 0.0)
                        vals[hit[0]][cname][did] += float(value)
                    elif did <= 2 and cname in ("FETCH_SIZE", "WRITE_SIZE"):  # the 1 GiB fill, then the copy
                        calib[cname] = max(calib.get(cname, 0.0), float(value))
            for k in kernels:
                for cname, per in vals[k].items():
                    out[k][cname] = sum(per.values()) / len(per)
                    detail.setdefault("dispatches", {})[k] = len(per)
            for cname, v in calib.items():
                if v > 0:
                    detail[cname.lower() + "_scale"] = (1 << 20) / v  # KiB counted for 1 GiB -> bytes per KiB
    for k in kernels:
        for cname in ("FETCH_SIZE", "WRITE_SIZE"):
            if cname in out[k] and cname.lower() + "_scale" in detail:
                out[k][cname + "_bytes"] = out[k][cname] * detail[cname.lower() + "_scale"] * 1024
    detail["method"] = ("rocprofv3 --pmc, two passes (" + "; ".join(" ".join(p) for p in PASSES) + "), means per "
                        "dispatch over 3 steps of the chosen form; FETCH_SIZE / WRITE_SIZE scaled by the same counter "
                        "on a 1 GiB device copy")
    return out, detail


def flop_rates(cnt, avg_ms):
    """FLOP/s from the VALU instruction counters of one kernel (64 lanes a wave
    instruction; FMA = 2 flops, add / mul / transcendental = 1; a BF16 MFMA 'mop' = 512
    flops)."""
    def per(kind):
        a = cnt.get(f"SQ_INSTS_VALU_ADD_{kind}", 0.0) + cnt.get(f"SQ_INSTS_VALU_MUL_{kind}", 0.0) + \
            cnt.get(f"SQ_INSTS_VALU_TRANS_{kind}", 0.0) + 2 * cnt.get(f"SQ_INSTS_VALU_FMA_{kind}", 0.0)
        return 64 * a
    f32, f64 = per("F32"), per("F64")
    mfma = 512 * cnt.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0)
    s = avg_ms * 1e-3
    return {"fp32_tflops": round(f32 / s / 1e12, 3), "fp64_tflops": round(f64 / s / 1e12, 3),
            "bf16_mfma_tflops": round(mfma / s / 1e12, 3),
            "fp32_frac": round(f32 / s / 1e12 / PEAK_TFLOPS["fp32"], 4),
            "fp64_frac": round(f64 / s / 1e12 / PEAK_TFLOPS["fp64"], 4),
            "bf16_mfma_frac": round(mfma / s / 1e12 / PEAK_TFLOPS["bf16_mfma"], 5),
            "flops_fp32": f32, "flops_fp64": f64, "flops_bf16_mfma": mfma}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # at least 100 timed steps after 10 warm-up steps, the measurement the GDOF/s figure is
    # defined on
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=("tgv96", "les25"), default="tgv96")
    ap.add_argument("--elements", type=int, default=0, help="elements per direction (default 96 / 25)")
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle steps of the CPU baseline (0: 100 / 40)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=("all", "fp64", "mixed", "learned"), default="all",
                    help="run one leg alone")
    ap.add_argument("--counters", choices=("on", "off"), default="on",
                    help="measure HBM bytes and VALU / MFMA instructions per launch with two rocprofv3 counter "
                         "passes (child processes; use off when the bench itself runs under a profiler)")
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
    from cmdg.dg import DGModel, to_device

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    wl = args.workload
    ne = args.elements or (96 if wl == "tgv96" else 25)
    t_build = time.perf_counter()
    grid, cfgs, aux, Q0, dt = (build_tgv if wl == "tgv96" else build_les)(ne)
    nodes = grid.Np * grid.nelem
    dof = nodes * 5
    print(f"[bench] {wl}: {grid.nelem} elements, {nodes} nodes, dt {dt:.4g} s "
          f"(host build {time.perf_counter() - t_build:.1f} s)", file=sys.stderr, flush=True)
    gd = lambda el: dof / (el / args.steps / 5) / 1e9
    out = {}
    Q0d = to_device(Q0)
    del Q0

    def change(Q):
        return (Q - Q0d).cpu().numpy()

    def rel_err(a, b):
        return [float(np.linalg.norm(a[:, s] - b[:, s]) / np.linalg.norm(b[:, s])) for s in range(5)]

    def state_err(a, b):
        # the same difference against the size of the plain run's final state
        return [float(np.linalg.norm(a[:, s] - b[:, s]) / out["fp64_state_norm"][s]) for s in range(5)]

    # plain FP64 leg (and the reference state change for the error statement)
    if args.only in ("all", "fp64"):
        dg = DGModel(cfgs["plain"], grid, aux=aux, precision=abi.FP64)
        Q = Q0d.clone()
        el, _ = timed(dg, Q, dt, args.steps, args.warmup)
        out["fp64"], out["fp64_ms"] = gd(el), el / args.steps * 1e3
        dQ_fp64 = change(Q) if args.only == "all" else None
        out["fp64_state_norm"] = [float(torch.linalg.vector_norm(Q[:, s])) for s in range(5)]
        dg.close()
        del Q
        torch.cuda.empty_cache()
        print(f"[bench] plain FP64: {out['fp64_ms']:.3f} ms/step, {out['fp64']:.2f} GDOF/s per stage",
              file=sys.stderr, flush=True)
    # mixed precision with Smagorinsky-Lilly: the error of the precision change
    if args.only in ("all", "mixed"):
        dg = DGModel(cfgs["plain"], grid, aux=aux, precision=abi.MIXED)
        Q = Q0d.clone()
        el, _ = timed(dg, Q, dt, args.steps, args.warmup)
        out["mixed"] = gd(el)
        if args.only == "all":
            dq = change(Q)
            out["mixed_err"], out["mixed_state_err"] = rel_err(dq, dQ_fp64), state_err(dq, dQ_fp64)
            del dq
        dg.close()
        del Q
        torch.cuda.empty_cache()
        print(f"[bench] mixed + Smagorinsky: {out['mixed']:.2f} GDOF/s per stage", file=sys.stderr, flush=True)
    # chosen form: mixed precision + learned closure (and its state change vs plain FP64)
    mlp, mlp_note = closure_weights(wl)
    if args.only in ("all", "learned"):
        dg = DGModel(cfgs["learned"], grid, aux=aux, precision=abi.MIXED)
        dg.set_closure(mlp)
        Q = Q0d.clone()
        el, solver = timed(dg, Q, dt, args.steps, args.warmup)
        out["learned"], out["learned_ms"] = gd(el), el / args.steps * 1e3
        if args.only == "all":
            dq = change(Q)
            out["learned_err"], out["learned_state_err"] = rel_err(dq, dQ_fp64), state_err(dq, dQ_fp64)
            del dq
        kt = kernel_pass(dg, Q, solver, args.steps)
        dg.close()
        del Q
        torch.cuda.empty_cache()
        print(f"[bench] mixed + learned closure: {out['learned_ms']:.3f} ms/step, {out['learned']:.2f} GDOF/s "
              f"per stage; k_tendency {kt['tendency'][0] * 1e3:.1f} us, k_gradient {kt['gradient'][0] * 1e3:.1f} us",
              file=sys.stderr, flush=True)
    if args.only != "all":
        print(json.dumps(out), flush=True)
        return
    del Q0d
    torch.cuda.empty_cache()

    ceiling = hbm_ceiling()
    print(f"[bench] HBM ceiling on this GPU: {ceiling}", file=sys.stderr, flush=True)
    B = BYTES[wl]
    t_avg, g_avg = kt["tendency"][0], kt["gradient"][0]
    kern = {"k_tendency_mix": (t_avg, lsrk_bytes(B["tendency"], [args.steps])),
            "k_gradient_mix": (g_avg, B["gradient"])}
    b_counted = {"k_tendency_mix": lsrk_bytes(B["tendency"], COUNTER_CALLS), "k_gradient_mix": B["gradient"]}
    dom = "k_tendency_mix" if t_avg >= g_avg else "k_gradient_mix"
    achieved = {k: b * nodes / (ms * 1e-3) / 1e9 for k, (ms, b) in kern.items()}
    counters, cdetail = None, "not measured (--counters off)"
    if args.counters == "on":
        counters, cdetail = counter_passes(list(kern), wl, ne)
        print(f"[bench] counters: {counters} ({cdetail})", file=sys.stderr, flush=True)
    traffic = None
    per_kernel = {}
    for k, (ms, b) in kern.items():
        e = {"avg_launch_us": round(ms * 1e3, 2), "algorithmic_bytes_per_node": round(b, 2),
             "achieved_GBs": round(achieved[k], 1), "frac": round(achieved[k] / HBM_PEAK_GBS, 4)}
        c = (counters or {}).get(k, {})
        if "FETCH_SIZE_bytes" in c and "WRITE_SIZE_bytes" in c:
            tb = c["FETCH_SIZE_bytes"] + c["WRITE_SIZE_bytes"]
            e.update(traffic=int(tb), traffic_ratio_to_algorithmic=round(tb / (b_counted[k] * nodes), 3),
                     algorithmic_bytes_per_node_counted_runs=round(b_counted[k], 2),
                     traffic_rate_GBs=round(tb / (ms * 1e-3) / 1e9, 1))
            if k == dom:
                traffic = int(tb)
        if any(n.startswith("SQ_INSTS_VALU_") for n in c):
            e["compute"] = flop_rates(c, ms)
            if "SQ_ACTIVE_INST_VALU" in c and "SQ_BUSY_CYCLES" in c and c["SQ_BUSY_CYCLES"] > 0:
                # SQ_ACTIVE_INST_VALU counts 4-cycle units summed over all SIMDs; SQ_BUSY_CYCLES
                # sums the 32 shader-engine sequencers (8 XCDs x 4) over 1024 SIMDs
                e["compute"]["valu_issue_frac"] = round(4 * c["SQ_ACTIVE_INST_VALU"] /
                                                        (c["SQ_BUSY_CYCLES"] / 32 * 1024), 3)
            if "SQ_WAVES" in c and c["SQ_WAVES"] > 0 and "SQ_INSTS_VALU" in c:
                e["compute"]["valu_insts_per_wave"] = round(c["SQ_INSTS_VALU"] / c["SQ_WAVES"], 1)
        per_kernel[k] = e
    if os.path.exists(COUNTS):
        with open(COUNTS) as f:
            alg = json.load(f)
        key = "tgv_learned" if wl == "tgv96" else "les_learned"
        if key in alg.get("flops_per_node_per_stage", {}):
            fl = alg["flops_per_node_per_stage"][key]
            for k, part in (("k_gradient_mix", "gradient"), ("k_tendency_mix", "tendency")):
                if part in fl and "compute" in per_kernel.get(k, {}):
                    per_kernel[k]["compute"]["algorithmic_flops_per_node"] = fl[part]
                    per_kernel[k]["compute"]["algorithmic_tflops"] = round(
                        fl[part] * nodes / (kern[k][0] * 1e-3) / 1e12, 3)
    workload = (f"Taylor-Green vortex, Mach 0.1, {ne}^3 elements, N=4, periodic 2 pi box, Rusanov, "
                f"LSRK54CarpenterKennedy: the full chosen form at scale against plain FP64 Smagorinsky-Lilly"
                if wl == "tgv96" else
                f"dry rising thermal bubble LES, {ne}^3 elements, N=4, gravity + hydrostatic reference, walls in z, "
                "Rusanov, LSRK54CarpenterKennedy")
    stage_bytes = lsrk_bytes(B["stage"], [args.steps])
    stage_ach = stage_bytes * nodes / (out["learned_ms"] / 5 * 1e-3) / 1e9
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
                    "steps against the plain FP64 Smagorinsky-Lilly run from the same state",
            "chosen_form": [float(f"{e:.3e}") for e in out["learned_err"]],
            "mixed_smagorinsky": [float(f"{e:.3e}") for e in out["mixed_err"]],
            "relative_to_state": {
                "what": "the same differences against the norm of the plain run's final state",
                "chosen_form": [float(f"{e:.3e}") for e in out["learned_state_err"]],
                "mixed_smagorinsky": [float(f"{e:.3e}") for e in out["mixed_state_err"]],
            },
            "note": ("the chosen form's figures include the change of closure (trained network with the Re 1600 "
                     "viscosity floor against Smagorinsky-Lilly), not only the precision change"
                     if wl == "tgv96" else "the network reproduces Smagorinsky-Lilly: precision change only"),
        },
        "roofline": {
            "kernel": dom + (" (LSRK-fused, mixed)" if dom == "k_tendency_mix" else " (learned closure fused)"),
            "bound": "hbm",
            "achieved": round(achieved[dom], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved[dom] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "measured_ceiling": ceiling,
            "frac_of_measured_read_ceiling": round(achieved[dom] / ceiling["read"], 4),
            "kernels": per_kernel,
            "counter_detail": cdetail,
            "stage": {
                "what": "algorithmic bytes of a whole RK stage (both kernels) over the stage time, ms_per_step / 5",
                "bytes_per_node": round(stage_bytes, 2),
                "achieved": round(stage_ach, 1),
                "frac": round(stage_ach / HBM_PEAK_GBS, 4),
            },
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        steps = args.cpu_steps or (100 if wl == "tgv96" else 40)
        print(f"[bench] CPU baseline: oracle, {steps} steps", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(wl, steps)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

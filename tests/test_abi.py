"""The C ABI: header, ctypes mirror and the built library agree (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from cmdg import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cmdg.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(cmdg_\w+)\s*\(", text, re.M)))


def test_header_declares_the_exported_set():
    assert header_functions() == sorted(abi.EXPORTED)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(abi.LIB_PATH), "run __graft_entry__.build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (cmdg_\w+)$", out, re.M))
    missing = set(header_functions()) - exported
    assert not missing, missing


def test_library_loads_and_reports_version():
    lib = abi.load_library()
    assert lib.cmdg_version().decode().startswith("cmdg-mi355x")


def test_descriptor_layout_matches_header():
    text = open(HEADER).read()
    body = text[text.index("typedef struct cmdg_desc {"):text.index("} cmdg_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    decl = []
    for line in body.splitlines()[1:]:
        line = line.strip().rstrip(";").replace("const ", "")
        if not line:
            continue
        _, names = line.split(None, 1)
        decl += [n.strip().lstrip("*").strip() for n in names.split(",")]
    assert decl == [f[0] for f in abi.CmdgDesc._fields_]


def test_create_rejects_bad_descriptors_without_gpu_work():
    lib = abi.load_library()
    d = abi.CmdgDesc()
    d.N = 9  # not compiled
    ctx = C.c_void_p()
    rc = lib.cmdg_create(C.byref(d), None, C.byref(ctx))
    assert rc == abi.CMDG_EUNSUPPORTED
    assert b"compiled" in lib.cmdg_last_error(ctx)
    lib.cmdg_destroy(ctx)
    assert lib.cmdg_create(None, None, C.byref(ctx)) == abi.CMDG_EINVAL


class _RecordingSolver:
    """A solver with the ODESolvers accessors that records what solve asks of it."""

    def __init__(self, dt, batches):
        self.t, self.dt, self.steps, self.calls = 0.0, dt, 0, []
        if batches:
            self.dosteps = lambda Q, p, t, n: self.calls.extend(("step", t + k * self.dt, self.dt) for k in range(n))

    def gettime(self):
        return self.t

    def getdt(self):
        return self.dt

    def updatedt(self, dt):
        self.dt = dt

    def updatetime(self, t):
        self.t = t
        return t

    def dostep(self, Q, p, t):
        self.calls.append(("step", t, self.dt))


@pytest.mark.parametrize("kw", [dict(numberofsteps=7), dict(timeend=0.95), dict(timeend=0.95, adjustfinalstep=False),
                                dict(timeend=0.95, numberofsteps=4), dict(numberofsteps=1)])
def test_solve_batches_steps_like_the_step_loop(kw):
    """solve runs the steps between callbacks and the adjusted final step as one
    batch; the batch visits the same start times, step sizes and final time as the
    one-step-at-a-time loop of solve! (ODESolvers.jl:110-158)."""
    from cmdg.dg import solve
    ref, got = _RecordingSolver(0.1, False), _RecordingSolver(0.1, True)
    t_ref = solve(None, ref, **kw)
    t_got = solve(None, got, **kw)
    assert t_ref == t_got and ref.steps == got.steps and ref.dt == got.dt
    assert [c[0] for c in got.calls] == [c[0] for c in ref.calls]
    for a, b in zip(got.calls, ref.calls):
        assert a[1] == pytest.approx(b[1], abs=1e-15) and a[2] == b[2]


def test_static_loss_gradient_rejects_oversized_batches_without_gpu_work():
    """The size check of the allocation-free loss gradient comes before any device call."""
    lib = abi.load_library()
    p = C.c_void_p(256)
    n = abi.MLP_STATIC_MAX_BATCH + 1
    assert lib.cmdg_mlp_loss_grad_static(p, p, p, p, n, None, n, 1.0, p, None) == abi.CMDG_ERANGE
    assert lib.cmdg_mlp_loss_grad_static(p, p, p, p, 10, None, 0, 1.0, p, None) == abi.CMDG_EINVAL


def test_header_constants_match_the_ctypes_mirror():
    text = open(HEADER).read()
    consts = dict(re.findall(r"^#define (CMDG_\w+) (\d+)", text, re.M))
    assert int(consts["CMDG_ERANGE"]) == abi.CMDG_ERANGE
    assert int(consts["CMDG_MLP_STATIC_MAX_BATCH"]) == abi.MLP_STATIC_MAX_BATCH
    assert int(consts["CMDG_MLP_NPARAM"]) == abi.MLP_NPARAM

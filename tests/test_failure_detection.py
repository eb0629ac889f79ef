"""Crash detection (show_not_finite_fields / the crash check of checked_wait,
src/Arrays/MPIStateArrays.jl:884-927): the host logic on the CPU with a stand-in
for the library call, and the device map-reduce on the GPU."""
import numpy as np
import pytest

from cmdg.dg import NotFiniteCheck, STATE_FIELDS, fields_of_mask


def test_fields_of_mask():
    assert fields_of_mask(0) == []
    assert fields_of_mask(1) == ["ρ"]
    assert fields_of_mask(0b10110) == ["ρu[1]", "ρu[2]", "energy.ρe"]
    assert fields_of_mask(31) == list(STATE_FIELDS)


class _FakeDG:
    def __init__(self, bad_from_step):
        self.bad_from_step, self.calls = bad_from_step, []

    def not_finite_fields(self, Q):
        self.calls.append(Q["step"])
        return ["energy.ρe"] if Q["step"] >= self.bad_from_step else []


class _FakeSolver:
    steps = 0


def test_not_finite_check_stops_on_the_first_checked_step():
    dg, solver = _FakeDG(bad_from_step=5), _FakeSolver()
    cb = NotFiniteCheck(dg, every=2)
    for step in range(1, 5):
        solver.steps = step
        assert cb(solver, {"step": step}, None, 0.1 * step) is None
    assert dg.calls == [2, 4]            # only every second step is checked
    solver.steps = 5
    assert cb(solver, {"step": 5}, None, 0.5) is None   # not a checked step
    solver.steps = 6
    with pytest.raises(FloatingPointError, match=r"step 6.*energy\.ρe"):
        cb(solver, {"step": 6}, None, 0.6)
    with pytest.raises(ValueError):
        NotFiniteCheck(dg, every=0)


@pytest.mark.gpu
def test_device_not_finite_fields():
    import torch
    from cmdg import abi, atmos
    from cmdg.dg import DGModel, LSRK54CarpenterKennedy, solve, to_device
    from cmdg.grid import uniform_brick
    print("[failure] not_finite_fields", flush=True)
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((3, 2, 2), (0, 0, 0), (3.0, 2.0, 2.0), N=4)
    for precision in (abi.FP64, abi.MIXED):
        cfg = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.2)
        dg = DGModel(cfg, grid, precision=precision)
        Q0 = atmos.taylor_green(prm, grid, mach=0.1, L=3.0 / (2 * np.pi))
        Q = to_device(Q0)
        assert dg.not_finite_fields(Q) == []
        for planted, want in (({(7, 3, 5): np.nan}, ["ρu[3]"]),
                              ({(0, 0, 0): np.inf, (124, 4, 11): -np.inf}, ["ρ", "energy.ρe"])):
            Qb = Q0.copy(order="F")
            for (n, s, e), v in planted.items():
                Qb[n, s, e] = v
            assert dg.not_finite_fields(to_device(Qb)) == want
        # as a solve callback: a NaN planted mid-run stops the integration with its name
        solver = LSRK54CarpenterKennedy(dg, Q, dt=1e-4, t0=0.0)
        solve(Q, solver, numberofsteps=2, callbacks=[NotFiniteCheck(dg)])
        Q.view(-1)[3] = float("nan")          # element 0, rho, node 3
        with pytest.raises(FloatingPointError, match="ρ"):
            solve(Q, solver, numberofsteps=1, callbacks=[NotFiniteCheck(dg)])
        dg.close()

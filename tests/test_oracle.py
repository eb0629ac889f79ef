"""Pins the CPU oracle to the reference's own known answers and to closed forms.

Known-answer values come from the reference test suite (file:line cited on each
test); nothing here runs the reference.
"""
import json
import math
import os

import numpy as np
import pytest

from cmdg import abi, atmos
from cmdg.grid import BrickGrid, uniform_brick
from oracle import oracle as O


def vortex_run(level, numflux=abi.RUSANOV, use_float=False):
    """test_run of test/Numerics/DGMethods/Euler/isentropicvortex.jl:306-442 (3-D)."""
    prm = atmos.earth_parameters()
    setup = atmos.IsentropicVortex()
    L = setup.domain_halflength
    ne = 2 ** (level - 1) * 5
    br = [np.linspace(-L, L, ne + 1), np.linspace(-L, L, ne + 1), np.linspace(-L, L, 2)]
    grid = BrickGrid(br, N=4, periodicity=(True, True, True))
    cfg = atmos.AtmosConfig(prm, turbulence=abi.CONSTANT_DYNAMIC_VISCOSITY, viscosity=0.0,
                            numerical_flux=numflux)
    ft = np.float32 if use_float else np.float64
    aux = atmos.init_aux(cfg, grid)
    orc = O.Oracle(cfg, grid, aux, use_float=use_float)
    Q = np.asfortranarray(setup(prm, grid, 0.0).astype(ft))
    timeend = ft(ft(ft(2) * ft(L)) / ft(10)) / ft(setup.translation_speed)
    c = np.sqrt(ft(prm.cp_d / prm.cv_d) * ft(prm.R_d) * ft(setup.T_inf)) if not use_float else \
        ft(np.sqrt(ft(ft(prm.cp_d) / ft(prm.cv_d)) * ft(prm.R_d) * ft(setup.T_inf)))
    elementsize = ft(min(np.diff(b).min() for b in br))
    dt = ft(ft(elementsize / c) / ft(16))
    nsteps = math.ceil(timeend / dt)
    dt = ft(timeend / ft(nsteps))
    O.solve(orc, Q, dt, timeend)
    Qe = setup(prm, grid, float(timeend))
    return O.euclidean_distance(Q, Qe, grid.mass)


KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def vortex_kat(table, level):
    """expected_error[FT, 3, flux, level] of isentropicvortex.jl (the line is in the fixture)
    and the reference's tolerance, sqrt(eps(FT))."""
    value, _line = KATS["isentropic_vortex_3d"]["expected_error"][table][str(level)]
    ft = np.float32 if table.startswith("Float32") else np.float64
    return value, math.sqrt(np.finfo(ft).eps)


def test_vortex_rusanov_fp64_level1():
    # isentropicvortex.jl:105  expected_error[Float64, 3, Rusanov, 1]
    value, rtol = vortex_kat("Float64 Rusanov", 1)
    assert vortex_run(1) == pytest.approx(value, rel=rtol)


def test_vortex_central_fp64_level1():
    # isentropicvortex.jl:110  expected_error[Float64, 3, Central, 1]
    value, rtol = vortex_kat("Float64 Central", 1)
    assert vortex_run(1, numflux=abi.CENTRAL) == pytest.approx(value, rel=rtol)


def test_vortex_rusanov_fp32_level1():
    # isentropicvortex.jl:195  expected_error[Float32, 3, Rusanov, 1], rtol sqrt(eps(Float32))
    value, rtol = vortex_kat("Float32 Rusanov", 1)
    assert vortex_run(1, use_float=True) == pytest.approx(value, rel=rtol)


def test_vortex_central_fp32_level1():
    # isentropicvortex.jl:200  expected_error[Float32, 3, Central, 1], rtol sqrt(eps(Float32))
    value, rtol = vortex_kat("Float32 Central", 1)
    assert vortex_run(1, numflux=abi.CENTRAL, use_float=True) == pytest.approx(value, rel=rtol)


@pytest.mark.slow
def test_vortex_rusanov_fp64_level2():
    # isentropicvortex.jl:106  expected_error[Float64, 3, Rusanov, 2]
    value, rtol = vortex_kat("Float64 Rusanov", 2)
    assert vortex_run(2) == pytest.approx(value, rel=rtol)


def _golden():
    import importlib.util
    path = os.path.join(os.path.dirname(__file__), "golden", "make_oracle_vectors.py")
    spec = importlib.util.spec_from_file_location("make_oracle_vectors", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_oracle_reproduces_golden_vectors():
    """The committed golden vectors (tests/golden/oracle_les_bubble_2x2x2.npz: the
    oracle's RHS and one LSRK54 step on a seeded 2 x 2 x 2 LES bubble) are what the
    oracle computes now: a change to the oracle that moves them fails here."""
    g = _golden()
    ref = np.load(g.FILE)
    cfg, grid, aux = g.case()
    orc = O.Oracle(cfg, grid, aux)
    Q0 = np.asfortranarray(ref["Q0"])
    np.testing.assert_allclose(orc.tendency(Q0), ref["tendency"], rtol=0, atol=1e-14 * np.abs(ref["tendency"]).max())
    Q1 = np.asfortranarray(Q0.copy())
    dQ = np.zeros_like(Q1, order="F")
    orc.lsrk54_step(Q1, dQ, 0.0, float(ref["dt"]))
    np.testing.assert_allclose(Q1, ref["Q1"], rtol=1e-15, atol=0)

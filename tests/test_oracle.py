"""Pins the CPU oracle to the reference's own known answers and to closed forms.

Known-answer values come from the reference test suite (file:line cited on each
test); nothing here runs the reference.
"""
import math

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


def test_vortex_rusanov_fp64_level1():
    # isentropicvortex.jl:105  expected_error[Float64, 3, Rusanov, 1]
    err = vortex_run(1)
    assert err == pytest.approx(3.7918869862613858e+00, rel=math.sqrt(np.finfo(np.float64).eps))


def test_vortex_central_fp64_level1():
    # isentropicvortex.jl:110  expected_error[Float64, 3, Central, 1]
    err = vortex_run(1, numflux=abi.CENTRAL)
    assert err == pytest.approx(6.5903683487905749e+00, rel=math.sqrt(np.finfo(np.float64).eps))


def test_vortex_rusanov_fp32_level1():
    # isentropicvortex.jl:195  expected_error[Float32, 3, Rusanov, 1], rtol sqrt(eps(Float32))
    err = vortex_run(1, use_float=True)
    assert err == pytest.approx(3.7918186187744141e+00, rel=math.sqrt(np.finfo(np.float32).eps))


@pytest.mark.slow
def test_vortex_rusanov_fp64_level2():
    # isentropicvortex.jl:106  expected_error[Float64, 3, Rusanov, 2]
    err = vortex_run(2)
    assert err == pytest.approx(6.5816485664822677e-01, rel=math.sqrt(np.finfo(np.float64).eps))

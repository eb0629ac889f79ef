"""The reference's 3-D manufactured-solution known answer reproduced by the oracle
(test/Numerics/DGMethods/compressible_Navier_Stokes/mms_bc_atmos.jl:92-214, level 1):
compressible Navier-Stokes with ConstantDynamicViscosity(1/100, WithDivergence()),
NoOrientation, NoReferenceState, Rusanov / central second-order / central gradient
fluxes, InitStateBC on every face of a 4 x 4 x 4-element unit cube warped by the
test's meshwarp, the manufactured source, T_0 = 0 and total_specific_enthalpy = 0,
LSRK54CarpenterKennedy with dt = 1/800 to t = 1.  The known answer is the
mass-weighted L2 distance to the exact solution, 3.3983777728925593e-02 (rtol
sqrt(eps)).  It pins, together: the constant-viscosity stress with divergence, the
curvilinear metric terms and the boundary-state path of all three numerical fluxes."""
import dataclasses
import json
import math
import os

import numpy as np
import pytest

from cmdg import abi, atmos
from cmdg.grid import uniform_brick
from oracle import oracle as O

import mms_solution as MMS
from test_grid_curvilinear import mms_warp

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def mms_setup(ne=4):
    prm = dataclasses.replace(atmos.earth_parameters(), T_0=0.0)   # mms_bc_atmos.jl:33-34
    grid = uniform_brick((ne, ne, ne), (0, 0, 0), (1.0, 1.0, 1.0), N=4, periodicity=(False,) * 3,
                         boundary=((1, 1), (1, 1), (1, 1)), meshwarp=mms_warp)
    cfg = atmos.AtmosConfig(prm, turbulence=abi.CONSTANT_DYNAMIC_VISCOSITY, viscosity=0.01,
                            with_divergence=True, numerical_flux=abi.RUSANOV)
    return prm, grid, cfg


def run_mms(ne=4, lib_path=O.LIB_PATH):
    """solve!(Q, LSRK54CarpenterKennedy; timeend = 1) with the stage times of dostep!
    (LowStorageRungeKuttaMethod.jl:102-144) and the final-step rule of
    general_dostep! (ODESolvers.jl:49-74); returns euclidean_distance(Q, Qe)."""
    prm, grid, cfg = mms_setup(ne)
    aux = atmos.init_aux(cfg, grid)
    orc = O.Oracle(cfg, grid, aux, lib_path=lib_path)
    X = np.stack(grid.coords)
    Q = MMS.exact_state(0.0, X)
    dQ = np.zeros_like(Q, order="F")
    RKA, RKB, RKC = O.lsrk54_coefficients()
    timeend = 1.0
    dt = 5e-3 / ne
    nsteps = math.ceil(timeend / dt)
    dt = timeend / nsteps
    time = 0.0
    while time < timeend:
        step_dt, final = dt, False
        if time + dt > timeend:
            step_dt, final = timeend - time, True
        for s in range(5):
            ts = time + RKC[s] * step_dt
            orc.set_prescribed(MMS.exact_state(ts, X), MMS.source(ts, X), htot_zero=True)
            orc.rhs(dQ, Q, 1.0, 1.0)
            Q += RKB[s] * step_dt * dQ
            dQ *= RKA[(s + 1) % 5]
        time = timeend if final else time + dt
    orc.close()
    Qe = MMS.exact_state(timeend, X)
    return O.euclidean_distance(Q, Qe, grid.mass)


@pytest.mark.slow
def test_mms_3d_known_answer():
    kat = KATS["mms_bc_atmos_3d"]
    err = run_mms(4)
    assert err == pytest.approx(kat["expected_error"], rel=kat["rtol"]), err

"""HIP general-mesh kernels against the oracle on curvilinear and non-orthogonal
meshes, where every off-diagonal metric term xi_a x_d is non-zero (FP64 bar: 1e-12
relative L2 per state component).  The oracle's use of those terms is pinned by the
reference's manufactured-solution known answer on the warped mesh of
mms_bc_atmos.jl and by closed forms on sheared elements (test_oracle_pins.py)."""
import math

import numpy as np
import pytest

from cmdg import abi, atmos
from cmdg.grid import uniform_brick
from oracle import oracle as O

import mms_solution as MMS
from test_grid_curvilinear import mms_warp
from test_oracle_pins import sheared_brick, wavy_brick

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _case(name):
    rng = np.random.default_rng(sum(map(ord, name)))
    prm = atmos.ideal_gas_parameters()
    if name == "sheared_viscous":
        grid = sheared_brick()
        cfg = atmos.AtmosConfig(prm, turbulence=abi.CONSTANT_DYNAMIC_VISCOSITY, viscosity=2e-2,
                                with_divergence=True)
    elif name == "wavy_smagorinsky":
        grid = wavy_brick()
        cfg = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.21)
    elif name == "wavy_les_walls":
        # periodic in x and y, curved free-slip insulating walls at the bottom and top,
        # gravity and the hydrostatic reference state
        L = (300.0, 300.0, 300.0)
        tau = 2 * math.pi
        def warp(x, y, z):
            sx, sy = np.sin(tau * x / L[0]), np.sin(tau * y / L[1])
            return x + 6.0 * sy * np.sin(math.pi * z / L[2]), y + 6.0 * sx * np.sin(math.pi * z / L[2]), \
                z + 8.0 * sx * sy
        grid = uniform_brick((3, 3, 3), (0, 0, 0), L, N=4, periodicity=(True, True, False),
                             boundary=((0, 0), (0, 0), (1, 2)), meshwarp=warp)
        cfg = atmos.AtmosConfig(prm, orientation=abi.FLAT_ORIENTATION, ref_profile=atmos.DryAdiabaticProfile(300., 0.),
                                gravity_source=True, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.23)
    elif name == "mms_walls_viscous":
        # the warped unit cube of the manufactured-solution test, free-slip walls on
        # every face, the viscosity of that test; the state is its exact solution at t = 0.3
        prm = atmos.earth_parameters()
        grid = uniform_brick((3, 3, 3), (0, 0, 0), (1.0, 1.0, 1.0), N=4, periodicity=(False,) * 3,
                             boundary=((1, 1), (1, 1), (1, 1)), meshwarp=mms_warp)
        cfg = atmos.AtmosConfig(prm, turbulence=abi.CONSTANT_DYNAMIC_VISCOSITY, viscosity=0.01,
                                with_divergence=True)
    else:
        raise KeyError(name)
    aux = atmos.init_aux(cfg, grid)
    X = np.stack(grid.coords)
    if name == "mms_walls_viscous":
        Q0 = MMS.exact_state(0.3, X)
        Q0[:, 4] += 2.5e5  # a realistic internal energy (T_0 = 273.16 here)
    elif name == "wavy_les_walls":
        Q0 = atmos.rising_bubble(prm, grid, radius=100.0, centre=(150.0, 150.0, 100.0), seed=4)
        Q0[:, 1:4, :] += rng.uniform(-1, 1, Q0[:, 1:4, :].shape)
    else:
        Q0 = atmos.taylor_green(prm, grid, mach=0.2)
        Q0[:, 1:4, :] *= 1 + 0.05 * rng.uniform(-1, 1, Q0[:, 1:4, :].shape)
    return cfg, grid, aux, np.asfortranarray(Q0)


CASES = ["sheared_viscous", "wavy_smagorinsky", "wavy_les_walls", "mms_walls_viscous"]


@pytest.mark.parametrize("name", CASES)
def test_general_kernels_rhs_on_curvilinear_mesh(name):
    import torch
    from cmdg.dg import DGModel, to_device, to_host
    print(f"[curvilinear] rhs {name}", flush=True)
    cfg, grid, aux, Q0 = _case(name)
    # off-diagonal metric terms are present and not negligible
    offd = np.abs(grid.vgeo[:, [1, 2, 3, 5, 6, 7]]).max()
    assert offd > 1e-3 * np.abs(grid.vgeo[:, [0, 4, 8]]).max()
    orc = O.Oracle(cfg, grid, aux)
    ref = orc.tendency(Q0)
    dg = DGModel(cfg, grid, aux=aux)
    assert dg.geometry_kind == "general"
    Q = to_device(Q0)
    dQ = torch.full_like(Q, float("nan"))
    dg(dQ, Q, None, 0.0)
    got = to_host(dQ)
    for s in range(5):
        r = rel_l2(got[:, s], ref[:, s])
        print(f"  state {s}: rel L2 {r:.3e}", flush=True)
        assert r < RTOL, (name, s, r)
    gf_dev = to_host(dg.state_gradient_flux)
    gf_ref = orc.gradient_flux()
    for s in range(cfg.ngradflux):
        assert rel_l2(gf_dev[:, s], gf_ref[:, s]) < RTOL, ("gradflux", s)
    lay = cfg.layout
    aux_dev = to_host(dg.state_auxiliary)
    for key in ("theta_v", "air_T"):
        assert rel_l2(aux_dev[:, lay[key]], orc.aux[:, lay[key]]) < 1e-14, key
    dg.close()
    orc.close()


@pytest.mark.parametrize("name", ["wavy_smagorinsky", "mms_walls_viscous"])
def test_general_kernels_lsrk54_on_curvilinear_mesh(name):
    from cmdg.dg import DGModel, LSRK54CarpenterKennedy, solve, to_device, to_host
    print(f"[curvilinear] lsrk54 {name}", flush=True)
    cfg, grid, aux, Q0 = _case(name)
    # acoustic Courant number 0.2 from the fastest node of the initial state
    prm = cfg.params
    rho = Q0[:, 0]
    ke = 0.5 * (Q0[:, 1] ** 2 + Q0[:, 2] ** 2 + Q0[:, 3] ** 2) / rho
    T = (Q0[:, 4] - ke) / rho / prm.cv_d + prm.T_0
    speed = np.sqrt(prm.cp_d / prm.cv_d * prm.R_d * T) + np.sqrt(2 * ke / rho)
    dt = 0.2 * grid.min_node_distance() / float(speed.max())
    orc = O.Oracle(cfg, grid, aux)
    Qr = np.asfortranarray(Q0.copy())
    dQr = np.zeros_like(Qr, order="F")
    for k in range(3):
        orc.lsrk54_step(Qr, dQr, k * dt, dt)
    assert np.isfinite(Qr).all(), "the oracle run itself is not finite: the time step is too large"
    dg = DGModel(cfg, grid, aux=aux)
    Q = to_device(Q0)
    solver = LSRK54CarpenterKennedy(dg, Q, dt=dt, t0=0.0)
    solve(Q, solver, numberofsteps=3)
    got = to_host(Q)
    for s in range(5):
        r = rel_l2(got[:, s] - Q0[:, s], Qr[:, s] - Q0[:, s])
        print(f"  state {s}: change rel L2 {r:.3e}", flush=True)
        assert r < 1e-11, (name, s, r)
    dg.close()
    orc.close()

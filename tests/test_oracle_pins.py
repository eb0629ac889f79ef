"""Closed-form pins of oracle functions that the reference's known answers reach only
in combination: the constant-viscosity stress (TurbulenceClosures.jl:357-404), the
buoyancy frequency N^2 of Smagorinsky-Lilly under gravity (:455-470), and the use
of off-diagonal metric terms (vgeo fields xi_a x_d, Grids.jl:76-92) in the volume
tendency and the volume gradient.  Each pin is a function of the oracle library;
it must hold for the oracle and fail for a copy of the oracle with the mistake
named beside it planted in its source (tests/oracle_mutants.py)."""
import math

import numpy as np
import pytest

from cmdg import abi, atmos
from cmdg.grid import uniform_brick
from oracle import oracle as O

from oracle_mutants import build as build_mutant


# ------------------------------------------------------------- constant viscosity
def _const_visc_flux(lib, turb, with_div, visc, q, S, gh):
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((1, 1, 1), (0, 0, 0), (1., 1., 1.), N=2)
    cfg = atmos.AtmosConfig(prm, turbulence=turb, viscosity=visc, with_divergence=with_div)
    orc = O.Oracle(cfg, grid, atmos.init_aux(cfg, grid), lib_path=lib)
    S11, S21, S31, S22, S32, S33 = S
    gf = np.array([*gh, S11, S21, S31, S22, S32, S33])
    F = orc.flux_second_order(np.asarray(q, float), gf)
    orc.close()
    return F, prm.inv_Pr_turb


def pin_constant_viscosity_stress(lib=O.LIB_PATH):
    """Textbook special cases of tau = -2 nu S (+ 2 nu / 3 tr(S) I with divergence,
    Stokes' hypothesis): F_momentum = rho tau, F_energy = tau rho u - rho (nu/Pr_t) grad h."""
    rho, nu, a = 1.3, 0.7, 0.4
    u = np.array([2.0, -1.0, 0.5])
    q = [rho, *(rho * u), 2.5e5]
    gh = np.array([0.2, -0.4, 0.9])
    for turb in (abi.CONSTANT_KINEMATIC_VISCOSITY, abi.CONSTANT_DYNAMIC_VISCOSITY):
        visc = nu if turb == abi.CONSTANT_KINEMATIC_VISCOSITY else nu * rho   # same nu
        # uniaxial strain S11 = a: with divergence tau = nu a diag(-4/3, 2/3, 2/3),
        # without it -2 nu a e1 e1
        for with_div, diag in ((True, (-4 / 3, 2 / 3, 2 / 3)), (False, (-2.0, 0.0, 0.0))):
            F, inv_Pr = _const_visc_flux(lib, turb, with_div, visc, q, (a, 0, 0, 0, 0, 0), gh)
            for d in range(3):
                for k in range(3):
                    want = rho * nu * a * diag[d] if d == k else 0.0
                    assert F[d, 1 + k] == pytest.approx(want, rel=1e-13, abs=1e-15), (turb, with_div, d, k)
                assert F[d, 0] == 0
                want_e = nu * a * diag[d] * rho * u[d] - rho * nu * inv_Pr * gh[d]
                assert F[d, 4] == pytest.approx(want_e, rel=1e-13), (turb, with_div, d)
            # traceless with divergence (no bulk viscosity), -2 nu a without
            tr = F[0, 1] + F[1, 2] + F[2, 3]
            assert tr == pytest.approx(0.0 if with_div else -2 * rho * nu * a, abs=1e-13)
        # simple shear S12 = a/2: rho tau12 = -rho nu a in either form; energy row
        # through y: -nu a rho u1 - rho (nu / Pr_t) dh/dy
        for with_div in (True, False):
            F, inv_Pr = _const_visc_flux(lib, turb, with_div, visc, q, (0, a / 2, 0, 0, 0, 0), gh)
            assert F[0, 2] == pytest.approx(-rho * nu * a, rel=1e-13)
            assert F[1, 1] == pytest.approx(-rho * nu * a, rel=1e-13)
            assert F[1, 4] == pytest.approx(-nu * a * rho * u[0] - rho * nu * inv_Pr * gh[1], rel=1e-13)
        # pure dilatation S = a I: no stress with divergence, -2 nu a I without
        for with_div in (True, False):
            F, _ = _const_visc_flux(lib, turb, with_div, visc, q, (a, 0, 0, a, 0, a), 0 * gh)
            want = 0.0 if with_div else -2 * rho * nu * a
            for d in range(3):
                assert F[d, 1 + d] == pytest.approx(want, abs=1e-13)
    # dynamic viscosity: nu = mu / rho, so at fixed mu the momentum flux rho tau = -mu a
    # (shear) does not depend on rho, and the enthalpy flux is -(mu / Pr_t) grad h
    mu = 0.35
    for rho2 in (0.5, 2.0):
        q2 = [rho2, *(rho2 * u), 2.5e5]
        F, inv_Pr = _const_visc_flux(lib, abi.CONSTANT_DYNAMIC_VISCOSITY, True, mu, q2, (0, a / 2, 0, 0, 0, 0), gh)
        assert F[0, 2] == pytest.approx(-mu * a, rel=1e-13)
        assert F[2, 4] == pytest.approx(-mu * inv_Pr * gh[2], rel=1e-13)


# ------------------------------------------------------------- N^2 under gravity
def pin_buoyancy_frequency(lib=O.LIB_PATH):
    """A rest atmosphere whose virtual potential temperature is linear in height,
    theta_v = theta_0 + G z: the DG gradient of a linear field is exact, so N^2 =
    grad theta_v . grad Phi / theta_v = g G / theta_v at every node."""
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((2, 2, 3), (0, 0, 0), (300., 300., 600.), N=4, periodicity=(True, True, False),
                         boundary=((0, 0), (0, 0), (1, 2)))
    cfg = atmos.AtmosConfig(prm, orientation=abi.FLAT_ORIENTATION, gravity_source=True,
                            turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.23)
    aux = atmos.init_aux(cfg, grid)
    z = grid.coords[2]
    theta0, G = 300.0, 3e-3
    theta = theta0 + G * z
    p = 1e5 - 11.0 * z                         # any positive pressure field
    T = theta * (p / prm.MSLP) ** prm.kappa_d  # theta_v = T (MSLP / p)^kappa
    rho = p / (prm.R_d * T)
    Phi = aux[:, cfg.layout["Phi"]]
    Q = np.zeros((grid.Np, 5, grid.nelem), order="F")
    Q[:, 0] = rho
    Q[:, 4] = rho * (prm.cv_d * (T - prm.T_0) + Phi)
    orc = O.Oracle(cfg, grid, aux, lib_path=lib)
    orc.tendency(Q)
    N2 = orc.gradient_flux()[:, 9]
    orc.close()
    want = prm.grav * G / theta
    assert np.abs(N2 - want).max() < 1e-9 * np.abs(want).max(), np.abs(N2 - want).max()


# ------------------------------------------------------------- off-diagonal metrics
A_SHEAR = np.array([[1.0, 0.3, -0.2], [0.1, 0.9, 0.25], [-0.15, 0.05, 1.2]])


def sheared_brick(**kw):
    """A 3 x 3 x 3 periodic brick mapped by the non-symmetric matrix A_SHEAR: affine
    elements whose metric terms have non-zero off-diagonal entries."""
    A = A_SHEAR
    warp = lambda x, y, z: tuple(A[i, 0] * x + A[i, 1] * y + A[i, 2] * z for i in range(3))
    return uniform_brick((3, 3, 3), (0, 0, 0), (3., 3., 1.5), N=4, meshwarp=warp, **kw)


def wavy_brick(n=(3, 3, 2), L=(3., 3., 2.), eps=0.08):
    """A periodic brick with a smooth periodic warp: curved elements."""
    tau = 2 * math.pi
    def warp(x, y, z):
        sx, sy, sz = np.sin(tau * x / L[0]), np.sin(tau * y / L[1]), np.sin(tau * z / L[2])
        return x + eps * sy * sz, y + eps * sx * sz, z + eps * sx * sy
    return uniform_brick(n, (0, 0, 0), L, N=4, meshwarp=warp)


def pin_free_stream_on_curved_mesh(lib=O.LIB_PATH):
    """A uniform flow has zero tendency on a periodic curved mesh: the curl-invariant
    metric terms satisfy the discrete identities sum_a D_a (J xi_a x_d) = 0."""
    prm = atmos.ideal_gas_parameters()
    grid = wavy_brick()
    cfg = atmos.AtmosConfig(prm, turbulence=abi.CONSTANT_DYNAMIC_VISCOSITY, viscosity=1e-2, with_divergence=True)
    orc = O.Oracle(cfg, grid, atmos.init_aux(cfg, grid), lib_path=lib)
    Q = np.zeros((grid.Np, 5, grid.nelem), order="F")
    u = np.array([3.0, -2.0, 1.0])
    Q[:, 0] = 1.2
    Q[:, 1:4] = 1.2 * u[None, :, None]
    Q[:, 4] = 1.2 * atmos.total_energy(prm, 0.5 * (u ** 2).sum(), 0.0, 290.0)
    T = orc.tendency(Q)
    orc.close()
    # roundoff of the flux divergence: 1e-11 of each component's flux magnitude
    # (elements of size ~1, so D / h ~ 10)
    p = prm.R_d * 1.2 * 290.0
    speed = float(np.sqrt((u ** 2).sum()))
    flux = [1.2 * speed, p + 1.2 * speed ** 2, p + 1.2 * speed ** 2, p + 1.2 * speed ** 2,
            speed * (float(Q[0, 4, 0]) + p)]
    for s in range(5):
        assert np.abs(T[:, s]).max() < 1e-11 * flux[s], (s, np.abs(T[:, s]).max(), flux[s])


def pin_linear_fields_on_sheared_mesh(lib=O.LIB_PATH):
    """On affine non-orthogonal elements the DG derivative of a linear field is
    exact: the strain rate of u = B x is sym(B), and the mass tendency of a linear
    momentum rho_0 B x is -rho_0 tr(B), in the centre element (the linear field jumps
    at the periodic seams, which it does not touch)."""
    prm = atmos.ideal_gas_parameters()
    grid = sheared_brick()
    cfg = atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.2)
    orc = O.Oracle(cfg, grid, atmos.init_aux(cfg, grid), lib_path=lib)
    B = np.array([[0.2, -0.5, 0.1], [0.3, 0.05, -0.25], [0.15, 0.4, -0.1]])
    X = np.stack(grid.coords)   # (3, Np, E)
    u = np.einsum("ij,jne->ine", B, X)
    rho0 = 1.1
    Q = np.zeros((grid.Np, 5, grid.nelem), order="F")
    Q[:, 0] = rho0
    Q[:, 1:4] = np.moveaxis(rho0 * u, 0, 1)
    Q[:, 4] = rho0 * atmos.total_energy(prm, 0.5 * (u ** 2).sum(axis=0), 0.0, 300.0)
    T = orc.tendency(Q)
    gf = orc.gradient_flux()
    orc.close()
    inner = (grid.ex == 1) & (grid.ey == 1) & (grid.ez == 1)
    Ssym = (B + B.T) / 2
    comp = {3: (0, 0), 4: (1, 0), 5: (2, 0), 6: (1, 1), 7: (2, 1), 8: (2, 2)}
    for s, (i, j) in comp.items():
        got = gf[:, s][:, inner]
        assert np.abs(got - Ssym[i, j]).max() < 1e-11, (s, np.abs(got - Ssym[i, j]).max())
    got = T[:, 0][:, inner]
    assert np.abs(got + rho0 * np.trace(B)).max() < 1e-11, np.abs(got + rho0 * np.trace(B)).max()


PINS = {
    "constant_viscosity_stress": (pin_constant_viscosity_stress,
                                  ["divergence_factor_dropped", "divergence_sign_flipped",
                                   "dynamic_viscosity_not_divided_by_rho"]),
    "buoyancy_frequency": (pin_buoyancy_frequency, ["buoyancy_with_unit_vertical"]),
    "free_stream_on_curved_mesh": (pin_free_stream_on_curved_mesh, ["tendency_metric_transposed"]),
    "linear_fields_on_sheared_mesh": (pin_linear_fields_on_sheared_mesh,
                                      ["tendency_metric_transposed", "gradient_metric_transposed"]),
}


@pytest.mark.parametrize("pin", sorted(PINS))
def test_pin_holds_for_the_oracle(pin):
    PINS[pin][0]()


@pytest.mark.parametrize("pin,mutant", [(p, m) for p in sorted(PINS) for m in PINS[p][1]])
def test_pin_fails_for_a_planted_mistake(pin, mutant):
    lib = build_mutant(mutant)
    with pytest.raises(AssertionError):
        PINS[pin][0](lib)

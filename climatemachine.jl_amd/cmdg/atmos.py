"""Dry AtmosModel configuration: parameters, auxiliary layout and set-up.

Mirrors the pieces of src/Atmos/Model that the DG hot path reads:

* the auxiliary variable layout of ``vars_state(::AtmosModel, ::Auxiliary)``
  (AtmosModel.jl:481-497): coord, orientation (Phi, grad Phi), ref_state
  (rho, p, T, rho e, rho q_tot, rho q_liq, rho q_ice), turbulence (Delta),
  moisture (theta_v, air_T);
* ``init_state_auxiliary!`` (AtmosModel.jl:908-925): Phi = g z and its DG
  gradient (Orientations.jl:101-127, 249-257), the hydrostatic reference state
  with density from the discrete pressure gradient (ref_state.jl:143-263),
  the Smagorinsky length scale (TurbulenceClosures.jl:437-444,
  Geometry.jl:121-122);
* initial conditions for the isentropic vortex
  (test/Numerics/DGMethods/Euler/isentropicvortex_setup.jl:14-62), the dry
  rising bubble, a rest atmosphere and the Taylor-Green vortex.

Set-up runs once on the host in FP64 (numpy); the time-stepping hot path is
the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from . import abi
from .grid import BrickGrid, MI_ID, M_ID, X1_ID, X2_ID, X3_ID, over_elements


# --------------------------------------------------------------------- parameters
@dataclass(frozen=True)
class Parameters:
    """Thermodynamic constants (CLIMAParameters Planet / SubgridScale names)."""
    R_d: float
    cp_d: float
    cv_d: float
    T_0: float = 273.16
    MSLP: float = 1.01325e5
    grav: float = 9.81
    C_smag: float = 0.21
    inv_Pr_turb: float = 3.0
    T_surf_ref: float = 290.0
    T_min_ref: float = 220.0

    @property
    def kappa_d(self) -> float:
        return self.R_d / self.cp_d


def earth_parameters() -> Parameters:
    """CLIMAParameters 0.2.0 values: R_d = R/M_d, cp_d = R_d/kappa_d, cv_d = cp_d - R_d."""
    R_d = 8.3144598 / 28.97e-3
    kappa_d = 2 / 7
    cp_d = R_d / kappa_d
    return Parameters(R_d=R_d, cp_d=cp_d, cv_d=cp_d - R_d)


def ideal_gas_parameters(R_d: float = 287.0, cp_d: float = 1004.5) -> Parameters:
    """Ideal-gas constants used by the benchmark workloads (R_d = 287, cp_d = 1004.5)."""
    return Parameters(R_d=R_d, cp_d=cp_d, cv_d=cp_d - R_d)


# --------------------------------------------------------------------- profiles
@dataclass(frozen=True)
class DryAdiabaticProfile:
    """TemperatureProfiles.jl:52-98: uniform potential temperature T_surface."""
    T_surface: float
    T_min_ref: float

    def __call__(self, prm: Parameters, z):
        Gamma = prm.grav / prm.cp_d
        T = np.maximum(self.T_surface - Gamma * z, self.T_min_ref)
        p = prm.MSLP * (T / self.T_surface) ** (prm.grav / (prm.R_d * Gamma))
        at_min = T == self.T_min_ref
        if np.any(at_min):
            z_top = (self.T_surface - self.T_min_ref) / Gamma
            H_min = prm.R_d * self.T_min_ref / prm.grav
            p = np.where(at_min, p * np.exp(-(z - z_top) / H_min), p)
        return T, p


@dataclass(frozen=True)
class DecayingTemperatureProfile:
    """TemperatureProfiles.jl:100-150 (IsothermalProfile is T_surf == T_min)."""
    T_virt_surf: float
    T_min_ref: float
    H_t: Optional[float] = None

    def __call__(self, prm: Parameters, z):
        H_t = self.H_t if self.H_t is not None else prm.R_d * self.T_virt_surf / prm.grav
        H_sfc = prm.R_d * self.T_virt_surf / prm.grav
        zp = z / H_t
        tz = np.tanh(zp)
        dT = self.T_virt_surf - self.T_min_ref
        Tv = self.T_virt_surf - dT * tz
        dTp = dT / self.T_virt_surf
        p = -H_t * (zp + dTp * (np.log(1 - dTp * tz) - np.log(1 + tz) + zp))
        p = p / (H_sfc * (1 - dTp ** 2))
        return Tv, prm.MSLP * np.exp(p)


# --------------------------------------------------------------------- model
@dataclass
class AtmosConfig:
    """The options of ``AtmosModel``/``AtmosPhysics`` that the hot path supports."""
    params: Parameters
    orientation: int = abi.NO_ORIENTATION
    ref_profile: Optional[object] = None     # HydrostaticState(profile) when set
    subtract_off: bool = True
    gravity_source: bool = False
    turbulence: int = abi.CONSTANT_DYNAMIC_VISCOSITY
    viscosity: float = 0.0
    with_divergence: bool = False
    C_smag: Optional[float] = None
    numerical_flux: int = abi.RUSANOV
    layout: dict = field(default_factory=dict, init=False)

    def __post_init__(self):
        self.layout = self.aux_layout()

    def aux_layout(self) -> dict:
        """Positions in ``vars_state(::AtmosModel, ::Auxiliary)`` order."""
        off = 0
        lay = {"coord": off}
        off += 3
        if self.orientation == abi.FLAT_ORIENTATION:
            lay["Phi"] = off
            lay["gradPhi"] = off + 1
            off += 4
        if self.ref_profile is not None:
            for i, name in enumerate(("ref_rho", "ref_p", "ref_T", "ref_rhoe", "ref_rhoq_tot",
                                      "ref_rhoq_liq", "ref_rhoq_ice")):
                lay[name] = off + i
            off += 7
        if self.turbulence in (abi.SMAGORINSKY_LILLY, abi.LEARNED_CLOSURE):
            lay["Delta"] = off
            off += 1
        lay["theta_v"] = off
        lay["air_T"] = off + 1
        off += 2
        lay["naux"] = off
        return lay

    @property
    def naux(self) -> int:
        return self.layout["naux"]

    @property
    def ngradflux(self) -> int:
        if self.turbulence == abi.LEARNED_CLOSURE:
            return 11      # Smagorinsky-Lilly's 10 + the learned nu_0 (mixed-precision path)
        return 10 if self.turbulence == abi.SMAGORINSKY_LILLY else 9

    def descriptor(self, grid: BrickGrid, aux: np.ndarray, precision: int = abi.FP64):
        """cmdg_desc for this model on ``grid``; the returned tuple keeps the arrays alive.
        On a lean grid the per-node geometry fields are null (cmdg_create_brick)."""
        lay = self.layout
        g = lambda k: lay.get(k, -1)
        prm = self.params
        per_node = (grid.vgeo, grid.sgeo, grid.vmap_minus, grid.vmap_plus, grid.elemtobndy)
        keep = [None if a is None else np.asfortranarray(a) for a in per_node + (grid.D, grid.omega, aux)]
        vgeo, sgeo, vmm, vmp, bnd, D, om, auxf = keep
        d = abi.CmdgDesc()
        d.N = grid.N
        d.nvgeo = 0 if vgeo is None else vgeo.shape[1]
        d.nelem = grid.nelem
        d.naux = self.naux
        d.aux_coord = g("coord")
        d.aux_Phi = g("Phi")
        d.aux_gradPhi = g("gradPhi")
        d.aux_ref_rho = g("ref_rho")
        d.aux_ref_p = g("ref_p")
        d.aux_Delta = g("Delta")
        d.aux_theta_v = g("theta_v")
        d.aux_air_T = g("air_T")
        d.vgeo = abi.as_dp(vgeo)
        d.sgeo = abi.as_dp(sgeo)
        d.vmap_minus = abi.as_ip(vmm)
        d.vmap_plus = abi.as_ip(vmp)
        d.elemtobndy = abi.as_ip(bnd)
        d.D = abi.as_dp(D)
        d.omega = abi.as_dp(om)
        d.aux_init = abi.as_dp(auxf)
        d.R_d, d.cp_d, d.cv_d, d.T_0 = prm.R_d, prm.cp_d, prm.cv_d, prm.T_0
        d.grav, d.MSLP = prm.grav, prm.MSLP
        d.orientation = self.orientation
        d.ref_state = abi.HYDROSTATIC_STATE if self.ref_profile is not None else abi.NO_REFERENCE_STATE
        d.subtract_off = int(bool(self.subtract_off))
        d.gravity_source = int(bool(self.gravity_source))
        d.turbulence = self.turbulence
        d.with_divergence = int(bool(self.with_divergence))
        d.viscosity = self.viscosity
        d.C_smag = self.C_smag if self.C_smag is not None else prm.C_smag
        d.inv_Pr_turb = prm.inv_Pr_turb
        d.numerical_flux = self.numerical_flux
        d.precision = precision
        return d, keep


# --------------------------------------------------------------------- DG helpers
def _tensor(f: np.ndarray, Nq: int) -> np.ndarray:
    """(Np, E) Fortran node order -> (Nq, Nq, Nq, E) indexed [i, j, k, e]."""
    return f.reshape(Nq, Nq, Nq, -1, order="F")


def field_gradient(grid: BrickGrid, f: np.ndarray) -> np.ndarray:
    """Volume-only strong derivative (dgsem_auxiliary_field_gradient!,
    DGModel_kernels.jl:3097-3232).  f (Np, E) -> (Np, 3, E)."""
    Nq, D = grid.Nq, grid.D
    ft = _tensor(f, Nq)
    g1 = np.einsum("im,mjke->ijke", D, ft).reshape(grid.Np, -1, order="F")
    g2 = np.einsum("jm,imke->ijke", D, ft).reshape(grid.Np, -1, order="F")
    g3 = np.einsum("km,ijme->ijke", D, ft).reshape(grid.Np, -1, order="F")
    v = grid.vgeo
    out = np.empty((grid.Np, 3, grid.nelem), order="F")
    for d in range(3):
        out[:, d, :] = (v[:, 0 + 3 * d, :] * g1 + v[:, 1 + 3 * d, :] * g2) + v[:, 2 + 3 * d, :] * g3
    return out


def pressure_gradient(grid: BrickGrid, p: np.ndarray) -> np.ndarray:
    """DG tendency of the flux -p I with central fluxes (PressureGradientModel,
    ref_state.jl:188-263): weak volume term plus surface lift; at walls the
    minus state is mirrored, so the face flux there is F(p-).n.  -> (Np, 3, E)."""
    Nq, Np, E = grid.Nq, grid.Np, grid.nelem
    v = grid.vgeo
    M, MI = v[:, M_ID, :], v[:, MI_ID, :]
    out = np.zeros((Np, 3, E), order="F")
    for c in range(3):
        # contravariant flux F~_a = M (xi_a x_c) (-p)
        Ft = [M * (v[:, a + 3 * c, :] * (-p)) for a in range(3)]
        t = np.einsum("mi,mjke->ijke", grid.D, _tensor(Ft[0], Nq))
        t = t + np.einsum("mj,imke->ijke", grid.D, _tensor(Ft[1], Nq))
        t = t + np.einsum("mk,ijme->ijke", grid.D, _tensor(Ft[2], Nq))
        out[:, c, :] = MI * t.reshape(Np, E, order="F")
    pflat = p.reshape(-1, order="F")
    for f in range(6):
        vm = grid.fmask[f]
        idp = grid.vmap_plus[:, f, :] - 1
        pp = pflat[idp]
        pm = p[vm, :]
        wall = grid.elemtobndy[f, :] != 0
        pp = np.where(wall[None, :], pm, pp)
        sM = grid.sgeo[3, :, f, :]
        vMI = grid.sgeo[4, :, f, :]
        for c in range(3):
            nc = grid.sgeo[c, :, f, :]
            flux = (-(pm + pp)) * (nc / 2)
            out[vm, c, :] -= vMI * sM * flux
    return out


def lengthscale(grid: BrickGrid) -> np.ndarray:
    """2 / (cbrt(det(d xi / d x)) N)  (Geometry.jl:121-122), (Np, E)."""
    if grid.lean:
        # the per-node formula below on diag(2/h): det = a00 (a11 a22), the same bits
        h1, h2, h3 = grid.element_sizes
        det = (2.0 / h1) * ((2.0 / h2) * (2.0 / h3))
        return np.broadcast_to((2.0 / (np.cbrt(det) * max(1, grid.N)))[None, :], (grid.Np, grid.nelem))
    v = grid.vgeo
    a = [[v[:, i + 3 * j, :] for j in range(3)] for i in range(3)]
    det = (a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
           - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
           + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]))
    return 2.0 / (np.cbrt(det) * max(1, grid.N))


def init_aux(cfg: AtmosConfig, grid: BrickGrid) -> np.ndarray:
    """Initial auxiliary state (Np, naux, nelem), Fortran order."""
    lay, prm = cfg.layout, cfg.params
    aux = np.zeros((grid.Np, cfg.naux, grid.nelem), order="F")
    x1, x2, x3 = grid.coords
    aux[:, lay["coord"] + 0, :] = x1
    aux[:, lay["coord"] + 1, :] = x2
    aux[:, lay["coord"] + 2, :] = x3
    if grid.lean and (cfg.orientation == abi.FLAT_ORIENTATION or cfg.ref_profile is not None):
        raise ValueError("a lean grid has no face geometry for the DG gradient of Phi and p_ref")
    if cfg.orientation == abi.FLAT_ORIENTATION:
        Phi = prm.grav * x3
        aux[:, lay["Phi"], :] = Phi
        aux[:, lay["gradPhi"]:lay["gradPhi"] + 3, :] = field_gradient(grid, Phi)
    if cfg.ref_profile is not None:
        if cfg.orientation != abi.FLAT_ORIENTATION:
            raise ValueError("a hydrostatic reference state needs FlatOrientation")
        Phi = aux[:, lay["Phi"], :]
        gPhi = aux[:, lay["gradPhi"]:lay["gradPhi"] + 3, :]
        z = Phi / prm.grav
        Tv, p = cfg.ref_profile(prm, z)
        gp = pressure_gradient(grid, p)
        k = gPhi / prm.grav
        kgp = (k[:, 0] * gp[:, 0] + k[:, 1] * gp[:, 1]) + k[:, 2] * gp[:, 2]
        kgPhi = (k[:, 0] * gPhi[:, 0] + k[:, 1] * gPhi[:, 1]) + k[:, 2] * gPhi[:, 2]
        rho = -kgp / kgPhi
        # ref_state_finalize_init! for the dry model: T from the ideal gas law
        T = p / (prm.R_d * rho)
        e_int = prm.cv_d * (T - prm.T_0)
        aux[:, lay["ref_rho"], :] = rho
        aux[:, lay["ref_p"], :] = p
        aux[:, lay["ref_T"], :] = prm.T_0 + e_int / prm.cv_d
        aux[:, lay["ref_rhoe"], :] = rho * ((0.0 + Phi) + e_int)
    if "Delta" in lay:
        aux[:, lay["Delta"], :] = lengthscale(grid)
    return aux


# --------------------------------------------------------------------- initial states
def total_energy(prm: Parameters, e_kin, e_pot, T):
    """Thermodynamics.total_energy for dry air: e_kin + e_pot + cv_d (T - T_0)."""
    return e_kin + e_pot + prm.cv_d * (T - prm.T_0)


@dataclass(frozen=True)
class IsentropicVortex:
    """IsentropicVortexSetup (isentropicvortex_setup.jl:3-62)."""
    p_inf: float = 1e5
    T_inf: float = 300.0
    rho_inf: Optional[float] = None
    translation_speed: float = 150.0
    translation_angle: float = np.pi / 4
    vortex_speed: float = 50.0
    vortex_radius: float = 1 / 200
    domain_halflength: float = 1 / 20

    def __call__(self, prm: Parameters, grid: BrickGrid, t: float = 0.0) -> np.ndarray:
        rho_inf = self.rho_inf if self.rho_inf is not None else self.p_inf / (prm.R_d * self.T_inf)
        L, R = self.domain_halflength, self.vortex_radius
        a = self.translation_angle
        uinf = np.array([self.translation_speed * np.cos(a), self.translation_speed * np.sin(a), 0.0])
        x = [c.copy() for c in grid.coords]
        for d in range(3):
            x[d] = x[d] - uinf[d] * t
            x[d] = x[d] - np.floor((x[d] + L) / (2 * L)) * (2 * L)
        r = np.sqrt(x[0] ** 2 + x[1] ** 2)
        du_x = -self.vortex_speed * x[1] / R * np.exp(-(r / R) ** 2 / 2)
        du_y = self.vortex_speed * x[0] / R * np.exp(-(r / R) ** 2 / 2)
        u = [uinf[0] + du_x, uinf[1] + du_y, uinf[2] + 0 * du_x]
        kappa = prm.kappa_d
        T = self.T_inf * (1 - kappa * self.vortex_speed ** 2 / 2 * rho_inf / self.p_inf * np.exp(-(r / R) ** 2))
        p = self.p_inf * (T / self.T_inf) ** (1 / kappa)
        rho = p / (prm.R_d * T)
        Q = np.empty((grid.Np, 5, grid.nelem), order="F")
        Q[:, 0, :] = rho
        for d in range(3):
            Q[:, 1 + d, :] = rho * u[d]
        e_kin = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / 2
        Q[:, 4, :] = rho * total_energy(prm, e_kin, 0.0, T)
        return Q


def rising_bubble(prm: Parameters, grid: BrickGrid, theta_ref: float = 300.0, dtheta: float = 2.0,
                  radius: float = 250.0, centre=(1250.0, 1250.0, 350.0), noise: float = 1e-3,
                  seed: int = 1) -> np.ndarray:
    """Dry rising thermal: cos^2 warm bubble on an isentropic background (after
    experiments/AtmosLES/rising_bubble_bryan.jl:45-82, dry, 3-D radius) plus
    seeded uniform white noise of amplitude ``noise`` K in theta."""
    x, y, z = grid.coords
    r = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2)
    dth = np.where(r <= radius, dtheta * np.cos(0.5 * np.pi * r / radius) ** 2, 0.0)
    rng = np.random.default_rng(seed)
    theta = theta_ref + dth + noise * rng.uniform(-1.0, 1.0, size=x.shape)
    exner = 1.0 - prm.grav / (prm.cp_d * theta) * z
    rho = prm.MSLP / (prm.R_d * theta) * exner ** (prm.cv_d / prm.R_d)
    T = theta * exner
    Q = np.zeros((grid.Np, 5, grid.nelem), order="F")
    Q[:, 0, :] = rho
    Q[:, 4, :] = rho * total_energy(prm, 0.0, prm.grav * z, T)
    return Q


def rest_state(cfg: AtmosConfig, aux: np.ndarray) -> np.ndarray:
    """State equal to the reference state at rest (discrete_hydrostatic_balance.jl:24-29)."""
    lay = cfg.layout
    Q = np.zeros((aux.shape[0], 5, aux.shape[2]), order="F")
    Q[:, 0, :] = aux[:, lay["ref_rho"], :]
    Q[:, 4, :] = aux[:, lay["ref_rhoe"], :]
    return Q


def taylor_green(prm: Parameters, grid: BrickGrid, mach: float = 0.1, rho0: float = 1.0,
                 T0: float = 300.0, phase=(0.0, 0.0, 0.0), L: float = 1.0) -> np.ndarray:
    """Taylor-Green vortex on a periodic (2 pi L)^3 box: u = U sin x cos y cos z,
    v = -U cos x sin y cos z, w = 0, p = p0 + rho0 U^2/16 (cos 2x + cos 2y)(cos 2z + 2),
    with U = Mach * c(T0) and seeded phase shifts (experiments/AtmosLES/taylor_green.jl:50-83)."""
    c0 = np.sqrt(prm.cp_d / prm.cv_d * prm.R_d * T0)
    U = mach * c0
    p0 = rho0 * prm.R_d * T0
    X = grid.coords
    Q = np.empty((grid.Np, 5, grid.nelem), order="F")

    def fill(e0, e1):
        x, y, z = (c[:, e0:e1] / L for c in X)
        x, y, z = x + phase[0], y + phase[1], z + phase[2]
        u = U * np.sin(x) * np.cos(y) * np.cos(z)
        v = -U * np.cos(x) * np.sin(y) * np.cos(z)
        w = 0 * u
        p = p0 + rho0 * U ** 2 / 16 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
        rho = rho0 * p / p0          # isothermal start
        T = p / (prm.R_d * rho)
        Q[:, 0, e0:e1] = rho
        Q[:, 1, e0:e1] = rho * u
        Q[:, 2, e0:e1] = rho * v
        Q[:, 3, e0:e1] = rho * w
        Q[:, 4, e0:e1] = rho * total_energy(prm, (u * u + v * v + w * w) / 2, 0.0, T)
    over_elements(grid.nelem, fill)
    return Q

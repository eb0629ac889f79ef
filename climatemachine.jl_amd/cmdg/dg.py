"""Host-side mirror of the reference's DG / ODE-solver interface.

* ``DGModel``: callable as ``dg(tendency, Q, param, t; increment)`` and
  ``dg(tendency, Q, param, t, alpha, beta)`` (SpaceDiscretization.jl:68-77,
  DGModel.jl:85), owning ``state_auxiliary`` like the reference's DGModel
  (DGModel.jl:3-65).  Every call goes to the HIP library.
* ``LSRK54CarpenterKennedy``: the 2N low-storage solver
  (LowStorageRungeKuttaMethod.jl:26-158, 293-327).  With a ``DGModel``
  right-hand side a step is one fused library call; any other callable
  right-hand side gets the reference's stage loop with the library's update!
  kernel.
* ``solve(Q, solver; timeend, numberofsteps, adjustfinalstep, callbacks)``:
  ``ODESolvers.solve!`` (ODESolvers.jl:49-158).

States live on the GPU as torch tensors of shape (nelem, nstate, Np), whose
memory is the reference's column-major (Np, nstate, nelem) MPIStateArray data.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from fractions import Fraction
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import abi
from .atmos import AtmosConfig, init_aux
from .grid import BrickGrid

# LowStorageRungeKuttaMethod.jl:302-324
_RKA = (Fraction(0), Fraction(-567301805773, 1357537059087), Fraction(-2404267990393, 2016746695238),
        Fraction(-3550918686646, 2091501179385), Fraction(-1275806237668, 842570457699))
_RKB = (Fraction(1432997174477, 9575080441755), Fraction(5161836677717, 13612068292357),
        Fraction(1720146321549, 2090206949498), Fraction(3134564353537, 4481467310338),
        Fraction(2277821191437, 14882151754819))
_RKC = (Fraction(0), Fraction(1432997174477, 9575080441755), Fraction(2526269341429, 6820363962896),
        Fraction(2006345519317, 3224310063776), Fraction(2802321613138, 2924317926251))


def _rt(x: Fraction) -> float:
    # RT(a // b) with RT = Float64: Float64(a) / Float64(b)
    return float(x.numerator) / float(x.denominator)


RKA = tuple(_rt(x) for x in _RKA)
RKB = tuple(_rt(x) for x in _RKB)
RKC = tuple(_rt(x) for x in _RKC)


class CmdgError(RuntimeError):
    pass


def to_device(a: np.ndarray, device="cuda") -> torch.Tensor:
    """(Np, nstate, nelem) Fortran array -> (nelem, nstate, Np) device tensor."""
    t = torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).T))
    return t.to(device=device, dtype=torch.float64).contiguous()


def to_host(t: torch.Tensor) -> np.ndarray:
    """(nelem, nstate, Np) tensor -> (Np, nstate, nelem) Fortran numpy array."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def stress_form_layout(ngradflux: int) -> dict:
    """Planes of the mixed-precision gradient flux (stress form) for its plane count:
    ``tau`` the viscous stress factor -2 nu_d S_d. (+ 2 nu / 3 tr S on the diagonal
    with divergence): 6 symmetric planes xx xy xz yy yz zz when the viscosity is
    isotropic (``iso``, 10 planes in all), else 9 row-major (15 planes); ``e`` the
    three (nu_d / Pr_t) dh_tot/dx_d; ``nu`` the viscosity (1 plane, or nu_d).  Row d of
    the second-order flux is (0, rho tau_d., tau_d . rho u - rho e_d)."""
    if ngradflux not in (10, 15):
        raise CmdgError(f"not a stress-form plane count: {ngradflux}")
    iso = ngradflux == 10
    ntau = 6 if iso else 9
    sym = ((0, 1, 2), (1, 3, 4), (2, 4, 5))
    tau = [[sym[d][c] if iso else 3 * d + c for c in range(3)] for d in range(3)]
    return {"iso": iso, "tau": tau, "e": [ntau, ntau + 1, ntau + 2], "nu": list(range(ntau + 3, ngradflux))}


def _ptr(t: torch.Tensor) -> int:
    if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
        raise CmdgError("state arrays must be contiguous float64 CUDA tensors")
    return t.data_ptr()


class DGModel:
    """DGModel(balance_law, grid, ...) backed by the HIP library."""

    def __init__(self, config: AtmosConfig, grid: BrickGrid, precision: int = abi.FP64,
                 aux: Optional[np.ndarray] = None, device: str = "cuda"):
        self.lib = abi.load_library()
        if not torch.cuda.is_available():
            raise CmdgError("no GPU: the DG right-hand side runs only through the HIP library")
        self.balance_law = config
        self.grid = grid
        self.device = torch.device(device)
        self.precision = precision
        aux0 = init_aux(config, grid) if aux is None else np.asfortranarray(aux)
        desc, keep = config.descriptor(grid, aux0, precision)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            ctx = C.c_void_p()
            if getattr(grid, "lean", False):
                brick, bkeep = abi.brick_descriptor(grid)
                rc = self.lib.cmdg_create_brick(C.byref(desc), C.byref(brick), C.c_void_p(stream), C.byref(ctx))
                del bkeep
            else:
                rc = self.lib.cmdg_create(C.byref(desc), C.c_void_p(stream), C.byref(ctx))
        self._ctx = ctx
        if rc != abi.CMDG_OK:
            msg = self.lib.cmdg_last_error(ctx).decode() if ctx else "cmdg_create failed"
            self.lib.cmdg_destroy(ctx)
            self._ctx = None
            raise CmdgError(f"cmdg_create: {msg} (status {rc})")
        del keep
        self.nelem, self.Np = grid.nelem, grid.Np

    # ----------------------------------------------------------------- helpers
    def _check(self, rc: int, what: str):
        if rc != abi.CMDG_OK:
            raise CmdgError(f"{what}: {self.lib.cmdg_last_error(self._ctx).decode()} (status {rc})")

    def new_state(self) -> torch.Tensor:
        return torch.zeros((self.nelem, 5, self.Np), dtype=torch.float64, device=self.device)

    @property
    def state_auxiliary(self) -> torch.Tensor:
        """Device view (nelem, naux, Np) of the context-owned auxiliary state."""
        p = C.c_void_p()
        self._check(self.lib.cmdg_aux(self._ctx, C.byref(p)), "cmdg_aux")
        naux = self.balance_law.naux
        return _wrap_device(p.value, (self.nelem, naux, self.Np), self.device)

    @property
    def state_gradient_flux(self) -> torch.Tensor:
        p = C.c_void_p()
        n = C.c_int32()
        self._check(self.lib.cmdg_gradient_flux(self._ctx, C.byref(p), C.byref(n)), "cmdg_gradient_flux")
        return _wrap_device(p.value, (self.nelem, n.value, self.Np), self.device)

    @property
    def state_gradient_flux_f32(self) -> torch.Tensor:
        """Device view (nelem, ngradflux, Np) of the mixed-precision FP32 gradient flux, in
        stress form (``stress_form_layout`` names its planes)."""
        p = C.c_void_p()
        n = C.c_int32()
        self._check(self.lib.cmdg_gradient_flux_f32(self._ctx, C.byref(p), C.byref(n)), "cmdg_gradient_flux_f32")
        return _wrap_device(p.value, (self.nelem, n.value, self.Np), self.device, "<f4")

    def set_closure(self, mlp) -> None:
        """Load a learned-closure network (an ``closure.MLPParams``) into the context."""
        w1, b1, w2, b2, w3, b3, shift, scale = mlp.export()
        keep = (w1, b1, w2, b2, w3, b3, shift, scale)
        ptrs = [C.c_void_p(a.ctypes.data) for a in keep]
        self._check(self.lib.cmdg_mlp_set_weights(self._ctx, *ptrs), "cmdg_mlp_set_weights")
        del keep

    def closure_forward(self, x_raw: torch.Tensor) -> torch.Tensor:
        """The context's network on raw input rows (n, 10) float32 CUDA -> (n,)."""
        if x_raw.dtype != torch.float32 or not x_raw.is_contiguous() or not x_raw.is_cuda:
            raise CmdgError("inputs must be a contiguous float32 CUDA tensor")
        y = torch.empty(x_raw.shape[0], dtype=torch.float32, device=x_raw.device)
        self._check(self.lib.cmdg_mlp_forward(self._ctx, C.c_void_p(x_raw.data_ptr()), C.c_void_p(y.data_ptr()),
                                              int(x_raw.shape[0])), "cmdg_mlp_forward")
        return y

    def velocity_gradient(self, Q: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """du_c/dx_d of Q as the gradient pass forms it: (nelem, 9, Np), entry 3c + d."""
        if out is None:
            out = torch.empty((self.nelem, 9, self.Np), dtype=torch.float64, device=Q.device)
        self._check(self.lib.cmdg_velocity_gradient(self._ctx, C.c_void_p(_ptr(Q)), C.c_void_p(_ptr(out))),
                    "cmdg_velocity_gradient")
        return out

    def init_ode_state(self, Q0: np.ndarray) -> torch.Tensor:
        """Upload an initial state built on the host (init_ode_state with init_on_cpu)."""
        return to_device(Q0, self.device)

    # ----------------------------------------------------------------- calls
    def __call__(self, tendency, Q, param, t, alpha=None, beta=None, *, increment: bool = False):
        if alpha is None:
            alpha, beta = 1.0, (1.0 if increment else 0.0)
        self._check(self.lib.cmdg_rhs(self._ctx, C.c_void_p(_ptr(tendency)), C.c_void_p(_ptr(Q)), float(t),
                                      float(alpha), float(beta)), "cmdg_rhs")

    def lsrk54_step(self, Q, dQ, t, dt):
        self._check(self.lib.cmdg_lsrk54_step(self._ctx, C.c_void_p(_ptr(Q)), C.c_void_p(_ptr(dQ)), float(t),
                                              float(dt)), "cmdg_lsrk54_step")

    def lsrk54_steps(self, Q, dQ, t, dt, nsteps: int):
        self._check(self.lib.cmdg_lsrk54_steps(self._ctx, C.c_void_p(_ptr(Q)), C.c_void_p(_ptr(dQ)), float(t),
                                               float(dt), int(nsteps)), "cmdg_lsrk54_steps")

    def lsrk_stage(self, Q, dQ, t_stage, dt, rka_next, rkb):
        self._check(self.lib.cmdg_lsrk_stage(self._ctx, C.c_void_p(_ptr(Q)), C.c_void_p(_ptr(dQ)),
                                             float(t_stage), float(dt), float(rka_next), float(rkb)),
                    "cmdg_lsrk_stage")

    def courant(self, Q, dt: float, kind: str = "nondiffusive", direction: str = "every") -> float:
        """courant(local_courant, dg, m, Q, dt, t, direction): the largest local
        Courant number; kind advective / nondiffusive / diffusive."""
        k = {"advective": 0, "nondiffusive": 1, "diffusive": 2}[kind]
        d = {"every": 0, "horizontal": 1, "vertical": 2}[direction]
        out = C.c_double()
        self._check(self.lib.cmdg_courant(self._ctx, C.c_void_p(_ptr(Q)), float(dt), k, d, C.byref(out)),
                    "cmdg_courant")
        return out.value

    def weightedsum(self, Q, state: int) -> float:
        """weightedsum(Q, state): sum of M * Q over all nodes (double-double)."""
        out = C.c_double()
        self._check(self.lib.cmdg_weightedsum(self._ctx, C.c_void_p(_ptr(Q)), int(state), C.byref(out)),
                    "cmdg_weightedsum")
        return out.value

    def not_finite_fields(self, Q) -> list:
        """Names of the state fields of Q that hold a NaN or Inf (the device map-reduce
        of show_not_finite_fields, MPIStateArrays.jl:884-903)."""
        m = C.c_int32()
        self._check(self.lib.cmdg_not_finite_fields(self._ctx, C.c_void_p(_ptr(Q)), C.byref(m)),
                    "cmdg_not_finite_fields")
        return fields_of_mask(m.value)

    def show_not_finite_fields(self, Q) -> list:
        """Warn about (and return) the fields of Q that are not finite."""
        bad = self.not_finite_fields(Q)
        if bad:
            warnings.warn(f"Field(s) ({', '.join(bad)}) are not finite (has NaNs or Inf)")
        return bad

    @property
    def geometry_kind(self) -> str:
        k = C.c_int32()
        self._check(self.lib.cmdg_geometry_kind(self._ctx, C.byref(k)), "cmdg_geometry_kind")
        return "affine" if k.value & 1 else "general"

    @property
    def static_aux_continuous(self) -> bool:
        """True when Phi and p_ref agree bitwise across every interior face of the
        affine brick, so the face kernels read no neighbour static trace."""
        k = C.c_int32()
        self._check(self.lib.cmdg_geometry_kind(self._ctx, C.byref(k)), "cmdg_geometry_kind")
        return bool(k.value & 2)

    @property
    def static_aux_classes(self) -> bool:
        """True when the affine kernels read the static auxiliary fields (Phi, grad Phi,
        rho_ref, p_ref, Delta) from shared per-class blocks: elements whose fields agree
        bitwise (one class per element layer on a brick with a height-only reference)."""
        k = C.c_int32()
        self._check(self.lib.cmdg_geometry_kind(self._ctx, C.byref(k)), "cmdg_geometry_kind")
        return bool(k.value & 4)

    @property
    def tiled_traversal(self) -> bool:
        """True when an affine pass (by default the tendency pass) walks the elements in
        lattice tiles (so that a block's face neighbours ran shortly before it) rather
        than in storage order."""
        k = C.c_int32()
        self._check(self.lib.cmdg_geometry_kind(self._ctx, C.byref(k)), "cmdg_geometry_kind")
        return bool(k.value & 8)

    def set_kernel_timing(self, enable: bool):
        self._check(self.lib.cmdg_set_kernel_timing(self._ctx, int(bool(enable))), "cmdg_set_kernel_timing")

    def kernel_timing(self) -> dict:
        """{kind: (total_ms, launches)} since the last call (synchronises)."""
        n = len(abi.KERNEL_KINDS)
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        self._check(self.lib.cmdg_kernel_timing(self._ctx, ms, cnt, n), "cmdg_kernel_timing")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(abi.KERNEL_KINDS)}

    def synchronize(self):
        self._check(self.lib.cmdg_sync(self._ctx), "cmdg_sync")

    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.cmdg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _wrap_device(ptr: int, shape, device, typestr: str = "<f8") -> torch.Tensor:
    """A torch view of library-owned device memory (valid while the context lives)."""
    n = int(np.prod(shape))

    class _Iface:
        __cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 2,
                                    "strides": None}
    with torch.cuda.device(device):
        return torch.as_tensor(_Iface(), device=device).view(*shape)


def lsrk_update(dQ: torch.Tensor, Q: torch.Tensor, rka: float, rkb: float, dt: float):
    """The update! kernel on any pair of float64 device tensors."""
    lib = abi.load_library()
    stream = torch.cuda.current_stream(Q.device).cuda_stream
    rc = lib.cmdg_lsrk_update(C.c_void_p(_ptr(dQ)), C.c_void_p(_ptr(Q)), Q.numel(), float(rka), float(rkb),
                              float(dt), C.c_void_p(stream))
    if rc != abi.CMDG_OK:
        raise CmdgError(f"cmdg_lsrk_update failed ({rc})")


class LSRK54CarpenterKennedy:
    """LowStorageRungeKutta2N with the Carpenter-Kennedy (5,4) coefficients."""

    RKA, RKB, RKC = RKA, RKB, RKC

    def __init__(self, rhs: Callable, Q: torch.Tensor, dt: float = 0.0, t0: float = 0.0):
        self.rhs = rhs
        self.dt = float(dt)
        self.t = float(t0)
        self.steps = 0
        self.dQ = torch.zeros_like(Q)

    # ODESolvers accessor names
    def gettime(self):
        return self.t

    def getdt(self):
        return self.dt

    def updatedt(self, dt):
        self.dt = float(dt)

    def updatetime(self, t):
        self.t = float(t)
        return self.t

    def dostep(self, Q: torch.Tensor, p, time: float):
        """dostep! (LowStorageRungeKuttaMethod.jl:102-144)."""
        if isinstance(self.rhs, DGModel):
            self.rhs.lsrk54_step(Q, self.dQ, time, self.dt)
            return
        for s in range(5):
            self.rhs(self.dQ, Q, p, time + RKC[s] * self.dt, increment=True)
            lsrk_update(self.dQ, Q, RKA[(s + 1) % 5], RKB[s], self.dt)

    def dosteps(self, Q: torch.Tensor, p, time: float, nsteps: int):
        """nsteps calls of dostep! with the same dt and nothing in between (the
        solve! loop without callbacks) in one library call."""
        if isinstance(self.rhs, DGModel):
            self.rhs.lsrk54_steps(Q, self.dQ, time, self.dt, nsteps)
            return
        for k in range(nsteps):
            self.dostep(Q, p, time + k * self.dt)


STATE_FIELDS = ("ρ", "ρu[1]", "ρu[2]", "ρu[3]", "energy.ρe")


def fields_of_mask(mask: int) -> list:
    """Field names (AtmosModel prognostic order, AtmosModel.jl:387-403) of the set bits."""
    return [name for s, name in enumerate(STATE_FIELDS) if mask >> s & 1]


class NotFiniteCheck:
    """solve callback that stops on a crash: every `every` steps it asks the library
    which fields of Q are not finite and raises FloatingPointError naming them (the
    reference stops through show_not_finite_fields and an exception,
    MPIStateArrays.jl:884-927)."""

    def __init__(self, dg, every: int = 1):
        if every < 1:
            raise ValueError("every must be >= 1")
        self.dg, self.every = dg, every

    def __call__(self, solver, Q, param, time):
        if solver.steps % self.every:
            return None
        bad = self.dg.not_finite_fields(Q)
        if bad:
            raise FloatingPointError(f"step {solver.steps}, t = {time}: field(s) ({', '.join(bad)}) are not finite")
        return None


def solve(Q: torch.Tensor, solver, param=None, *, timeend: float = math.inf, adjustfinalstep: bool = True,
          numberofsteps: int = 0, callbacks: Sequence[Callable] = ()) -> float:
    """ODESolvers.solve! (ODESolvers.jl:110-158) with general_dostep! (:49-74).

    Callbacks are called as cb(solver, Q, param, time) after every step; a
    return value > 0 stops the integration."""
    if not (math.isfinite(timeend) or numberofsteps > 0):
        raise ValueError("give timeend or numberofsteps")
    step = 0
    time = solver.gettime()
    while time < timeend:
        if not callbacks and hasattr(solver, "dosteps"):
            # the whole steps ahead that no callback and no final-step adjustment
            # separate run as one batch; time advances by the same additions
            t, dt = solver.gettime(), solver.getdt()
            tt, n = t, 0
            while tt < timeend and (numberofsteps <= 0 or step + n < numberofsteps) \
                    and not (adjustfinalstep and tt + dt > timeend):
                tt += dt
                n += 1
            if n > 1:
                solver.dosteps(Q, param, t, n)
                step += n
                solver.steps = step
                time = solver.updatetime(tt)
                if step == numberofsteps:
                    break
                continue
        step += 1
        solver.steps = step
        t, dt = solver.gettime(), solver.getdt()
        final_step = False
        if adjustfinalstep and t + dt > timeend:
            orig_dt = dt
            dt = timeend - t
            solver.updatedt(dt)
            final_step = True
        assert dt > 0
        solver.dostep(Q, param, t)
        if not final_step:
            time = solver.updatetime(t + dt)
        else:
            solver.updatedt(orig_dt)
            time = solver.updatetime(timeend)
        for cb in callbacks:
            val = cb(solver, Q, param, time)
            if val is not None and val > 0:
                return solver.gettime()
        if step == numberofsteps:
            break
    return solver.gettime()

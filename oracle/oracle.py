"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, and only as the checker / CPU baseline.  The
oracle (oracle/cmdg_oracle.cpp) is a plain restatement of the reference's DG
right-hand side and LSRK54 stepper; see that file's header for what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "climatemachine.jl_amd"))

from cmdg import abi  # noqa: E402  (descriptor layout shared with the library)
from cmdg.atmos import AtmosConfig  # noqa: E402
from cmdg.grid import BrickGrid  # noqa: E402

LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
_libs = {}


def load(path: str = LIB_PATH):
    """The oracle library at `path` (default: the built oracle) with its argument
    types declared; other paths are test builds of modified oracle sources."""
    lib = _libs.get(path)
    if lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        lib = C.CDLL(path)
        vp = C.c_void_p
        lib.oracle_create.argtypes = [C.POINTER(abi.CmdgDesc), C.c_int]
        lib.oracle_create.restype = vp
        lib.oracle_destroy.argtypes = [vp]
        lib.oracle_rhs.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double]
        lib.oracle_lsrk54_step.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double]
        lib.oracle_gradient_flux.argtypes = [vp, vp]
        lib.oracle_field_gradient.argtypes = [vp, vp, C.c_int, C.c_int]
        lib.oracle_pressure_gradient.argtypes = [vp, vp, vp]
        lib.oracle_lsrk54_coefficient.argtypes = [C.c_int, C.c_int, C.c_int]
        lib.oracle_lsrk54_coefficient.restype = C.c_double
        lib.oracle_num_threads.restype = C.c_int
        lib.oracle_flux_second_order.argtypes = [vp, vp, vp, vp, C.c_double, vp]
        lib.oracle_mlp_set.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, vp, vp]
        lib.oracle_mlp_forward.argtypes = [vp, vp, C.c_int64, vp]
        lib.oracle_courant.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int]
        lib.oracle_courant.restype = C.c_double
        lib.oracle_weightedsum.argtypes = [vp, vp, C.c_int]
        lib.oracle_weightedsum.restype = C.c_double
        lib.oracle_set_prescribed.argtypes = [vp, vp, vp, C.c_int]
        _libs[path] = lib
    return lib


def _p(a: np.ndarray):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1
    return C.c_void_p(a.ctypes.data)


class Oracle:
    """Oracle DG model on ``grid`` with the options of ``cfg``."""

    def __init__(self, cfg: AtmosConfig, grid: BrickGrid, aux: np.ndarray, use_float: bool = False,
                 lib_path: str = LIB_PATH):
        self.lib = load(lib_path)
        self.cfg, self.grid = cfg, grid
        self.dtype = np.float32 if use_float else np.float64
        desc, keep = cfg.descriptor(grid, np.asfortranarray(aux, dtype=np.float64))
        self._ctx = self.lib.oracle_create(C.byref(desc), int(use_float))
        if not self._ctx:
            raise RuntimeError("oracle_create failed")
        self.aux = np.asfortranarray(aux, dtype=self.dtype).copy(order="F")

    def close(self):
        if self._ctx:
            self.lib.oracle_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        assert a.dtype == self.dtype and a.flags["F_CONTIGUOUS"], (a.dtype, a.flags)
        return _p(a)

    def rhs(self, tend: np.ndarray, Q: np.ndarray, alpha: float = 1.0, beta: float = 0.0):
        self.lib.oracle_rhs(self._ctx, self._arr(tend), self._arr(Q), self._arr(self.aux), alpha, beta)

    def tendency(self, Q: np.ndarray) -> np.ndarray:
        out = np.zeros_like(Q, order="F")
        self.rhs(out, np.asfortranarray(Q, dtype=self.dtype))
        return out

    def lsrk54_step(self, Q: np.ndarray, dQ: np.ndarray, t: float, dt: float):
        self.lib.oracle_lsrk54_step(self._ctx, self._arr(Q), self._arr(dQ), self._arr(self.aux), t, dt)

    @property
    def ngradflux(self) -> int:
        # the learned closure also keeps the assembled velocity gradient (9) before nu_0
        return 20 if self.cfg.turbulence == abi.LEARNED_CLOSURE else self.cfg.ngradflux

    def gradient_flux(self) -> np.ndarray:
        """(Np, ngradflux, nelem); learned closure: [grad h, S, N2, grad u (9), nu_0]."""
        out = np.zeros((self.grid.Np, self.ngradflux, self.grid.nelem), order="F")
        self.lib.oracle_gradient_flux(self._ctx, _p(out))
        return out

    def field_gradient(self, f: np.ndarray) -> np.ndarray:
        """DG volume gradient of a nodal field f (Np, E) -> (Np, 3, E)."""
        tmp = np.zeros((self.grid.Np, self.cfg.naux, self.grid.nelem), order="F")
        tmp[:, 0, :] = f
        assert self.cfg.naux >= 4
        self.lib.oracle_field_gradient(self._ctx, _p(tmp), 0, 1)
        return np.asfortranarray(tmp[:, 1:4, :])

    def flux_second_order(self, q, gf, gPhi=(0.0, 0.0, 0.0), Delta=0.0) -> np.ndarray:
        """Second-order flux F[d, s] at one node."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        gf = np.ascontiguousarray(gf, dtype=np.float64)
        gp = np.ascontiguousarray(gPhi, dtype=np.float64)
        F = np.zeros(15)
        self.lib.oracle_flux_second_order(self._ctx, _p(q), _p(gf), _p(gp), float(Delta), _p(F))
        return F.reshape(3, 5)

    def courant(self, Q, dt, kind=1, direction=0) -> float:
        """Largest local Courant number (kind 0 advective, 1 nondiffusive, 2 diffusive
        with the gradient flux of the last RHS; direction 0 every, 1 horizontal, 2 vertical)."""
        Qf = np.asfortranarray(Q, dtype=np.float64)
        return self.lib.oracle_courant(self._ctx, _p(Qf), _p(self.aux), float(dt), int(kind), int(direction))

    def weightedsum(self, Q, state: int) -> float:
        Qf = np.asfortranarray(Q, dtype=np.float64)
        return self.lib.oracle_weightedsum(self._ctx, _p(Qf), int(state))

    def set_prescribed(self, qbnd=None, src=None, htot_zero=False):
        """Manufactured-solution set-up: qbnd (Np, 5, E), the state InitStateBC
        prescribes (read at boundary-face nodes), src (Np, 5, E) a per-node source,
        htot_zero the zero total specific enthalpy override.  The arrays are kept
        referenced here and read by later rhs calls."""
        self._prescribed = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (qbnd, src)]
        qb, sr = self._prescribed
        self.lib.oracle_set_prescribed(self._ctx, None if qb is None else _p(qb), None if sr is None else _p(sr),
                                       int(bool(htot_zero)))

    def set_mlp(self, mlp):
        """Load a closure.MLPParams network (weights rounded to BF16 inside)."""
        f = mlp.flat.astype(np.float64)
        parts = [np.ascontiguousarray(a) for a in (f[0:320], f[320:352], f[352:1376], f[1376:1408], f[1408:1440],
                                                    mlp.shift.astype(np.float64), mlp.scale.astype(np.float64))]
        w1, b1, w2, b2, w3, sh, sc = parts
        self.lib.oracle_mlp_set(self._ctx, _p(w1), _p(b1), _p(w2), _p(b2), _p(w3), float(f[1440]), _p(sh), _p(sc))

    def mlp_forward(self, raw: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw, dtype=np.float64)
        y = np.zeros(raw.shape[0])
        self.lib.oracle_mlp_forward(self._ctx, _p(raw.ravel()), raw.shape[0], _p(y))
        return y

    def pressure_gradient(self, p: np.ndarray) -> np.ndarray:
        out = np.zeros((self.grid.Np, 3, self.grid.nelem), order="F")
        pf = np.asfortranarray(p, dtype=np.float64)  # keep the (possibly copied) input alive
        self.lib.oracle_pressure_gradient(self._ctx, _p(pf), _p(out))
        return out


def lsrk54_coefficients(use_float: bool = False):
    lib = load()
    return [[lib.oracle_lsrk54_coefficient(w, k, int(use_float)) for k in range(5)] for w in range(3)]


def solve(orc: Oracle, Q: np.ndarray, dt: float, timeend: float, t0: float = 0.0):
    """ODESolvers.solve! loop (ODESolvers.jl:110-158) around the oracle step."""
    dQ = np.zeros_like(Q, order="F")
    dtype = orc.dtype
    t = dtype(t0)
    dt = dtype(dt)
    timeend = dtype(timeend)
    while t < timeend:
        step_dt = dt
        final = False
        if t + dt > timeend:
            step_dt = dtype(timeend - t)
            final = True
        orc.lsrk54_step(Q, dQ, float(t), float(step_dt))
        t = timeend if final else dtype(t + dt)
    return float(t)


def euclidean_distance(A: np.ndarray, B: np.ndarray, M: np.ndarray) -> float:
    """MPIStateArrays.euclidean_distance: sqrt(sum (A-B)^2 * M) (MPIStateArrays.jl:628-644)."""
    E = (A.astype(np.float64) - B.astype(np.float64)) ** 2 * M[:, None, :]
    return float(np.sqrt(E.sum()))


def oracle_init_aux(cfg: AtmosConfig, grid: BrickGrid) -> np.ndarray:
    """Auxiliary state from the oracle's own DG operators (grad Phi, grad p_ref)."""
    from cmdg import atmos
    lay, prm = cfg.layout, cfg.params
    aux = np.zeros((grid.Np, cfg.naux, grid.nelem), order="F")
    x1, x2, x3 = grid.coords
    aux[:, lay["coord"]:lay["coord"] + 3, :] = np.stack([x1, x2, x3], axis=1)
    orc = Oracle(cfg, grid, aux)
    if cfg.orientation == abi.FLAT_ORIENTATION:
        Phi = prm.grav * x3
        aux[:, lay["Phi"], :] = Phi
        aux[:, lay["gradPhi"]:lay["gradPhi"] + 3, :] = orc.field_gradient(Phi)
    if cfg.ref_profile is not None:
        Phi = aux[:, lay["Phi"], :]
        gPhi = aux[:, lay["gradPhi"]:lay["gradPhi"] + 3, :]
        Tv, p = cfg.ref_profile(prm, Phi / prm.grav)
        gp = orc.pressure_gradient(p)
        k = gPhi / prm.grav
        rho = -((k[:, 0] * gp[:, 0] + k[:, 1] * gp[:, 1]) + k[:, 2] * gp[:, 2]) / (
            (k[:, 0] * gPhi[:, 0] + k[:, 1] * gPhi[:, 1]) + k[:, 2] * gPhi[:, 2])
        T = p / (prm.R_d * rho)
        e_int = prm.cv_d * (T - prm.T_0)
        aux[:, lay["ref_rho"], :] = rho
        aux[:, lay["ref_p"], :] = p
        aux[:, lay["ref_T"], :] = prm.T_0 + e_int / prm.cv_d
        aux[:, lay["ref_rhoe"], :] = rho * ((0.0 + Phi) + e_int)
    if "Delta" in lay:
        aux[:, lay["Delta"], :] = atmos.lengthscale(grid)
    orc.close()
    return aux

"""Learned eddy-viscosity closure and the tools that train and evaluate it.

The closure replaces nu_0 of SmagorinskyLilly
(src/Common/TurbulenceClosures/TurbulenceClosures.jl:478-491) by

    nu_0 = max(nu_floor + clamp(y, -e^-1, e^-1) * Delta^2 * |S|, 0),   y = MLP((raw - shift) * scale),

raw = (du_c/dx_d at 3c + d, Delta), a 10 -> 32 -> 32 -> 1 network with SiLU
hidden units, nu_floor the context's kinematic viscosity (1e-5 when zero): a signed
eddy-viscosity coefficient, whose negative values take viscosity away (backscatter)
down to none in total.  The right-hand side evaluates it with BF16 MFMA tiles fused into
the mixed-precision gradient kernel; this module holds the FP32 master
parameters, exports them as BF16, and drives the library's training, filter,
sample and spectrum kernels.  It has no counterpart in the reference except the
spectrum (power_spectrum_3d, src/Common/Spectra/power_spectrum_les.jl:13-75) and
the DG -> uniform-grid interpolation (InterpolationBrick,
src/Numerics/Mesh/Interpolation.jl:66-571) it feeds.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np
import torch

from . import abi
from .grid import BrickGrid, baryweights, lglpoints

NPARAM = abi.MLP_NPARAM
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 320, 352, 1376, 1408, 1440
NIN, NH = 10, 32
SAMPLE_W = 13        # raw inputs (10), then P, Q, R of the stress loss
CMAX = math.exp(-1.0)  # output clamp: |nu_0 - nu_floor| <= 0.37 Delta^2 |S| (|C_s| <= 0.6)
NMOM = 11            # filtered fields: rho, rho u (3), rho e, rho u_a u_b (6)


class ClosureError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != abi.CMDG_OK:
        raise ClosureError(f"{what} failed (status {rc})")


def _vp(t: torch.Tensor) -> C.c_void_p:
    if not t.is_cuda or not t.is_contiguous():
        raise ClosureError("device arrays must be contiguous CUDA tensors")
    return C.c_void_p(t.data_ptr())


def _stream(device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


# ----------------------------------------------------------------- BF16
def bf16_bits(a) -> np.ndarray:
    """Round float32 values to BF16 (nearest, ties to even); returns the uint16 bits."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_round(a) -> np.ndarray:
    """The float32 values of bf16_bits(a)."""
    return (bf16_bits(a).astype(np.uint32) << 16).view(np.float32)


# ----------------------------------------------------------------- network
class MLPParams:
    """FP32 master parameters (flat, CMDG_MLP_NPARAM) plus the input normalisation."""

    def __init__(self, flat, shift=None, scale=None):
        self.flat = np.ascontiguousarray(flat, dtype=np.float32).copy()
        if self.flat.shape != (NPARAM,):
            raise ValueError(f"expected {NPARAM} parameters")
        self.shift = np.zeros(NIN, np.float32) if shift is None else np.asarray(shift, np.float32).copy()
        self.scale = np.ones(NIN, np.float32) if scale is None else np.asarray(scale, np.float32).copy()

    @classmethod
    def random(cls, seed: int = 0, y0: float = 0.0, shift=None, scale=None) -> "MLPParams":
        """Glorot-uniform weights, zero hidden biases, output bias y0."""
        rng = np.random.default_rng(seed)
        flat = np.zeros(NPARAM, np.float32)
        for off, fan_out, fan_in in ((OFF_W1, NH, NIN), (OFF_W2, NH, NH), (OFF_W3, 1, NH)):
            lim = math.sqrt(6.0 / (fan_in + fan_out))
            flat[off:off + fan_out * fan_in] = rng.uniform(-lim, lim, fan_out * fan_in)
        flat[OFF_B3] = y0
        return cls(flat, shift, scale)

    @classmethod
    def constant(cls, y0: float) -> "MLPParams":
        """A network whose output is y0 everywhere: with y0 = C_s^2 (and the floor 1e-5
        of a context without viscosity) it is Smagorinsky-Lilly's
        nu_0 = |S| (C_s Delta)^2 + 1e-5."""
        flat = np.zeros(NPARAM, np.float32)
        flat[OFF_B3] = y0
        return cls(flat)

    # views
    @property
    def W1(self):
        return self.flat[OFF_W1:OFF_B1].reshape(NH, NIN)

    @property
    def b1(self):
        return self.flat[OFF_B1:OFF_W2]

    @property
    def W2(self):
        return self.flat[OFF_W2:OFF_B2].reshape(NH, NH)

    @property
    def b2(self):
        return self.flat[OFF_B2:OFF_W3]

    @property
    def W3(self):
        return self.flat[OFF_W3:OFF_B3]

    @property
    def b3(self):
        return float(self.flat[OFF_B3])

    def export(self):
        """(w1 bits, b1, w2 bits, b2, w3 bits, b3, shift, scale) for cmdg_mlp_set_weights."""
        return (bf16_bits(self.W1.ravel()), self.b1.copy(), bf16_bits(self.W2.ravel()), self.b2.copy(),
                bf16_bits(self.W3), np.array([self.b3], np.float32), self.shift.copy(), self.scale.copy())

    def save(self, path: str):
        np.savez(path, flat=self.flat, shift=self.shift, scale=self.scale)

    @classmethod
    def load(cls, path: str) -> "MLPParams":
        with np.load(path, allow_pickle=False) as z:
            return cls(z["flat"], z["shift"], z["scale"])


# ----------------------------------------------------------------- 1-D operators
def lagrange_matrix(nodes: np.ndarray, x: np.ndarray) -> np.ndarray:
    """L[m, i] = l_i(x_m) for the Lagrange basis on `nodes` (barycentric form)."""
    wb = baryweights(nodes)
    L = np.zeros((len(x), len(nodes)))
    for m, xm in enumerate(x):
        d = xm - nodes
        hit = np.nonzero(d == 0)[0]
        if hit.size:
            L[m, hit[0]] = 1.0
        else:
            t = wb / d
            L[m] = t / t.sum()
    return L


def filter_projection(N: int) -> np.ndarray:
    """P[h, i, a]: L2 projection (exact mass matrix) of the fine-element basis
    function a on half h of [-1, 1] (h = 0 lower) onto the coarse basis function i."""
    xi, _ = lglpoints(N)
    g, gw = np.polynomial.legendre.leggauss(2 * N + 4)
    Lg = lagrange_matrix(xi, g)
    M = (Lg * gw[:, None]).T @ Lg
    P = np.zeros((2, N + 1, N + 1))
    for h in range(2):
        xs = (g - 1.0) / 2.0 if h == 0 else (g + 1.0) / 2.0     # coarse coordinate on the half
        B = (lagrange_matrix(xi, xs) * (gw / 2.0)[:, None]).T @ Lg   # fine coordinate is g itself
        P[h] = np.linalg.solve(M, B)
    return P


def children_map(coarse: BrickGrid, fine: BrickGrid) -> np.ndarray:
    """children[8 c + hx + 2 hy + 4 hz]: fine element of half (hx, hy, hz) of coarse c."""
    nx, ny, nz = coarse.ne
    fx, fy, fz = fine.ne
    if (fx, fy, fz) != (2 * nx, 2 * ny, 2 * nz):
        raise ValueError("the fine brick must have twice the coarse elements per direction")
    out = np.empty((coarse.nelem, 8), np.int32)
    for h in range(8):
        hx, hy, hz = h & 1, (h >> 1) & 1, h >> 2
        kx, ky, kz = 2 * coarse.ex + hx, 2 * coarse.ey + hy, 2 * coarse.ez + hz
        out[:, h] = kz + fz * (kx + fx * ky)
    return np.ascontiguousarray(out.ravel())


def element_ijk(grid: BrickGrid) -> np.ndarray:
    return np.ascontiguousarray(np.stack([grid.ex, grid.ey, grid.ez], axis=1).astype(np.int32).ravel())


def interpolation_matrix(N: int, p: int) -> np.ndarray:
    """L[m, i] = l_i(xi_m), xi_m = -1 + (2m + 1)/p (cell centres of p per element)."""
    xi, _ = lglpoints(N)
    return lagrange_matrix(xi, -1.0 + (2.0 * np.arange(p) + 1.0) / p)


# ----------------------------------------------------------------- device drivers
def box_filter(Qf: torch.Tensor, fine: BrickGrid, coarse: BrickGrid, out: Optional[torch.Tensor] = None):
    """Filtered moments (Ec, 11, Np) of the fine state Qf (Ef, 5, Np) on the coarse brick."""
    lib = abi.load_library()
    ch = torch.from_numpy(children_map(coarse, fine)).to(Qf.device)
    P = np.ascontiguousarray(filter_projection(fine.N).ravel())
    if out is None:
        out = torch.empty((coarse.nelem, NMOM, coarse.Np), dtype=torch.float64, device=Qf.device)
    _check(lib.cmdg_box_filter(_vp(Qf), _vp(ch), coarse.nelem, fine.N, C.c_void_p(P.ctypes.data), _vp(out),
                               _stream(Qf.device)), "cmdg_box_filter")
    return out


def closure_samples(M: torch.Tensor, Gu: torch.Tensor, grid: BrickGrid, Delta: float, nu_mol: float,
                    table: torch.Tensor, offset: int, element_average: bool = False, nu_floor: float = 1e-5):
    """Append the grid's samples to the SoA table (13, ld) at column `offset`;
    nu_floor is the closure's viscosity floor (its context's `viscosity`, 1e-5 when
    zero); element_average replaces P, Q, R by their element means."""
    lib = abi.load_library()
    ld = table.shape[1]
    _check(lib.cmdg_closure_samples(_vp(M), _vp(Gu), grid.nelem, grid.N, float(Delta), float(nu_mol),
                                    float(nu_floor), int(bool(element_average)), _vp(table), ld, int(offset),
                                    _stream(M.device)), "cmdg_closure_samples")
    return offset + grid.nelem * grid.Np


def equispaced_velocity(Q: torch.Tensor, grid: BrickGrid, p: int) -> torch.Tensor:
    """u = (rho u)/rho on the uniform grid of p cell-centred points per element:
    (3, nz p, ny p, nx p) float64."""
    lib = abi.load_library()
    nx, ny, nz = grid.ne
    L = np.ascontiguousarray(interpolation_matrix(grid.N, p).ravel())
    eijk = torch.from_numpy(element_ijk(grid)).to(Q.device)
    out = torch.empty((3, nz * p, ny * p, nx * p), dtype=torch.float64, device=Q.device)
    _check(lib.cmdg_interpolate_equispaced(_vp(Q), _vp(eijk), grid.nelem, grid.N, p, C.c_void_p(L.ctypes.data),
                                           nx * p, ny * p, nz * p, _vp(out), _stream(Q.device)),
           "cmdg_interpolate_equispaced")
    return out


# ----------------------------------------------------------------- spectrum
def shell_index(dim: int) -> np.ndarray:
    """Shell (1-based) of every FFT mode of a dim^3 grid as power_spectrum_3d bins
    them (power_spectrum_les.jl:23-71): integer wavenumber vectors in FFT order,
    r = |k|; shell 1 takes r <= 1.5, shell N takes (N - 1/2, N + 1/2], the last
    shell (N = k_nyq) takes (k_nyq - 1/2, k_nyq]; 0 marks modes in no shell."""
    if dim % 2 == 0:
        freq = np.concatenate([np.arange(0, dim // 2 + 1), np.arange(-dim // 2 + 1, 0)])
        k_nyq = dim // 2
    else:
        freq = np.concatenate([np.arange(0, (dim - 1) // 2 + 1), np.arange(-(dim - 1) // 2, 0)])
        k_nyq = (dim - 1) // 2
    r = np.sqrt(freq[:, None, None] ** 2 + freq[None, :, None] ** 2 + freq[None, None, :] ** 2)
    shell = np.ceil(r - 0.5).astype(np.int64)
    shell[r <= 1.5] = 1
    shell[r > k_nyq] = 0
    return shell


def power_spectrum_3d(u, v, w, L: float, dim: int, nor: float = 1.0):
    """Shell-binned kinetic-energy spectrum of equispaced velocity fields
    (power_spectrum_3d(::AtmosLESConfigType, ...), power_spectrum_les.jl:13-75):
    spectrum[N] = sum_shell(|u^|^2 + |v^|^2 + |w^|^2) * 2 pi k_N^2 / (count_N dx^3) / nor
    with |.^| = |fft| / dim^3, dx = 2 pi / L, k_N = N dx.  Returns (spectrum, k)."""
    def mag2(f):
        t = torch.as_tensor(f)
        return (torch.fft.fftn(t).abs() / t.shape[0] ** 3).pow(2).cpu().numpy()
    e = mag2(u) + mag2(v) + mag2(w)
    shell = shell_index(dim)
    k_nyq = dim // 2 if dim % 2 == 0 else (dim - 1) // 2
    dx = 2 * np.pi / L
    sums = np.bincount(shell.ravel(), weights=e.ravel(), minlength=k_nyq + 1)[1:k_nyq + 1]
    cnt = np.bincount(shell.ravel(), minlength=k_nyq + 1)[1:k_nyq + 1]
    k = np.arange(1, k_nyq + 1) * dx
    return sums * 2 * np.pi * k ** 2 / (cnt * dx ** 3) / nor, k


def spectrum_error(E, E_ref, kmax: Optional[int] = None, floor: float = 1e-8) -> dict:
    """Errors of a spectrum against a reference over shells 1..kmax:
    rel_l2 = |E - E_ref|_2 / |E_ref|_2, and log_rms = RMS of log(E / E_ref) over the
    shells whose reference energy exceeds floor * max(E_ref) (shells holding only
    roundoff are left out)."""
    n = len(E_ref) if kmax is None else kmax
    E, R = np.asarray(E[:n], np.float64), np.asarray(E_ref[:n], np.float64)
    keep = R > floor * R.max()
    return dict(rel_l2=float(np.linalg.norm(E - R) / np.linalg.norm(R)),
                log_rms=float(np.sqrt(np.mean(np.log(np.maximum(E[keep], 1e-300) / R[keep]) ** 2))),
                shells=int(keep.sum()))


# ----------------------------------------------------------------- training
class ClosureTrainer:
    """Adam on the library's loss-gradient kernels; parameters stay on the device."""

    def __init__(self, mlp: MLPParams, device="cuda", max_batch: int = 1 << 16):
        self.lib = abi.load_library()
        self.device = torch.device(device)
        f32 = dict(dtype=torch.float32, device=self.device)
        self.params = torch.from_numpy(mlp.flat.copy()).to(**f32)
        self.shift = torch.from_numpy(mlp.shift.copy()).to(**f32)
        self.scale = torch.from_numpy(mlp.scale.copy()).to(**f32)
        self.m = torch.zeros(NPARAM, **f32)
        self.v = torch.zeros(NPARAM, **f32)
        self.grad = torch.zeros(NPARAM + 1, **f32)
        self.max_batch = max_batch
        self.work = torch.zeros(((max_batch + 63) // 64) * (NPARAM + 1), **f32)
        self.step = 0

    def loss_grad(self, table: torch.Tensor, batch: int, idx: Optional[torch.Tensor] = None,
                  inv_norm: float = 1.0):
        """Loss and gradient over `batch` samples (rows idx, or the first `batch` columns)."""
        if batch > self.max_batch:
            raise ClosureError("batch larger than the trainer's workspace")
        ip = C.c_void_p(0) if idx is None else _vp(idx)
        _check(self.lib.cmdg_mlp_loss_grad(_vp(self.params), _vp(self.shift), _vp(self.scale), _vp(table),
                                           table.shape[1], ip, int(batch), float(inv_norm), _vp(self.grad),
                                           _vp(self.work), self.work.numel(), _stream(self.device)),
               "cmdg_mlp_loss_grad")
        return self.grad

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        _check(self.lib.cmdg_mlp_adam(_vp(self.params), _vp(self.grad), _vp(self.m), _vp(self.v), NPARAM,
                                      float(lr), float(beta1), float(beta2), float(eps), self.step,
                                      _stream(self.device)), "cmdg_mlp_adam")

    def fit(self, table: torch.Tensor, nsamples: int, steps: int, batch: int, lr: float, inv_norm: float,
            seed: int = 0, lr_final: Optional[float] = None, log=None, log_every: int = 200):
        """Minibatch Adam with uniformly sampled rows; cosine decay to lr_final."""
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        hist = []
        lr_final = lr if lr_final is None else lr_final
        for it in range(steps):
            idx = torch.randint(0, nsamples, (batch,), device=self.device, generator=gen)
            self.loss_grad(table, batch, idx, inv_norm / batch)
            a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / max(1, steps - 1)))
            self.adam(a)
            if (it + 1) % log_every == 0 or it == 0:
                loss = float(self.grad[NPARAM].item())
                hist.append((it + 1, loss))
                if log:
                    log(f"[train] step {it + 1}/{steps} loss {loss:.5e} lr {a:.2e}")
        return hist

    def mlp(self) -> MLPParams:
        return MLPParams(self.params.cpu().numpy(), self.shift.cpu().numpy(), self.scale.cpu().numpy())

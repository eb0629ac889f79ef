"""cmdg_mlp_loss_grad_static, the loss-gradient entry point that allocates nothing
(its partial gradients live in the library's fixed device array), against
cmdg_mlp_loss_grad with a caller-allocated workspace: the same kernels, so the same
result bit for bit, on consecutive calls of decreasing size, and both against the
numpy oracle of the loss and its gradient."""
import ctypes as C

import numpy as np
import pytest

from cmdg import abi, closure
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _table(n, ld, seed):
    rng = np.random.default_rng(seed)
    table = np.zeros((13, ld), np.float32)
    table[:10, :n] = rng.standard_normal((10, n)) * 3
    P = rng.uniform(0.1, 2.0, n)
    Qv = -rng.uniform(0, 1, n) * P
    table[10, :n], table[11, :n], table[12, :n] = P, Qv, Qv * Qv / P + rng.uniform(0, 0.5, n)
    return table


def test_static_loss_gradient_equals_the_workspace_entry_point():
    import torch
    lib = abi.load_library()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(11)
    mlp = closure.MLPParams.random(seed=4, y0=0.02, shift=rng.normal(0, 0.1, 10), scale=rng.uniform(0.3, 1.0, 10))
    mlp.flat[closure.OFF_W3:closure.OFF_B3] *= 0.1   # outputs inside the clamp: non-zero gradients
    f32 = dict(dtype=torch.float32, device="cuda")
    params = torch.from_numpy(mlp.flat.astype(np.float32)).to(**f32)
    shift = torch.from_numpy(mlp.shift.astype(np.float32)).to(**f32)
    scale = torch.from_numpy(mlp.scale.astype(np.float32)).to(**f32)
    nmax = abi.MLP_STATIC_MAX_BATCH
    table = _table(nmax, nmax + 100, seed=2)
    T = torch.from_numpy(table).cuda()
    work = torch.empty((nmax // 64) * (abi.MLP_NPARAM + 1), **f32)
    # largest batch first, then smaller ones: the later calls find the earlier calls'
    # partial rows in the shared array and must not see them
    idx_small = torch.from_numpy(rng.permutation(nmax)[:3001].astype(np.int64)).cuda()
    cases = ((nmax, None), (20000, None), (3001, idx_small), (37, None))
    for batch, idx in cases:
        print(f"[no-alloc] loss gradient, batch {batch}, {'rows' if idx is not None else 'first columns'}",
              flush=True)
        ip = C.c_void_p(0) if idx is None else _vp(idx)
        g_ws = torch.full((abi.MLP_NPARAM + 1,), float("nan"), **f32)
        g_st = torch.full((abi.MLP_NPARAM + 1,), float("nan"), **f32)
        rc = lib.cmdg_mlp_loss_grad(_vp(params), _vp(shift), _vp(scale), _vp(T), T.shape[1], ip, batch,
                                    1.0 / batch, _vp(g_ws), _vp(work), work.numel(), stream)
        assert rc == abi.CMDG_OK
        rc = lib.cmdg_mlp_loss_grad_static(_vp(params), _vp(shift), _vp(scale), _vp(T), T.shape[1], ip, batch,
                                           1.0 / batch, _vp(g_st), stream)
        assert rc == abi.CMDG_OK
        torch.cuda.synchronize()
        a, b = g_ws.cpu().numpy(), g_st.cpu().numpy()
        assert np.isfinite(b).all() and np.abs(b[:abi.MLP_NPARAM]).max() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "static entry point differs from the workspace one"
        rows = np.arange(batch) if idx is None else idx.cpu().numpy()
        x = MO.normalise(table, mlp.shift, mlp.scale, rows)
        L, gref = MO.loss_grad(mlp.flat, x, table[10, rows], table[11, rows], table[12, rows], 1.0 / batch)
        gerr = np.linalg.norm(b[:abi.MLP_NPARAM].astype(np.float64) - gref) / np.linalg.norm(gref)
        print(f"  loss {b[abi.MLP_NPARAM]:.6e} vs oracle {L:.6e}, gradient rel L2 {gerr:.2e}", flush=True)
        assert abs(b[abi.MLP_NPARAM] - L) < 1e-5 * abs(L)
        assert gerr < 1e-4


def test_static_loss_gradient_rejects_a_batch_above_its_maximum():
    import torch
    lib = abi.load_library()
    one = torch.zeros(16, dtype=torch.float32, device="cuda")
    p = _vp(one)
    rc = lib.cmdg_mlp_loss_grad_static(p, p, p, p, abi.MLP_STATIC_MAX_BATCH + 1, C.c_void_p(0),
                                       abi.MLP_STATIC_MAX_BATCH + 1, 1.0, p, C.c_void_p(0))
    assert rc == abi.CMDG_ERANGE

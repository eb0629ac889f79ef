"""Checkpoint/restart of the state: the contract of src/Driver/Checkpoint/Checkpoint.jl:37-125.

write_checkpoint stores what the reference's `@save cfull h_Q h_aux t` stores: the
state `h_Q` and auxiliary state `h_aux` as FP64 arrays of shape (Np, nstate, nelem)
and (Np, naux, nelem), and the solver time `t`, in
"<name>_checkpoint_mpirank%04d_num%04d.jld2" (spaces in the name become "_",
Checkpoint.jl:47-53).  The file is HDF5, as JLD2 files are: each array is a
contiguous IEEE float64 little-endian dataset at the root named as the Julia
variable, with the dimensions listed in reverse (Julia arrays are column-major, HDF5
dataspaces row-major, so the bytes are the Julia array's own), `t` is a scalar
dataset, and the first 512 bytes (the HDF5 user block) hold JLD2's file-header
string.  The file format objects are those of HDF5 1.8 (superblock version 2,
version-2 object headers), the layout JLD2 writes, rather than the library's
earliest-compatible default (superblock version 0).  read_checkpoint returns (h_Q, h_aux, t) as Fortran-ordered NumPy arrays
(Checkpoint.jl:103-125), ready for DGModel(cfg, grid, aux=h_aux) and to_device(h_Q).

The HDF5 library itself is the system libhdf5 (1.10, C API through ctypes):
CMDG_HDF5_LIBRARY names it, else the first of ctypes.util.find_library("hdf5") and
/opt/conda/lib/libhdf5.so that loads.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

# JLD2's file header (the first bytes of every .jld2 file; the rest of the 512-byte
# user block is zero).
JLD2_HEADER = b"HDF5-based Julia Data Format, version 0.1.1\x00 (Julia 1.5.4 64-bit LE)\x00"
USERBLOCK = 512

_lib = None
_hid = C.c_int64          # hid_t (HDF5 >= 1.10)
_hsize = C.c_ulonglong    # hsize_t
H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5F_LIBVER_V18 = 1        # H5F_libver_t: file format objects of HDF5 1.8
H5P_DEFAULT, H5S_ALL, H5S_SCALAR = 0, 0, 0


def _load():
    global _lib
    if _lib is not None:
        return _lib
    cands = [os.environ.get("CMDG_HDF5_LIBRARY"), ctypes.util.find_library("hdf5"), "/opt/conda/lib/libhdf5.so"]
    err = []
    for c in cands:
        if not c:
            continue
        try:
            lib = C.CDLL(c)
            break
        except OSError as ex:
            err.append(f"{c}: {ex}")
    else:
        raise OSError("no HDF5 library found (set CMDG_HDF5_LIBRARY): " + "; ".join(err))
    if lib.H5open() < 0:
        raise OSError("H5open failed")
    for name, res, args in [
        ("H5Pcreate", _hid, [_hid]), ("H5Pset_userblock", C.c_int, [_hid, _hsize]),
        ("H5Pset_libver_bounds", C.c_int, [_hid, C.c_int, C.c_int]),
        ("H5Pclose", C.c_int, [_hid]), ("H5Fcreate", _hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        ("H5Fopen", _hid, [C.c_char_p, C.c_uint, _hid]), ("H5Fclose", C.c_int, [_hid]),
        ("H5Screate_simple", _hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
        ("H5Screate", _hid, [C.c_int]), ("H5Sclose", C.c_int, [_hid]),
        ("H5Sget_simple_extent_ndims", C.c_int, [_hid]),
        ("H5Sget_simple_extent_dims", C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
        ("H5Dcreate2", _hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        ("H5Dopen2", _hid, [_hid, C.c_char_p, _hid]), ("H5Dget_space", _hid, [_hid]),
        ("H5Dwrite", C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        ("H5Dread", C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]), ("H5Dclose", C.c_int, [_hid]),
    ]:
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _lib = lib
    return lib


def _global(name: str) -> int:
    """A predefined HDF5 identifier (H5T_IEEE_F64LE and the like are globals set by H5open)."""
    return _hid.in_dll(_load(), name + "_g").value


def _check(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


def checkpoint_filename(name: str, rank: int, num: int) -> str:
    """Checkpoint.jl:47-53."""
    return "%s_checkpoint_mpirank%04d_num%04d.jld2" % (name.replace(" ", "_"), rank, num)


def _host(a) -> np.ndarray:
    """A NumPy (Np, nvar, nelem) array, or a torch tensor in the library's
    (nelem, nvar, Np) view of the same memory (cmdg.dg.to_device), as a Fortran array."""
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy().T
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 3:
        raise ValueError(f"expected an (Np, nvar, nelem) array, got shape {a.shape}")
    return np.asfortranarray(a)


def write_checkpoint(Q, aux, t: float, checkpoint_dir: str, name: str, num: int, rank: int = 0) -> str:
    """write_checkpoint (Checkpoint.jl:37-72); returns the file written.  Q and aux are
    NumPy (Np, nvar, nelem) arrays or the library's device tensors."""
    lib = _load()
    path = os.path.join(checkpoint_dir, checkpoint_filename(name, rank, num))
    arrays = {"h_Q": _host(Q), "h_aux": _host(aux)}
    fcpl = _check(lib.H5Pcreate(_global("H5P_CLS_FILE_CREATE_ID")), "H5Pcreate")
    fapl = _check(lib.H5Pcreate(_global("H5P_CLS_FILE_ACCESS_ID")), "H5Pcreate")
    try:
        _check(lib.H5Pset_userblock(fcpl, USERBLOCK), "H5Pset_userblock")
        _check(lib.H5Pset_libver_bounds(fapl, H5F_LIBVER_V18, H5F_LIBVER_V18), "H5Pset_libver_bounds")
        fid = _check(lib.H5Fcreate(path.encode(), H5F_ACC_TRUNC, fcpl, fapl), "H5Fcreate " + path)
    finally:
        lib.H5Pclose(fcpl)
        lib.H5Pclose(fapl)
    f64le, native = _global("H5T_IEEE_F64LE"), _global("H5T_NATIVE_DOUBLE")
    try:
        for key, a in arrays.items():
            dims = (_hsize * 3)(*reversed(a.shape))
            sid = _check(lib.H5Screate_simple(3, dims, None), "H5Screate_simple")
            try:
                did = _check(lib.H5Dcreate2(fid, key.encode(), f64le, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                             "H5Dcreate2 " + key)
                try:
                    # the Fortran-ordered buffer is the row-major array of the reversed dims
                    _check(lib.H5Dwrite(did, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "H5Dwrite " + key)
                finally:
                    lib.H5Dclose(did)
            finally:
                lib.H5Sclose(sid)
        sid = _check(lib.H5Screate(H5S_SCALAR), "H5Screate")
        try:
            did = _check(lib.H5Dcreate2(fid, b"t", f64le, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Dcreate2 t")
            try:
                tv = C.c_double(float(t))
                _check(lib.H5Dwrite(did, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.addressof(tv)), "H5Dwrite t")
            finally:
                lib.H5Dclose(did)
        finally:
            lib.H5Sclose(sid)
    finally:
        _check(lib.H5Fclose(fid), "H5Fclose")
    with open(path, "r+b") as fh:  # the user block is the file's own: JLD2's header goes there
        fh.write(JLD2_HEADER.ljust(USERBLOCK, b"\x00"))
    return path


def read_checkpoint(checkpoint_dir: str, name: str, num: int, rank: int = 0):
    """read_checkpoint (Checkpoint.jl:103-125): (h_Q, h_aux, t)."""
    lib = _load()
    path = os.path.join(checkpoint_dir, checkpoint_filename(name, rank, num))
    if not os.path.isfile(path):
        raise FileNotFoundError(f"Cannot restore from checkpoint in {path}, file not found")
    native = _global("H5T_NATIVE_DOUBLE")
    fid = _check(lib.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen " + path)
    out = {}
    try:
        for key in ("h_Q", "h_aux", "t"):
            did = _check(lib.H5Dopen2(fid, key.encode(), H5P_DEFAULT), "H5Dopen2 " + key)
            try:
                sid = _check(lib.H5Dget_space(did), "H5Dget_space")
                try:
                    nd = _check(lib.H5Sget_simple_extent_ndims(sid), "ndims")
                    dims = (_hsize * max(nd, 1))()
                    if nd:
                        _check(lib.H5Sget_simple_extent_dims(sid, dims, None), "dims")
                finally:
                    lib.H5Sclose(sid)
                shape = tuple(reversed([int(dims[i]) for i in range(nd)]))
                a = np.empty(shape, dtype=np.float64, order="F")
                _check(lib.H5Dread(did, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "H5Dread " + key)
                out[key] = a
            finally:
                lib.H5Dclose(did)
    finally:
        lib.H5Fclose(fid)
    return out["h_Q"], out["h_aux"], float(out["t"])


def rm_checkpoint(checkpoint_dir: str, name: str, num: int, rank: int = 0):
    """rm_checkpoint (Checkpoint.jl:79-96)."""
    p = os.path.join(checkpoint_dir, checkpoint_filename(name, rank, num))
    if os.path.exists(p):
        os.remove(p)

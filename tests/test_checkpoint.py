"""Checkpoint/restart files (src/Driver/Checkpoint/Checkpoint.jl:37-125): names, the
HDF5 layout JLD2 uses for Julia arrays, and round trips.  h5dump (the HDF5 tools) is an
independent reader of what cmdg.checkpoint writes."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cmdg import checkpoint as ck

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def test_filename_follows_the_reference():
    assert ck.checkpoint_filename("dry rising bubble", 0, 7) == "dry_rising_bubble_checkpoint_mpirank0000_num0007.jld2"
    assert ck.checkpoint_filename("x", 12, 345) == "x_checkpoint_mpirank0012_num0345.jld2"


def test_round_trip_is_bitwise(tmp_path):
    rng = np.random.default_rng(4)
    Q = np.asfortranarray(rng.standard_normal((125, 5, 9)))
    A = np.asfortranarray(rng.standard_normal((125, 17, 9)))
    Q[0, 0, 0], Q[1, 1, 1] = -0.0, np.nextafter(1.0, 2.0)
    p = ck.write_checkpoint(Q, A, 3.25, str(tmp_path), "case a", 2)
    assert os.path.basename(p) == "case_a_checkpoint_mpirank0000_num0002.jld2"
    q2, a2, t2 = ck.read_checkpoint(str(tmp_path), "case a", 2)
    assert q2.shape == Q.shape and a2.shape == A.shape and q2.flags["F_CONTIGUOUS"]
    assert q2.tobytes(order="F") == Q.tobytes(order="F") and a2.tobytes(order="F") == A.tobytes(order="F")
    assert t2 == 3.25
    with open(p, "rb") as f:
        head = f.read(ck.USERBLOCK)
    assert head.startswith(b"HDF5-based Julia Data Format, version ")
    ck.rm_checkpoint(str(tmp_path), "case a", 2)
    assert not os.path.exists(p)
    with pytest.raises(FileNotFoundError):
        ck.read_checkpoint(str(tmp_path), "case a", 2)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_independent_reader_sees_julia_layout(tmp_path):
    """h5dump lists a dataset in row-major order of its (reversed) dimensions, which is
    the column-major order of the Julia array: values 0..23 of a (2, 3, 4) array."""
    Q = np.asfortranarray(np.arange(24, dtype=np.float64).reshape((2, 3, 4), order="F"))
    A = np.asfortranarray(np.full((2, 1, 4), 7.5))
    p = ck.write_checkpoint(Q, A, -1.5, str(tmp_path), "tiny", 0)
    head = subprocess.run([H5DUMP, "-H", p], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "h_Q"' in head and 'DATASET "h_aux"' in head and 'DATASET "t"' in head
    assert "H5T_IEEE_F64LE" in head and "SIMPLE { ( 4, 3, 2 ) / ( 4, 3, 2 ) }" in head
    assert "SIMPLE { ( 4, 1, 2 ) / ( 4, 1, 2 ) }" in head and "SCALAR" in head
    out = subprocess.run([H5DUMP, "-d", "/h_Q", "-y", "-w", "0", p], capture_output=True, text=True, check=True).stdout
    body = out[out.index("DATA {") + 6:out.rindex("}")]
    vals = [float(v) for v in body.replace("}", " ").replace(",", " ").split()]
    assert vals == list(range(24))
    tout = subprocess.run([H5DUMP, "-d", "/t", p], capture_output=True, text=True, check=True).stdout
    assert "-1.5" in tout
    blk = subprocess.run([H5DUMP, "-B", "-H", p], capture_output=True, text=True, check=True).stdout
    assert "USER_BLOCK" in blk and "512" in blk


def test_tensor_input_uses_the_device_view(tmp_path):
    """A tensor in the library's (nelem, nvar, Np) view (cmdg.dg.to_device on the CPU
    here) is written as the (Np, nvar, nelem) array it views."""
    import torch
    Q = np.asfortranarray(np.random.default_rng(2).standard_normal((8, 5, 3)))
    A = np.asfortranarray(np.random.default_rng(3).standard_normal((8, 4, 3)))
    tQ = torch.from_numpy(np.ascontiguousarray(Q.T))
    tA = torch.from_numpy(np.ascontiguousarray(A.T))
    ck.write_checkpoint(tQ, tA, 1.0, str(tmp_path), "t", 1)
    q2, a2, _ = ck.read_checkpoint(str(tmp_path), "t", 1)
    assert np.array_equal(q2, Q) and np.array_equal(a2, A)


def test_superblock_version_2_after_the_user_block(tmp_path):
    """JLD2 writes HDF5 superblock version 2 (file format objects of HDF5 1.8): the
    signature sits right after the 512-byte user block, followed by the version byte."""
    Q = np.asfortranarray(np.ones((125, 5, 2)))
    p = ck.write_checkpoint(Q, Q, 0.0, str(tmp_path), "sb", 0)
    with open(p, "rb") as f:
        raw = f.read(ck.USERBLOCK + 9)
    assert raw[ck.USERBLOCK:ck.USERBLOCK + 8] == b"\x89HDF\r\n\x1a\n"
    assert raw[ck.USERBLOCK + 8] == 2

"""cmdg: MI355X DG right-hand side + LSRK54 for the dry AtmosModel.

The compute path is the HIP library ``lib/libcmdg_hip.so`` (C ABI in
include/cmdg.h); this package is the host side that mirrors the reference's
``DGModel`` / ``LSRK54CarpenterKennedy`` / ``solve!`` interface and builds the
grid and auxiliary state once on the host.
"""
from . import abi, atmos, grid  # noqa: F401

__all__ = ["abi", "atmos", "grid", "dg"]


def __getattr__(name):
    # torch is imported lazily so grid/atmos stay usable without it
    if name == "dg":
        from . import dg
        return dg
    raise AttributeError(name)

"""Writes the golden vectors of the parity tests: the CPU oracle's right-hand side and
one LSRK54CarpenterKennedy step on a small seeded LES case, stored with the input state.

    python tests/golden/make_oracle_vectors.py

Case: a 2 x 2 x 2 brick of 100 m elements, N = 4, periodic in x and y, FreeSlip /
Insulating walls in z, gravity with the hydrostatic reference of a dry adiabatic
profile (theta = 300 K), Smagorinsky-Lilly (C_s = 0.23), Rusanov; the state is a cos^2
warm bubble (+2 K, radius 80 m) with 1e-3 K of seeded noise.  The oracle restates the
reference's data flow (oracle/cmdg_oracle.cpp); the vectors pin it against drift and
give the HIP path a committed fixture to meet.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "climatemachine.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

FILE = os.path.join(HERE, "oracle_les_bubble_2x2x2.npz")
DT = 0.05


def case():
    """(cfg, grid, aux) of the golden case; deterministic."""
    from cmdg import abi, atmos
    from cmdg.grid import uniform_brick
    prm = atmos.ideal_gas_parameters()
    grid = uniform_brick((2, 2, 2), (0.0, 0.0, 0.0), (200.0, 200.0, 200.0), N=4,
                         periodicity=(True, True, False), boundary=((0, 0), (0, 0), (1, 2)))
    cfg = atmos.AtmosConfig(prm, orientation=abi.FLAT_ORIENTATION,
                            ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0), gravity_source=True,
                            turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.23, numerical_flux=abi.RUSANOV)
    aux = atmos.init_aux(cfg, grid)
    return cfg, grid, aux


def initial_state(grid):
    from cmdg import atmos
    prm = atmos.ideal_gas_parameters()
    return atmos.rising_bubble(prm, grid, theta_ref=300.0, dtheta=2.0, radius=80.0,
                               centre=(100.0, 100.0, 60.0), noise=1e-3, seed=3)


def main():
    from oracle import oracle as O
    cfg, grid, aux = case()
    Q0 = np.asfortranarray(initial_state(grid))
    orc = O.Oracle(cfg, grid, aux)
    tend = orc.tendency(Q0)
    Q1 = np.asfortranarray(Q0.copy())
    dQ = np.zeros_like(Q1, order="F")
    orc.lsrk54_step(Q1, dQ, 0.0, DT)
    np.savez_compressed(FILE, Q0=Q0, tendency=tend, Q1=Q1, dt=np.float64(DT))
    print(f"wrote {FILE}: {os.path.getsize(FILE)} bytes")


if __name__ == "__main__":
    main()

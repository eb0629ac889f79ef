"""The 3-D manufactured solution of the reference's compressible Navier-Stokes test,
restated from its definition (test/Numerics/DGMethods/compressible_Navier_Stokes/
mms_solution.jl:14-102) and differentiated here with sympy: the exact state and the
source that makes it a solution of the compressible Navier-Stokes equations with
constant dynamic viscosity mu = 1/100 and the Stokes stress tau = 2 mu (eps - tr(eps)/3 I).
Test infrastructure: it drives the oracle's reproduction of the reference's known
answer (mms_bc_atmos.jl:201-214)."""
import functools

import numpy as np
import sympy as sp

x, y, z, t = sp.symbols("x y z t", real=True)
MU = sp.Rational(1, 100)
GAMMA = sp.Rational(7, 5)   # cp_d / cv_d with kappa_d = 2/7


@functools.lru_cache(maxsize=None)
def _expressions():
    pi = sp.pi
    rho = sp.cos(pi * t) * sp.sin(pi * x) * sp.cos(pi * y) * sp.cos(pi * z) + 3
    U = sp.cos(pi * t) * rho * sp.sin(pi * x) * sp.cos(pi * y) * sp.cos(pi * z)
    V = U
    W = sp.cos(pi * t) * rho * sp.sin(pi * x) * sp.cos(pi * y) * sp.sin(pi * z)
    E = sp.cos(pi * t) * sp.sin(pi * x) * sp.cos(pi * y) * sp.cos(pi * z) + 100
    P = (GAMMA - 1) * (E - (U ** 2 + V ** 2 + W ** 2) / (2 * rho))
    vel = [U / rho, V / rho, W / rho]
    X = (x, y, z)
    grad = [[sp.diff(vel[i], X[j]) for j in range(3)] for i in range(3)]
    eps = [[(grad[i][j] + grad[j][i]) / 2 for j in range(3)] for i in range(3)]
    tr = eps[0][0] + eps[1][1] + eps[2][2]
    tau = [[2 * MU * (eps[i][j] - (tr / 3 if i == j else 0)) for j in range(3)] for i in range(3)]
    mom = [U, V, W]
    # flux of (rho, rho u, rho e) through direction j
    def flux(j):
        return [mom[j]] + [vel[j] * mom[i] + (P if i == j else 0) - tau[j][i] for i in range(3)] + \
               [vel[j] * (E + P) - sum(vel[i] * tau[j][i] for i in range(3))]
    F = [flux(j) for j in range(3)]
    state = [rho, U, V, W, E]
    source = [sp.diff(state[s], t) + sum(sp.diff(F[j][s], X[j]) for j in range(3)) for s in range(5)]
    return state, source


@functools.lru_cache(maxsize=None)
def _functions():
    state, source = _expressions()
    args = (t, x, y, z)
    return sp.lambdify(args, state, "numpy", cse=True), sp.lambdify(args, source, "numpy", cse=True)


def _evaluate(f, tt, X):
    vals = f(float(tt), X[0], X[1], X[2])
    return np.stack([np.broadcast_to(np.asarray(v, dtype=np.float64), X[0].shape) for v in vals], axis=1)


def exact_state(tt, X):
    """(Np, 5, E) state (rho, rho u, rho e) at time tt at node coordinates X (3, Np, E)."""
    return np.asfortranarray(_evaluate(_functions()[0], tt, X))


def source(tt, X):
    """(Np, 5, E) manufactured source at time tt."""
    return np.asfortranarray(_evaluate(_functions()[1], tt, X))

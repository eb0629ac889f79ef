"""Brick topology + spectral-element grid in the reference's storage layout.

Builds, once and on the host, the arrays a ``DiscontinuousSpectralElementGrid``
holds (src/Numerics/Mesh/Grids.jl:267-413) for an axis-aligned brick:

* ``vgeo``  (Np, 25, nelem)  metric terms, mass matrix, coordinates
  (field ids Grids.jl:76-92 and GeometricFactors.jl:18-67),
* ``sgeo``  (5, Nfp, 6, nelem)  outward normal, surface mass, volume mass
  inverse at face nodes (Grids.jl:129-130, 1077-1105),
* ``vmap_minus`` / ``vmap_plus`` (Nfp, 6, nelem) 1-based volume dof indices
  (``mappings``, Grids.jl:559-637), ``elemtobndy`` (6, nelem) boundary tags.

All arrays are Fortran-ordered so their memory matches Julia's column-major
arrays.  Elements are stacked: the vertical index runs fastest, as in
``StackedBrickTopology`` (src/Numerics/Mesh/Topologies.jl:632-).  Metric terms
of an axis-aligned brick element are the exact affine values (d xi / d x =
2/h), the limit of the curl-invariant construction of
src/Numerics/Mesh/Metrics.jl:431-722 for straight-sided hexahedra.
"""
from __future__ import annotations

import numpy as np

# vgeo field ids, 0-based (Grids.jl:76-92; GeometricFactors.jl:18-67)
XI1X1, XI2X1, XI3X1, XI1X2, XI2X2, XI3X2, XI1X3, XI2X3, XI3X3 = range(9)
M_ID, MI_ID, MH_ID, X1_ID, X2_ID, X3_ID, JCV_ID = 9, 10, 11, 12, 13, 14, 15
NVGEO = 25
# sgeo field ids (Grids.jl:129-130)
N1_ID, N2_ID, N3_ID, SM_ID, VMI_ID = range(5)


def legendre(N: int, x: np.ndarray):
    """P_N(x) and P_{N-1}(x) by the three-term recurrence."""
    p0 = np.ones_like(x)
    if N == 0:
        return p0, np.zeros_like(x)
    p1 = x.copy()
    for n in range(2, N + 1):
        p0, p1 = p1, ((2 * n - 1) * x * p1 - (n - 1) * p0) / n
    return p1, p0


def lglpoints(N: int):
    """Legendre-Gauss-Lobatto nodes and weights (Elements.jl:11-14).

    Nodes are +-1 and the roots of P_N'; Newton iteration on
    (1 - x^2) P_N'(x) = N (P_{N-1} - x P_N), weights 2 / (N (N+1) P_N(x)^2).
    """
    if N < 1:
        raise ValueError("LGL needs N >= 1")
    x = -np.cos(np.pi * np.arange(N + 1) / N)
    for _ in range(100):
        pN, pNm1 = legendre(N, x)
        xold = x.copy()
        # Newton step for f(x) = x P_N - P_{N-1} (zero at the LGL interior nodes
        # and at +-1), f'(x) = (N+1) P_N
        x = x - (x * pN - pNm1) / ((N + 1) * pN)
        if np.max(np.abs(x - xold)) < 1e-16:
            break
    x[0], x[-1] = -1.0, 1.0
    x[N // 2] = 0.0 if N % 2 == 0 else x[N // 2]
    # symmetrise
    x = (x - x[::-1]) / 2
    pN, _ = legendre(N, x)
    w = 2.0 / (N * (N + 1) * pN ** 2)
    return x, w


def baryweights(r: np.ndarray) -> np.ndarray:
    """Barycentric weights (Elements.jl:33-46)."""
    n = len(r)
    wb = np.ones(n)
    for j in range(n):
        for i in range(n):
            if i != j:
                wb[j] = wb[j] * (r[j] - r[i])
        wb[j] = 1.0 / wb[j]
    return wb


def spectralderivative(r: np.ndarray) -> np.ndarray:
    """D[j, k] = l_k'(r_j) (Elements.jl:60-82)."""
    n = len(r)
    wb = baryweights(r)
    D = np.zeros((n, n))
    for k in range(n):
        for j in range(n):
            if k == j:
                for l in range(n):
                    if l != k:
                        D[j, k] = D[j, k] + 1.0 / (r[k] - r[l])
            else:
                D[j, k] = (wb[k] / wb[j]) / (r[j] - r[k])
    return D


def face_vids(Nq: int) -> np.ndarray:
    """fmask of Grids.jl:581-588: (6, Nfp) 0-based volume node per face node."""
    N = Nq - 1
    a = np.arange(Nq * Nq) % Nq
    b = np.arange(Nq * Nq) // Nq
    return np.stack([
        0 + Nq * (a + Nq * b),
        N + Nq * (a + Nq * b),
        a + Nq * (0 + Nq * b),
        a + Nq * (N + Nq * b),
        a + Nq * (b + Nq * 0),
        a + Nq * (b + Nq * N),
    ]).astype(np.int64)


class BrickGrid:
    """Axis-aligned brick DG grid.

    ``brickrange``: three 1-D arrays of element break points (like the
    reference's ``brickrange`` tuple); ``periodicity``: three bools;
    ``boundary``: ((xmin_tag, xmax_tag), (ymin, ymax), (zmin, zmax)) tags used on
    non-periodic faces (AtmosLESConfiguration default ((0,0),(0,0),(1,2))).
    """

    def __init__(self, brickrange, N: int = 4, periodicity=(True, True, True),
                 boundary=((1, 1), (1, 1), (1, 2))):
        self.N = N
        self.Nq = Nq = N + 1
        self.Np = Np = Nq ** 3
        self.Nfp = Nfp = Nq * Nq
        self.brickrange = [np.asarray(b, dtype=np.float64) for b in brickrange]
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.boundary = boundary
        self.ne = tuple(len(b) - 1 for b in self.brickrange)
        nx, ny, nz = self.ne
        self.nelem = E = nx * ny * nz
        self.xi, self.omega = lglpoints(N)
        self.D = spectralderivative(self.xi)
        self.fmask = face_vids(Nq)

        # element (ix, iy, iz) -> e with the vertical index fastest
        ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        eid = lambda i, j, k: k + nz * (i + nx * j)
        self.eidx = (ix, iy, iz)
        e_of = np.empty(E, dtype=np.int64)
        ex = np.empty(E, dtype=np.int64)
        ey = np.empty(E, dtype=np.int64)
        ez = np.empty(E, dtype=np.int64)
        e_all = eid(ix, iy, iz).ravel()
        ex[e_all] = ix.ravel()
        ey[e_all] = iy.ravel()
        ez[e_all] = iz.ravel()
        self.ex, self.ey, self.ez = ex, ey, ez
        del e_of

        lo = [self.brickrange[d][:-1] for d in range(3)]
        hi = [self.brickrange[d][1:] for d in range(3)]
        h1 = (hi[0] - lo[0])[ex]
        h2 = (hi[1] - lo[1])[ey]
        h3 = (hi[2] - lo[2])[ez]
        x0 = lo[0][ex]
        y0 = lo[1][ey]
        z0 = lo[2][ez]

        w = self.omega
        r = self.xi
        # node weights in kron order of Grids.jl:1068-1071: (w_k * w_j) * w_i
        W = np.empty(Np)
        ii = np.arange(Np) % Nq
        jj = (np.arange(Np) // Nq) % Nq
        kk = np.arange(Np) // (Nq * Nq)
        W[:] = (w[kk] * w[jj]) * w[ii]
        self.node_ijk = (ii, jj, kk)

        vgeo = np.zeros((Np, NVGEO, E), order="F")
        J = (h1 / 2) * (h2 / 2) * (h3 / 2)
        vgeo[:, XI1X1, :] = (2.0 / h1)[None, :]
        vgeo[:, XI2X2, :] = (2.0 / h2)[None, :]
        vgeo[:, XI3X3, :] = (2.0 / h3)[None, :]
        M = J[None, :] * W[:, None]
        vgeo[:, M_ID, :] = M
        vgeo[:, MI_ID, :] = 1.0 / M
        vgeo[:, MH_ID, :] = ((h1 / 2) * (h2 / 2))[None, :] * (w[jj] * w[ii])[:, None]
        vgeo[:, X1_ID, :] = x0[None, :] + ((r[ii] + 1) / 2)[:, None] * h1[None, :]
        vgeo[:, X2_ID, :] = y0[None, :] + ((r[jj] + 1) / 2)[:, None] * h2[None, :]
        vgeo[:, X3_ID, :] = z0[None, :] + ((r[kk] + 1) / 2)[:, None] * h3[None, :]
        vgeo[:, JCV_ID, :] = (h3 / 2)[None, :]
        vgeo[:, 16, :] = (h1 / 2)[None, :]   # x1 xi1
        vgeo[:, 20, :] = (h2 / 2)[None, :]   # x2 xi2
        vgeo[:, 24, :] = (h3 / 2)[None, :]   # x3 xi3
        self.vgeo = vgeo

        # surface geometry
        a = np.arange(Nfp) % Nq
        b = np.arange(Nfp) // Nq
        sgeo = np.zeros((5, Nfp, 6, E), order="F")
        sJ = [(h2 / 2) * (h3 / 2), (h1 / 2) * (h3 / 2), (h1 / 2) * (h2 / 2)]
        # face quadrature weights (Grids.jl:1078-1088): x-faces w_k*w_j at
        # fn = j + Nq k, y-faces w_k*w_i at fn = i + Nq k, z-faces w_j*w_i at fn = i + Nq j
        sw = [w[b] * w[a], w[b] * w[a], w[b] * w[a]]
        for f in range(6):
            d = f // 2
            sgn = -1.0 if f % 2 == 0 else 1.0
            sgeo[N1_ID + d, :, f, :] = sgn
            sgeo[SM_ID, :, f, :] = sJ[d][None, :] * sw[d][:, None]
            sgeo[VMI_ID, :, f, :] = vgeo[self.fmask[f], MI_ID, :]
        self.sgeo = sgeo

        # connectivity
        vmm = np.empty((Nfp, 6, E), dtype=np.int64, order="F")
        vmp = np.empty((Nfp, 6, E), dtype=np.int64, order="F")
        bndy = np.zeros((6, E), dtype=np.int64, order="F")
        eall = np.arange(E)
        idx = [ex, ey, ez]
        for f in range(6):
            d = f // 2
            step = -1 if f % 2 == 0 else 1
            nb = [ex.copy(), ey.copy(), ez.copy()]
            nb[d] = idx[d] + step
            n_d = self.ne[d]
            outside = (nb[d] < 0) | (nb[d] >= n_d)
            if self.periodicity[d]:
                nb[d] = np.mod(nb[d], n_d)
                e2 = eid(nb[0], nb[1], nb[2])
                bnd = np.zeros(E, dtype=bool)
            else:
                nb[d] = np.clip(nb[d], 0, n_d - 1)
                e2 = np.where(outside, eall, eid(nb[0], nb[1], nb[2]))
                bnd = outside
            f2 = f ^ 1
            vmm[:, f, :] = Np * eall[None, :] + self.fmask[f][:, None] + 1
            vmp[:, f, :] = np.where(bnd[None, :], vmm[:, f, :],
                                    Np * e2[None, :] + self.fmask[f2][:, None] + 1)
            tag = self.boundary[d][f % 2]
            bndy[f, :] = np.where(bnd, tag, 0)
        self.vmap_minus = vmm
        self.vmap_plus = vmp
        self.elemtobndy = bndy

    # ------------------------------------------------------------------ views
    @property
    def coords(self):
        """(Np, nelem) arrays x1, x2, x3."""
        return (self.vgeo[:, X1_ID, :], self.vgeo[:, X2_ID, :], self.vgeo[:, X3_ID, :])

    @property
    def mass(self):
        """(Np, nelem) mass matrix: the MPIStateArray weights (create_states.jl:14-16)."""
        return self.vgeo[:, M_ID, :]

    @property
    def ndof_nodes(self) -> int:
        return self.Np * self.nelem

    def min_node_distance(self) -> float:
        """Smallest LGL node gap over all elements and directions."""
        gaps = np.diff(self.xi).min() / 2
        return float(min(np.min(np.diff(b)) for b in self.brickrange) * gaps)


def uniform_brick(n, lo, hi, N=4, periodicity=(True, True, True),
                  boundary=((1, 1), (1, 1), (1, 2))) -> BrickGrid:
    """Brick of n = (nx, ny, nz) equal elements on [lo, hi]."""
    br = [np.linspace(lo[d], hi[d], n[d] + 1) for d in range(3)]
    return BrickGrid(br, N=N, periodicity=periodicity, boundary=boundary)

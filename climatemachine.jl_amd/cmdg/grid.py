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
src/Numerics/Mesh/Metrics.jl:431-722 for straight-sided hexahedra.  With a
``meshwarp`` (the reference grid's keyword, Grids.jl:267-325, 1028-1124) the
node coordinates are warped and every metric term comes from the curl-invariant
construction itself (``curvilinear_geometry``).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def over_elements(nelem: int, fn, min_chunk: int = 4096) -> None:
    """Call fn(e0, e1) on consecutive element ranges covering [0, nelem), on a thread
    pool when the mesh is large.  numpy's element-wise operations release the GIL, so
    host arrays of a large mesh fill in parallel, with the same values a single call
    fn(0, nelem) gives (every range evaluates the same operations on its own nodes)."""
    workers = min(16, os.cpu_count() or 1)
    if workers <= 1 or nelem < 2 * min_chunk:
        fn(0, nelem)
        return
    n = max(min_chunk, -(-nelem // (4 * workers)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(lambda e0: fn(e0, min(nelem, e0 + n)), range(0, nelem, n)))

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


def _d_along(D: np.ndarray, f: np.ndarray, axis: int) -> np.ndarray:
    """sum_n D[i, n] f[.., n, ..] along reference axis `axis` of f (Nq, Nq, Nq, E),
    accumulated over n in order as the reference's loops do."""
    Nq = D.shape[0]
    out = np.zeros_like(f)
    for n in range(Nq):
        sl = [slice(None)] * 4
        sl[axis] = slice(n, n + 1)
        shape = [1, 1, 1, 1]
        shape[axis] = Nq
        out += D[:, n].reshape(shape) * f[tuple(sl)]
    return out


def curvilinear_geometry(x: np.ndarray, D: np.ndarray, w: np.ndarray, fmask: np.ndarray):
    """vgeo (Np, 25, E) and sgeo (5, Nfp, 6, E) of hexahedra with node coordinates
    x (3, Np, E) (node n = i + Nq j + Nq^2 k), as computegeometry builds them
    (Grids.jl:1028-1124): x_xi by the derivative matrix
    (compute_reference_to_physical_coord_jacobian!, Metrics.jl:215-266), the
    curl-invariant d xi / d x = D (x_m d x_l / d xi) / 2J and then x_xi replaced by
    the inverse of that discrete metric (computemetric!, Metrics.jl:431-722), face
    normals and surface Jacobians from J d xi_a / d x at the face nodes, mass
    matrices with the LGL weights, and the horizontal mass (horizontal_metrics!,
    Grids.jl:1135-1157)."""
    Nq = D.shape[0]
    Np, Nfp = Nq ** 3, Nq * Nq
    E = x.shape[2]
    X = [x[d].reshape(Nq, Nq, Nq, E, order="F") for d in range(3)]
    # xr[d][a] = d x_d / d xi_a
    xr = [[_d_along(D, X[d], a) for a in range(3)] for d in range(3)]
    JcV = np.sqrt(xr[0][2] ** 2 + xr[1][2] ** 2 + xr[2][2] ** 2)
    J = (xr[0][0] * (xr[1][1] * xr[2][2] - xr[2][1] * xr[1][2])
         + xr[1][0] * (xr[2][1] * xr[0][2] - xr[0][1] * xr[2][2])
         + xr[2][0] * (xr[0][1] * xr[1][2] - xr[1][1] * xr[0][2]))
    JI2 = 1.0 / (2 * J)
    # cross terms c[l][a] = x_m d x_n / d xi_a - x_n d x_m / d xi_a for the cyclic
    # (l, m, n) = (1, 2, 3), (2, 3, 1), (3, 1, 2)
    cyc = ((1, 2), (2, 0), (0, 1))
    c = [[X[m] * xr[n][a] - X[n] * xr[m][a] for a in range(3)] for (m, n) in cyc]
    # xi_a x_l: the (a+1, a+2) cyclic pair of reference derivatives of c[l]
    xix = [[None] * 3 for _ in range(3)]  # xix[a][l] = d xi_a / d x_l
    for l in range(3):
        for a in range(3):
            b, g = (a + 1) % 3, (a + 2) % 3
            xix[a][l] = (_d_along(D, c[l][g], b) - _d_along(D, c[l][b], g)) * JI2
    # the x_xi fields are then overwritten with (1/det) A A^T, A the adjugate of the
    # discrete metric, as computemetric! does (Metrics.jl:606-666); that is not the
    # inverse Jacobian (on a unit brick it gives 2, not 1/2), and no DG kernel reads it
    A = [[None] * 3 for _ in range(3)]
    A[0][0] = xix[1][1] * xix[2][2] - xix[1][2] * xix[2][1]
    A[0][1] = xix[0][2] * xix[2][1] - xix[0][1] * xix[2][2]
    A[0][2] = xix[0][1] * xix[1][2] - xix[0][2] * xix[1][1]
    A[1][0] = xix[1][2] * xix[2][0] - xix[1][0] * xix[2][2]
    A[1][1] = xix[0][0] * xix[2][2] - xix[0][2] * xix[2][0]
    A[1][2] = xix[0][2] * xix[1][0] - xix[0][0] * xix[1][2]
    A[2][0] = xix[1][0] * xix[2][1] - xix[1][1] * xix[2][0]
    A[2][1] = xix[0][1] * xix[2][0] - xix[0][0] * xix[2][1]
    A[2][2] = xix[0][0] * xix[1][1] - xix[0][1] * xix[1][0]
    det = xix[0][0] * A[0][0] + xix[1][0] * A[0][1] + xix[2][0] * A[0][2]
    xinv = [[(1.0 / det) * (A[d][0] * A[a][0] + A[d][1] * A[a][1] + A[d][2] * A[a][2]) for a in range(3)]
            for d in range(3)]  # xinv[d][a] = x_d xi_a
    W = np.empty(Np)
    ii, jj, kk = np.arange(Np) % Nq, (np.arange(Np) // Nq) % Nq, np.arange(Np) // (Nq * Nq)
    W[:] = (w[kk] * w[jj]) * w[ii]
    flat = lambda f: f.reshape(Np, E, order="F")
    vgeo = np.zeros((Np, NVGEO, E), order="F")
    for l in range(3):
        for a in range(3):
            vgeo[:, a + 3 * l, :] = flat(xix[a][l])
    Jf = flat(J)
    vgeo[:, M_ID, :] = Jf * W[:, None]
    vgeo[:, MI_ID, :] = 1.0 / vgeo[:, M_ID, :]
    MH = (w[jj] * w[ii])[:, None]
    vgeo[:, MH_ID, :] = MH * np.sqrt((Jf * flat(xix[2][0])) ** 2 + (Jf * flat(xix[2][1])) ** 2
                                     + (Jf * flat(xix[2][2])) ** 2)
    for d in range(3):
        vgeo[:, X1_ID + d, :] = x[d]
    vgeo[:, JCV_ID, :] = flat(JcV)
    for a in range(3):
        for d in range(3):
            vgeo[:, 16 + d + 3 * a, :] = flat(xinv[d][a])
    # surface: n = -/+ J d xi_a / d x on faces 2a+1 / 2a+2, sJ = |n|, n /= sJ
    a_, b_ = np.arange(Nfp) % Nq, np.arange(Nfp) // Nq
    sgeo = np.zeros((5, Nfp, 6, E), order="F")
    for f in range(6):
        ax = f // 2
        sgn = -1.0 if f % 2 == 0 else 1.0
        vid = fmask[f]
        nv = [sgn * Jf[vid] * vgeo[vid, ax + 3 * l, :] for l in range(3)]
        sJ = np.sqrt(nv[0] ** 2 + nv[1] ** 2 + nv[2] ** 2)
        for l in range(3):
            sgeo[N1_ID + l, :, f, :] = nv[l] / sJ
        sgeo[SM_ID, :, f, :] = sJ * (w[b_] * w[a_])[:, None]
        sgeo[VMI_ID, :, f, :] = vgeo[vid, MI_ID, :]
    return vgeo, sgeo


class BrickGrid:
    """Brick-topology DG grid, axis-aligned unless warped.

    ``brickrange``: three 1-D arrays of element break points (like the
    reference's ``brickrange`` tuple); ``periodicity``: three bools;
    ``boundary``: ((xmin_tag, xmax_tag), (ymin, ymax), (zmin, zmax)) tags used on
    non-periodic faces (AtmosLESConfiguration default ((0,0),(0,0),(1,2)));
    ``meshwarp``: optional vectorised map (x1, x2, x3) -> (x1', x2', x3') applied to
    the node coordinates, as the reference grid's ``meshwarp`` keyword.  A warp
    must map periodic faces onto their images shifted by the period.
    ``lean``: keep only the brick description (break points, periodicity, tags) and
    per-element sizes, no per-node vgeo/sgeo/vmaps (the library builds its element
    records from the description, cmdg_create_brick); coordinates and the mass matrix
    are formed on demand.  For large meshes: a 96^3 brick's per-node arrays are ~30 GB.
    """

    def __init__(self, brickrange, N: int = 4, periodicity=(True, True, True),
                 boundary=((1, 1), (1, 1), (1, 2)), meshwarp=None, lean: bool = False):
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
        self.lean = bool(lean)
        self.warped = meshwarp is not None
        self._h, self._x0, self._W = (h1, h2, h3), (x0, y0, z0), W
        self._coords = None
        if self.lean:
            if self.warped:
                raise ValueError("a lean grid is an unwarped brick")
            self.vgeo = self.sgeo = self.vmap_minus = self.vmap_plus = self.elemtobndy = None
            return

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
        if self.warped:
            # creategrid! (Metrics.jl:85-120): trilinear blend of the element's corners,
            # then the warp of every node (Grids.jl:1049-1060), then the metrics
            xs = []
            for d, (l0, l1) in enumerate(((x0, x0 + h1), (y0, y0 + h2), (z0, z0 + h3))):
                t = r[(ii, jj, kk)[d]][:, None]
                xs.append(((1 - t) * l0[None, :] + (1 + t) * l1[None, :]) / 2)
            xw = meshwarp(*xs)
            xw = np.stack([np.broadcast_to(np.asarray(c, dtype=np.float64), xs[0].shape) for c in xw])
            self.vgeo, self.sgeo = curvilinear_geometry(xw, self.D, w, self.fmask)

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
        if not self.lean:
            return (self.vgeo[:, X1_ID, :], self.vgeo[:, X2_ID, :], self.vgeo[:, X3_ID, :])
        if self._coords is None:
            r = self.xi
            xs = tuple(np.empty((self.Np, self.nelem), order="F") for _ in range(3))
            rr = [((r[self.node_ijk[d]] + 1) / 2)[:, None] for d in range(3)]

            def fill(e0, e1):
                for d in range(3):
                    xs[d][:, e0:e1] = self._x0[d][None, e0:e1] + rr[d] * self._h[d][None, e0:e1]
            over_elements(self.nelem, fill)
            self._coords = xs
        return self._coords

    @property
    def mass(self):
        """(Np, nelem) mass matrix: the MPIStateArray weights (create_states.jl:14-16)."""
        if not self.lean:
            return self.vgeo[:, M_ID, :]
        h1, h2, h3 = self._h
        return ((h1 / 2) * (h2 / 2) * (h3 / 2))[None, :] * self._W[:, None]

    @property
    def element_sizes(self):
        """(h1, h2, h3): each element's edge lengths (unwarped bricks)."""
        if self.warped:
            raise ValueError("a warped grid has no element edge lengths")
        return self._h

    @property
    def ndof_nodes(self) -> int:
        return self.Np * self.nelem

    def min_node_distance(self) -> float:
        """Smallest distance between neighbouring nodes of an element along a
        reference direction (the distance of kernel_min_neighbor_distance!,
        Grids.jl:1228-1334); the smallest LGL gap times the smallest element edge on
        an unwarped brick."""
        if not self.warped:
            gaps = np.diff(self.xi).min() / 2
            return float(min(np.min(np.diff(b)) for b in self.brickrange) * gaps)
        Nq, E = self.Nq, self.nelem
        X = np.stack([self.vgeo[:, X1_ID + d, :].reshape(Nq, Nq, Nq, E, order="F") for d in range(3)])
        best = np.inf
        for a in range(3):
            d = np.diff(X, axis=1 + a)
            best = min(best, float(np.sqrt((d ** 2).sum(axis=0)).min()))
        return best


def uniform_brick(n, lo, hi, N=4, periodicity=(True, True, True),
                  boundary=((1, 1), (1, 1), (1, 2)), meshwarp=None, lean: bool = False) -> BrickGrid:
    """Brick of n = (nx, ny, nz) equal elements on [lo, hi] (optionally warped or lean)."""
    br = [np.linspace(lo[d], hi[d], n[d] + 1) for d in range(3)]
    return BrickGrid(br, N=N, periodicity=periodicity, boundary=boundary, meshwarp=meshwarp, lean=lean)

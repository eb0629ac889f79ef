"""Curvilinear grid geometry (computegeometry with a meshwarp, Grids.jl:1028-1124,
Metrics.jl:431-722) against closed forms: affine non-orthogonal elements, where the
curl-invariant metric is the exact inverse Jacobian; and the warp of the reference's
3-D MMS test (mms_bc_atmos.jl:178-188), where faces must be watertight, the discrete
metric identities hold to roundoff and the mass matrix integrates the warped volume."""
import math

import numpy as np
import pytest

from cmdg.grid import (JCV_ID, MH_ID, MI_ID, M_ID, N1_ID, SM_ID, VMI_ID, X1_ID, lglpoints, uniform_brick,
                       _d_along)

A = np.array([[1.0, 0.3, -0.2], [0.1, 0.9, 0.25], [-0.15, 0.05, 1.2]])


def mms_warp(x1, x2, x3):
    """The 3-D warp of mms_bc_atmos.jl:178-188."""
    return (x1 + (x1 - 0.5) * np.cos(2 * np.pi * x2 * x3) / 4,
            x2 + np.exp(np.sin(2 * np.pi * (x1 * x2 + x3))) / 20,
            x3 + x1 / 4 + x2 ** 2 / 2 + np.sin(x1 * x2 * x3))


def test_affine_non_orthogonal_elements_closed_form():
    n, L = (2, 3, 2), (2.0, 3.0, 1.0)
    warp = lambda x, y, z: tuple(A[i, 0] * x + A[i, 1] * y + A[i, 2] * z for i in range(3))
    g = uniform_brick(n, (0, 0, 0), L, N=4, meshwarp=warp)
    h = np.array(L) / np.array(n)
    dxdxi = A @ np.diag(h / 2)            # x_d xi_a at [d, a]
    xix = np.linalg.inv(dxdxi)            # xi_a x_d at [a, d]
    J = np.linalg.det(dxdxi)
    for a in range(3):
        for d in range(3):
            assert np.abs(g.vgeo[:, a + 3 * d] - xix[a, d]).max() < 1e-12
    _, w = lglpoints(4)
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    assert np.allclose(g.vgeo[:, M_ID], J * W[:, None], rtol=1e-13)
    assert np.allclose(g.vgeo[:, MI_ID] * g.vgeo[:, M_ID], 1.0, rtol=1e-15)
    assert g.mass.sum() == pytest.approx(np.prod(L) * np.linalg.det(A), rel=1e-13)
    # JcV = |d x / d xi_3|; MH = J |grad xi_3| w_i w_j
    assert np.allclose(g.vgeo[:, JCV_ID], np.linalg.norm(dxdxi[:, 2]), rtol=1e-13)
    Wh = np.einsum("j,i->ji", w, w).ravel()
    mh = J * np.linalg.norm(xix[2]) * np.tile(Wh, 5)
    assert np.allclose(g.vgeo[:, MH_ID], mh[:, None], rtol=1e-12)
    # faces: outward unit normal along +-grad xi_a; the surface mass of a face sums to
    # the area of the parallelogram spanned by the two in-face edge vectors
    for f in range(6):
        ax, sgn = f // 2, (-1.0 if f % 2 == 0 else 1.0)
        nrm = sgn * xix[ax] / np.linalg.norm(xix[ax])
        assert np.abs(g.sgeo[N1_ID:N1_ID + 3, :, f, :] - nrm[:, None, None]).max() < 1e-12
        b, c = [x for x in range(3) if x != ax]
        area = np.linalg.norm(np.cross(dxdxi[:, b] * 2, dxdxi[:, c] * 2))
        assert np.allclose(g.sgeo[SM_ID, :, f, :].sum(axis=0), area, rtol=1e-12)
        assert np.array_equal(g.sgeo[VMI_ID, :, f, :], g.vgeo[g.fmask[f], MI_ID, :])


@pytest.fixture(scope="module")
def mms_grid():
    return uniform_brick((4, 4, 4), (0, 0, 0), (1.0, 1.0, 1.0), N=4, periodicity=(False,) * 3,
                         boundary=((1, 1), (1, 1), (1, 1)), meshwarp=mms_warp)


def test_curved_faces_watertight(mms_grid):
    """At every interior face node both sides sit at the same point, with opposite
    unit normals and equal surface mass."""
    g = mms_grid
    X = np.stack(g.coords).reshape(3, -1, order="F")
    for f in range(6):
        inner = g.elemtobndy[f] == 0
        m = (g.vmap_minus[:, f, :] - 1)[:, inner]
        p = (g.vmap_plus[:, f, :] - 1)[:, inner]
        assert np.abs(X[:, m] - X[:, p]).max() < 1e-14
        e2 = p // g.Np
        en = np.nonzero(inner)[0]
        nm = g.sgeo[N1_ID:N1_ID + 3, :, f, :][:, :, en]
        sm = g.sgeo[SM_ID, :, f, :][:, en]
        # the neighbour's face f^1 lists the same nodes in the same order (conforming brick)
        np_ = g.sgeo[N1_ID:N1_ID + 3, :, f ^ 1, :][:, :, e2[0]]
        sp = g.sgeo[SM_ID, :, f ^ 1, :][:, e2[0]]
        assert np.abs(nm + np_).max() < 1e-12
        assert np.allclose(sm, sp, rtol=1e-12)


def test_curl_invariant_metric_identities(mms_grid):
    """sum_a d/d xi_a (J d xi_a / d x_d) = 0 at every node for d = 1, 2, 3 (to
    roundoff): the property that makes a uniform flow a discrete steady state."""
    g = mms_grid
    Nq, E = g.Nq, g.nelem
    J = (g.vgeo[:, M_ID] / np.einsum("k,j,i->kji", g.omega, g.omega, g.omega).ravel()[:, None])
    for d in range(3):
        tot = 0
        for a in range(3):
            f = (J * g.vgeo[:, a + 3 * d]).reshape(Nq, Nq, Nq, E, order="F")
            tot = tot + _d_along(g.D, f, a)
        assert np.abs(tot).max() < 1e-12, (d, np.abs(tot).max())
    # and it approximates the inverse of the interpolated Jacobian (not exactly: up to
    # 9 % where this strong warp is least resolved by 4^3 elements of degree 4)
    X = [g.vgeo[:, X1_ID + d].reshape(Nq, Nq, Nq, E, order="F") for d in range(3)]
    for a in range(3):
        for b in range(3):
            prod = sum(g.vgeo[:, a + 3 * d].reshape(Nq, Nq, Nq, E, order="F") * _d_along(g.D, X[d], b)
                       for d in range(3))
            assert np.abs(prod - (a == b)).max() < 0.15


def test_mass_matrix_integrates_warped_volume(mms_grid):
    """sum of M = volume of the warped unit cube = integral of det(d warp / d x),
    computed independently with a 40-point Gauss rule per direction."""
    xg, wg = np.polynomial.legendre.leggauss(40)
    x = (xg + 1) / 2
    wq = wg / 2
    X1, X2, X3 = np.meshgrid(x, x, x, indexing="ij")
    eps = 1e-6
    cols = []
    for d in range(3):
        dp = [X1, X2, X3]
        dm = [X1, X2, X3]
        dp = [c + (eps if k == d else 0) for k, c in enumerate(dp)]
        dm = [c - (eps if k == d else 0) for k, c in enumerate(dm)]
        cols.append((np.stack(mms_warp(*dp)) - np.stack(mms_warp(*dm))) / (2 * eps))
    Jm = np.stack(cols, axis=1)  # (3 out, 3 in, ...)
    det = np.linalg.det(np.moveaxis(Jm, (0, 1), (-2, -1)))
    vol = np.einsum("i,j,k,ijk->", wq, wq, wq, det)
    assert mms_grid.mass.sum() == pytest.approx(vol, rel=1e-6)
    assert abs(vol - 1.0) > 1e-3  # the warp does change the volume

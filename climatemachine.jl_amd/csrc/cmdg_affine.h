// cmdg_affine.h -- DG kernels for meshes of axis-aligned affine bricks.
//
// The context detects (bitwise, at cmdg_create) that every element's metric terms
// are a constant diagonal d xi / d x, that M = J * W[n] and sM = sJ_f * Wf[fn] with
// the tensor-product LGL weights, that face normals are +-e_d and that face
// neighbours are conforming (vmap+ = the opposite face of one element).  Geometry
// then shrinks from 11 doubles per node plus 5 per face node to 11 doubles and 6
// ints per element, and every product with an exactly-zero metric or normal
// component of the general formula is dropped; for finite data this changes no
// result bit.
//
// Neighbour data reaches a face through face traces: face-major arrays
// (nelem, 6, nvar, Nfp) in which the 25 values a neighbour needs for one face and
// one variable are contiguous.  The state trace is written by whoever produces
// the state (the LSRK epilogue for the next stage, or k_trace_aff), the trace of
// the second-order flux row by the gradient kernel, and the trace of the static
// fields (Phi, p_ref) once at set-up.  A neighbour read is then two 128-byte lines
// per variable instead of one line per node.
#pragma once
#include "cmdg_physics.h"

namespace cmdg {

// per-element record: xi11, xi22, xi33, J, sJ (6 faces), Delta
constexpr int AFF_GEO = 11;
// trace widths
constexpr int TR_Q = 5;   // rho, rhou (3), rhoe
constexpr int TR_F2 = 4;  // second-order flux row d, states 1..4 (the mass entry is 0)
constexpr int TR_S = 2;   // Phi, p_ref

struct AffArrays {
    const double *geo;    // (nelem, AFF_GEO)
    const int32_t *nbe;   // (nelem, 6) neighbour element across face f, -1 at a wall
    const double *W;      // (Np)  (w_k w_j) w_i
    const double *Wf;     // (Nfp) w_b w_a for face node a + Nq b
    const double *D;      // (Nq, Nq)
    const double *str;    // (nelem, 6, TR_S, Nfp) static trace
};

template <int Nfp> __device__ __forceinline__ int64_t tri(int64_t e, int f, int v, int nv, int fn) {
    return ((e * 6 + f) * nv + v) * (int64_t)Nfp + fn;
}

// volume node of the neighbour that faces node n across face f (conforming bricks)
template <int Nq> __device__ __forceinline__ int opposite_node(int n, int f) {
    constexpr int N = Nq - 1;
    switch (f) {
    case 0: return n + N;
    case 1: return n - N;
    case 2: return n + N * Nq;
    case 3: return n - N * Nq;
    case 4: return n + N * Nq * Nq;
    default: return n - N * Nq * Nq;
    }
}

// ------------------------------------------------------------------ state trace
// Writes the face-node trace of Q (and optionally a copy of Q).
template <int Nq>
__global__ void __launch_bounds__(threads_for<Nq>())
k_trace_aff(int64_t nelem, const double *__restrict__ Q, double *__restrict__ qtr, double *__restrict__ Qcopy) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    const int64_t e = xcd_element(blockIdx.x, nelem);
    if (e >= nelem) return;
    const int n = threadIdx.x;
    if (n >= Np) return;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    double q[5];
    const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
    for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
    if (Qcopy) {
        double *qc = Qcopy + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) qc[Np * s] = q[s];
    }
    int fl[3], fnl[3];
    const int nf = node_faces<Nq>(i, j, k, fl, fnl);
    for (int c = 0; c < nf; ++c)
#pragma unroll
        for (int s = 0; s < 5; ++s) qtr[tri<Nfp>(e, fl[c], s, TR_Q, fnl[c])] = q[s];
}

// ------------------------------------------------------------------ gradient
template <int Nq, bool SMAG>
__global__ void __launch_bounds__(threads_for<Nq>())
k_gradient_aff(KParams P, AffArrays A, int64_t nelem, const double *__restrict__ Q, const double *__restrict__ qtr,
               double *__restrict__ aux, double *__restrict__ gfo, double *__restrict__ f2tr, int write_aux) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NG = SMAG ? 5 : 4, NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sG[NG][Np];
    const int64_t e = xcd_element(blockIdx.x, nelem);
    if (e >= nelem) return;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double *ge = A.geo + (int64_t)AFF_GEO * e;
    double q[5], thv = 0.0, G[5];
    Aux a;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        a = load_aux(P, aux, Np, n, e);
        Thermo t = thermo(P, q, a.Phi);
        thv = theta_v(P, t);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = thv;
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        grad_arg<SMAG>(P, q, t, thv, G);
#pragma unroll
        for (int s = 0; s < NG; ++s) sG[s][n] = G[s];
    }
    __syncthreads();
    if (!act) return;

    // horizontal pass then vertical pass, transform applied to each and summed
    const double xi0 = ge[0], xi1 = ge[1], xi2 = ge[2];
    double gH[3][5], gV[3][5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        double g1 = 0.0, g2 = 0.0, g3 = -0.0;
        if (s < NG) {
#pragma unroll
            for (int m = 0; m < Nq; ++m) {
                g1 += sD[i + Nq * m] * sG[s][m + Nq * (j + Nq * k)];
                g2 += sD[j + Nq * m] * sG[s][i + Nq * (m + Nq * k)];
            }
#pragma unroll
            for (int m = 0; m < Nq; ++m) g3 += sD[k + Nq * m] * sG[s][i + Nq * (j + Nq * m)];
        }
        gH[0][s] = xi0 * g1;
        gH[1][s] = xi1 * g2;
        gH[2][s] = 0.0;
        gV[0][s] = 0.0;
        gV[1][s] = 0.0;
        gV[2][s] = xi2 * g3;
    }
    double gf[10], tmp[10];
    grad_flux<SMAG>(gH, a, thv, gf);
    grad_flux<SMAG>(gV, a, thv, tmp);
#pragma unroll
    for (int s = 0; s < NGF; ++s) gf[s] += tmp[s];

    // lifted central jumps M^-1 Mf (Phi(n G*) - Phi(n G-)) of the node's faces
    const double MI = 1.0 / (ge[3] * A.W[n]);
    int fl[3], fnl[3];
    const int nf = node_faces<Nq>(i, j, k, fl, fnl);
    for (int c = 0; c < nf; ++c) {
        const int f = fl[c], d = f >> 1, fn = fnl[c];
        const double nd = (f & 1) ? 1.0 : -1.0;
        const double w = MI * (ge[4 + f] * A.Wf[fn]);
        const int32_t ne = A.nbe[6 * e + f];
        double Gp[5], Gs[5], nG[5];
        if (ne >= 0) {
            const int fo = f ^ 1;
            double qp[5];
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = qtr[tri<Nfp>(ne, fo, s, TR_Q, fn)];
            const double Php = A.str[tri<Nfp>(ne, fo, 0, TR_S, fn)];
            Thermo tp = thermo(P, qp, Php);
            grad_arg<SMAG>(P, qp, tp, SMAG ? theta_v(P, tp) : 0.0, Gp);
#pragma unroll
            for (int s = 0; s < 5; ++s) Gs[s] = nd * (Gp[s] + G[s]) * 0.5;
        } else {
            const double nv[3] = {d == 0 ? nd : 0.0, d == 1 ? nd : 0.0, d == 2 ? nd : 0.0};
            double qb[5];
            wall_state(q, nv, 1.0, qb);
            Thermo tb = thermo(P, qb, a.Phi);
            grad_arg<SMAG>(P, qb, tb, SMAG ? theta_v(P, tb) : 0.0, Gp);
#pragma unroll
            for (int s = 0; s < 5; ++s) Gs[s] = nd * Gp[s];
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) nG[s] = nd * G[s];
        double gfs[10], gfm[10];
        grad_flux_row<SMAG>(d, Gs, a.gPhi, thv, gfs);
        grad_flux_row<SMAG>(d, nG, a.gPhi, thv, gfm);
#pragma unroll
        for (int s = 0; s < NGF; ++s) gf[s] += w * (gfs[s] - gfm[s]);
    }
    double *go = gfo + (int64_t)Np * NGF * e + n;
#pragma unroll
    for (int s = 0; s < NGF; ++s) go[Np * s] = gf[s];

    // trace of the second-order flux row of every face this node lies on
    if (nf > 0) {
        double nu[3];
        viscosity<SMAG>(P, q[0], gf, a.gPhi, ge[10], nu);
        for (int c = 0; c < nf; ++c) {
            double F2[5];
            flux2_row<SMAG>(P, q, gf, nu, fl[c] >> 1, F2);
#pragma unroll
            for (int v = 0; v < TR_F2; ++v) f2tr[tri<Nfp>(e, fl[c], v, TR_F2, fnl[c])] = F2[1 + v];
        }
    }
}

// ------------------------------------------------------------------ tendency
// MODE 0: tend = alpha*RHS + beta*tend.   MODE 1: LSRK stage (tend is dQ):
//   dQ <- RHS + dQ; Qout = Qin + bdt*dQ; dQ *= rka; the state trace of Qout is
//   written to qtr_out when given.
template <int Nq, bool DIFF, bool SMAG, int MODE>
__global__ void __launch_bounds__(threads_for<Nq>())
k_tendency_aff(KParams P, AffArrays A, int64_t nelem, const double *__restrict__ Q, const double *__restrict__ qtr,
               const double *__restrict__ gfi, const double *__restrict__ f2tr, double *__restrict__ aux,
               double *__restrict__ tend, double alpha, double beta, double *__restrict__ Qout,
               double *__restrict__ qtr_out, double rka, double bdt, int write_aux) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sF[3][5][Np];
    const int64_t e = xcd_element(blockIdx.x, nelem);
    if (e >= nelem) return;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double *ge = A.geo + (int64_t)AFF_GEO * e;

    double q[5], gf[10], u[3], nu[3], p = 0.0, pm = 0.0, MI = 0.0, temp = 0.0;
    Aux a;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        a = load_aux(P, aux, Np, n, e);
        a.Delta = ge[10];
        Thermo t = thermo(P, q, a.Phi);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = theta_v(P, t);
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        p = t.p;
        temp = t.temp;
        pm = P.ref_sub ? (p - a.ref_p) : p;
        u[0] = q[1] / q[0];
        u[1] = q[2] / q[0];
        u[2] = q[3] / q[0];
        if (DIFF) {
            const double *gb = gfi + (int64_t)Np * NGF * e + n;
#pragma unroll
            for (int s = 0; s < NGF; ++s) gf[s] = gb[Np * s];
            viscosity<SMAG>(P, q[0], gf, a.gPhi, a.Delta, nu);
        }
        const double M = ge[3] * A.W[n];
        MI = 1.0 / M;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            double F[5], F2[5];
            flux1_row(q, u, p, pm, ax, F);
            if (DIFF) {
                flux2_row<SMAG>(P, q, gf, nu, ax, F2);
#pragma unroll
                for (int s = 0; s < 5; ++s) F[s] += F2[s];
            }
            const double x = ge[ax];
#pragma unroll
            for (int s = 0; s < 5; ++s) sF[ax][s][n] = M * (x * F[s]);
        }
    }
    __syncthreads();
    if (!act) return;

    double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (P.gravity) {
        // Gravity() source (tendencies_momentum.jl:60-72)
        const double drho = P.ref_src ? (q[0] - a.ref_rho) : q[0];
#pragma unroll
        for (int d = 0; d < 3; ++d) src[1 + d] = (-drho) * a.gPhi[d];
    }
    double *tb = tend + (int64_t)Np * 5 * e + n;
    double T[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        double th = 0.0;
#pragma unroll
        for (int m = 0; m < Nq; ++m) {
            th += MI * sD[m + Nq * i] * sF[0][s][m + Nq * (j + Nq * k)];
            th += MI * sD[m + Nq * j] * sF[1][s][i + Nq * (m + Nq * k)];
        }
        double out;
        if (MODE == 0) out = (beta != 0.0) ? alpha * th + beta * tb[Np * s] : alpha * th;
        else out = th + tb[Np * s];
        double tv = 0.0;
#pragma unroll
        for (int m = 0; m < Nq; ++m) {
            tv += MI * sD[m + Nq * k] * sF[2][s][i + Nq * (j + Nq * m)];
            if (m == k) tv += src[s];
        }
        T[s] = (MODE == 0 ? alpha * tv : tv) + out;
    }

    // face fluxes of the faces this node lies on
    int fl[3], fnl[3];
    const int nf = node_faces<Nq>(i, j, k, fl, fnl);
    const double rim = 1.0 / q[0];
    const double cm = P.rusanov ? sqrt(P.gammaR * temp) : 0.0;
    for (int c = 0; c < nf; ++c) {
        const int f = fl[c], d = f >> 1, fn = fnl[c];
        const double nd = (f & 1) ? 1.0 : -1.0;
        const double h = nd * 0.5;
        const double w = (MODE == 0 ? alpha * MI : MI) * (ge[4 + f] * A.Wf[fn]);
        const int32_t ne = A.nbe[6 * e + f];
        const bool wall = ne < 0;
        const int fo = f ^ 1;
        double qp[5];
        double Php = a.Phi, refpp = a.ref_p;
        if (!wall) {
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = qtr[tri<Nfp>(ne, fo, s, TR_Q, fn)];
            Php = A.str[tri<Nfp>(ne, fo, 0, TR_S, fn)];
            refpp = A.str[tri<Nfp>(ne, fo, 1, TR_S, fn)];
        } else {
            const double nv[3] = {d == 0 ? nd : 0.0, d == 1 ? nd : 0.0, d == 2 ? nd : 0.0};
            wall_state(q, nv, 2.0, qp);
        }
        const Thermo tp = thermo(P, qp, Php);
        const double up[3] = {qp[1] / qp[0], qp[2] / qp[0], qp[3] / qp[0]};
        double Fm[5], Fp[5], flux[5];
        flux1_row(q, u, p, pm, d, Fm);
        flux1_row(qp, up, tp.p, P.ref_sub ? (tp.p - refpp) : tp.p, d, Fp);
#pragma unroll
        for (int s = 0; s < 5; ++s) flux[s] = (Fm[s] + Fp[s]) * h;
        if (P.rusanov) {
            const double rip = 1.0 / qp[0];
            const double qmd = d == 0 ? q[1] : (d == 1 ? q[2] : q[3]);
            const double qpd = d == 0 ? qp[1] : (d == 1 ? qp[2] : qp[3]);
            const double wm = fabs(nd * (rim * qmd)) + cm;
            const double wp = fabs(nd * (rip * qpd)) + sqrt(P.gammaR * tp.temp);
            const double mw = fmax(wm, wp);
#pragma unroll
            for (int s = 0; s < 5; ++s) flux[s] += (mw * (q[s] - qp[s])) * 0.5;
        }
        if (DIFF && !wall) {
            flux2_row<SMAG>(P, q, gf, nu, d, Fm);
            Fp[0] = 0.0;
#pragma unroll
            for (int v = 0; v < TR_F2; ++v) Fp[1 + v] = f2tr[tri<Nfp>(ne, fo, v, TR_F2, fn)];
#pragma unroll
            for (int s = 0; s < 5; ++s) flux[s] += (Fm[s] + Fp[s]) * h;
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) T[s] -= w * flux[s];
    }

    if (MODE == 0) {
#pragma unroll
        for (int s = 0; s < 5; ++s) tb[Np * s] = T[s];
    } else {
        double *qo = Qout + (int64_t)Np * 5 * e + n;
        double qn[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            qn[s] = q[s] + bdt * T[s];
            qo[Np * s] = qn[s];
            tb[Np * s] = T[s] * rka;
        }
        if (qtr_out)
            for (int c = 0; c < nf; ++c)
#pragma unroll
                for (int s = 0; s < 5; ++s) qtr_out[tri<Nfp>(e, fl[c], s, TR_Q, fnl[c])] = qn[s];
    }
}

}  // namespace cmdg

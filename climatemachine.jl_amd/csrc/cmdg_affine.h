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
// A face reads its neighbour's state straight from the neighbour's volume state
// (25 values per face and variable; the kernels that update the state write it to
// another buffer than the one they read).  Derived neighbour data reaches a face
// through face traces: face-major arrays (nelem, 6, nvar, Nfp) in which the 25
// values a neighbour needs for one face and one variable are contiguous -- the
// second-order flux rows, written by the gradient kernel, and the static fields
// (Phi, p_ref), written once at set-up.
#pragma once
#include "cmdg_physics.h"

namespace cmdg {

// per-element record: xi11, xi22, xi33, J, sJ (6 faces), Delta
constexpr int AFF_GEO = 11;
// trace widths
constexpr int TR_F2 = 4;  // second-order flux row d, states 1..4 (the mass entry is 0)
constexpr int TR_S = 2;   // Phi, p_ref

struct AffArrays {
    const double *geo;    // (nelem, AFF_GEO)
    const int32_t *nbe;   // (nelem, 6) neighbour element across face f, -1 at a wall
    const double *W;      // (Np)  (w_k w_j) w_i
    const double *Wf;     // (Nfp) w_b w_a for face node a + Nq b
    const double *D;      // (Nq, Nq)
    const double *str;    // (nelem, 6, TR_S, Nfp) static trace
    int32_t aux_cont;     // 1: Phi and p_ref agree bitwise on both sides of every interior face
    // Static auxiliary fields (Phi, grad Phi, rho_ref, p_ref, Delta): elements whose
    // static fields agree bitwise share one class, and the kernels read those fields
    // from the class's block of sa, (Np, naux, nclass) in the aux layout.  Without
    // classes, cls is null and sa is the aux array itself (class = element).
    const int32_t *cls;   // (nelem) class of each element, or null
    const double *sa;
    const float *gphi32;  // (Np, 3, nclass) grad Phi rounded to FP32 (mixed path; null without it)
};

// the class whose static fields element e reads
__device__ __forceinline__ int64_t static_class(const AffArrays &A, int64_t e) {
    return A.cls ? (int64_t)A.cls[e] : e;
}

// The elements one launch works on: the range first .. first + count - 1, or, with a
// list, list[0 .. count - 1].  Block b takes the element at position xcd_element(b, count):
// each XCD gets one contiguous run of the set.
struct ElemSet {
    const int32_t *list;
    int64_t first, count;
};
__device__ __forceinline__ int64_t set_element(const ElemSet &S, int64_t b) {
    const int64_t p = xcd_element(b, S.count);
    if (p >= S.count) return -1;
    return S.list ? (int64_t)S.list[p] : S.first + p;
}

template <int Nfp> __device__ __forceinline__ int64_t tri(int64_t e, int f, int v, int nv, int fn) {
    return ((e * 6 + f) * nv + v) * (int64_t)Nfp + fn;
}

// volume node of face node fn = a + Nq b of face f (the inverse of node_face_dir): the
// coordinate along the face's normal is 0 or N, the other two are (a, b) in order.
// Selects rather than a switch: the face tasks of one wave lie on up to three faces, and
// a switch over f branches once per face present.
template <int Nq> __device__ __forceinline__ int face_node(int f, int fn) {
    const int a = fn % Nq, b = fn / Nq, d = f >> 1, o = (f & 1) * (Nq - 1);
    const int i = d == 0 ? o : a;
    const int j = d == 1 ? o : (d == 0 ? a : b);
    const int k = d == 2 ? o : b;
    return i + Nq * (j + Nq * k);
}

// Face work is spread over all lanes of the workgroup: the 6 Nfp face nodes are
// tasks t = f Nfp + fn dealt round-robin (150 tasks on 128 lanes at N = 4 take 2 + 1
// wave-iterations, where one lane per volume node would run its up-to-3 faces in 3 +
// 3), each task's result goes to LDS, and each node then adds the results of its
// faces in face order (the reference's order).

// A static auxiliary field (Phi, p_ref) at volume node vn of element e, or dflt when
// the model has no such field.
template <int Nq>
__device__ __forceinline__ double own_aux(const KParams &P, const double *__restrict__ aux, int64_t e, int vn,
                                          int field, double dflt) {
    constexpr int Np = Nq * Nq * Nq;
    return field >= 0 ? aux[(int64_t)Np * (P.naux * e + field) + vn] : dflt;
}

// ------------------------------------------------------------------ gradient
// Lifted-jump term Phi(n G*) - Phi(n G-) of face node fn of face f of element e
// (DGModel_kernels.jl:1471-1647, CentralNumericalFluxGradient NumericalFluxes.jl:67-123,
// walls: G* = n G(q_wall), bc_momentum.jl:24-43).  G- and theta_v- come from the
// element's own gradient arguments in LDS, q+ from the neighbour's volume state.
template <int Nq, bool SMAG, bool BUOY>
__device__ __forceinline__ void face_gradient_aff(const KParams &P, const AffArrays &A, int64_t e, int f, int fn,
                                                  int32_t ne, const double *__restrict__ Q,
                                                  const double (*sG)[Nq * Nq * Nq], double *dgf) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq, NG = SMAG ? 5 : 4, NGF = SMAG ? 10 : 9;
    const int d = f >> 1, vn = face_node<Nq>(f, fn);
    const double nd = (f & 1) ? 1.0 : -1.0;
    double G[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) G[s] = s < NG ? sG[s][vn] : 0.0;
    const double thv = SMAG ? G[4] : 0.0;
    double gPhi[3] = {0.0, 0.0, 0.0};
    const int64_t ce = static_class(A, e);
    if (P.a_gPhi >= 0) {
        const double *ab = A.sa + (int64_t)Np * P.naux * ce + vn;
#pragma unroll
        for (int c = 0; c < 3; ++c) gPhi[c] = ab[(int64_t)Np * (P.a_gPhi + c)];
    }
    double Gp[5], Gs[5], nG[5];
    if (ne >= 0) {
        const int fo = f ^ 1;
        // the neighbour's state at the shared node, from its volume (the gradient pass
        // writes no state, so every block reads the stage's state)
        double qp[5];
        const int64_t qb = (int64_t)Np * 5 * ne + face_node<Nq>(fo, fn);
#pragma unroll
        for (int s = 0; s < 5; ++s) qp[s] = Q[qb + Np * s];
        const double Php = A.aux_cont ? own_aux<Nq>(P, A.sa, ce, vn, P.a_Phi, -0.0) : A.str[tri<Nfp>(ne, fo, 0, TR_S, fn)];
        Thermo tp = thermo(P, qp, Php);
        grad_arg<SMAG>(P, qp, tp, SMAG ? (BUOY ? theta_v(P, tp) : 1.0) : 0.0, Gp);
#pragma unroll
        for (int s = 0; s < 5; ++s) Gs[s] = nd * (Gp[s] + G[s]) * 0.5;
    } else {
        double q[5], qb[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = Q[(int64_t)Np * (s + 5 * e) + vn];
        const double Phi = own_aux<Nq>(P, A.sa, ce, vn, P.a_Phi, -0.0);
        const double nv[3] = {d == 0 ? nd : 0.0, d == 1 ? nd : 0.0, d == 2 ? nd : 0.0};
        wall_state(q, nv, 1.0, qb);
        Thermo tb = thermo(P, qb, Phi);
        grad_arg<SMAG>(P, qb, tb, SMAG ? (BUOY ? theta_v(P, tb) : 1.0) : 0.0, Gp);
#pragma unroll
        for (int s = 0; s < 5; ++s) Gs[s] = nd * Gp[s];
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) nG[s] = nd * G[s];
    double gfs[10], gfm[10];
    grad_flux_row<SMAG>(d, Gs, gPhi, thv, gfs);
    grad_flux_row<SMAG>(d, nG, gPhi, thv, gfm);
#pragma unroll
    for (int s = 0; s < NGF; ++s) dgf[s] = gfs[s] - gfm[s];
}

// BUOY: as for k_gradient_mix, theta_v and its gradient are evaluated for N^2 only when
// the context has a gravitational potential.
template <int Nq, bool SMAG, bool BUOY>
__global__ void __launch_bounds__(threads_for<Nq>())
k_gradient_aff(KParams P, AffArrays A, ElemSet ES, const double *__restrict__ Q, double *__restrict__ aux,
               double *__restrict__ gfo, double *__restrict__ f2tr, int write_aux) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq, NT = threads_for<Nq>();
    constexpr int NG = SMAG ? 5 : 4, NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sG[5][Np];
    // lifted-jump terms (6 Nfp x NGF)
    __shared__ double sJbuf[6 * Nfp * NGF];
    double(*sJ)[NGF] = reinterpret_cast<double(*)[NGF]>(sJbuf);
    const int64_t e = set_element(ES, blockIdx.x);
    if (e < 0) return;
    __shared__ int32_t nb[6];  // the element's face neighbours
    if (threadIdx.x < 6) nb[threadIdx.x] = A.nbe[6 * e + threadIdx.x];
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double *ge = A.geo + (int64_t)AFF_GEO * e;
    double q[5], thv = 0.0, G[5];
    Aux a;
    constexpr bool buoy = SMAG && BUOY;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        a = load_aux(P, A.sa, Np, n, static_class(A, e));
        Thermo t = thermo(P, q, a.Phi);
        // theta_v enters N^2 = grad theta_v . grad Phi / theta_v only, zero without a
        // gravitational potential (grad Phi = 0): then it is evaluated only for the aux
        // refresh, and 1 stands in the gradient argument (the face terms divide by it)
        thv = (buoy || (write_aux && P.a_theta_v >= 0)) ? theta_v(P, t) : 1.0;
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
    // lifted central jumps of all face nodes, spread over every lane
    for (int t = n; t < 6 * Nfp; t += NT) {
        double dgf[NGF];
        face_gradient_aff<Nq, SMAG, BUOY>(P, A, e, t / Nfp, t % Nfp, nb[t / Nfp], Q, sG, dgf);
#pragma unroll
        for (int s = 0; s < NGF; ++s) sJ[t][s] = dgf[s];
    }
    __syncthreads();

    double gf[10];
    if (act) {
        // horizontal pass then vertical pass, transform applied to each and summed
        const double xi0 = ge[0], xi1 = ge[1], xi2 = ge[2];
        double gH[3][5], gV[3][5];
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            double g1 = 0.0, g2 = 0.0, g3 = -0.0;
            if (s < NG && (s < 4 || buoy)) {
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
        double tmp[10];
        grad_flux<SMAG>(gH, a, thv, gf);
        grad_flux<SMAG>(gV, a, thv, tmp);
#pragma unroll
        for (int s = 0; s < NGF; ++s) gf[s] += tmp[s];
        // the lifted jumps of this node's faces, in face order: M^-1 Mf (...)
        const double MI = 1.0 / (ge[3] * A.W[n]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int fn;
            const int f = node_face_dir<Nq>(i, j, k, d, fn);
            if (f < 0) continue;
            const double w = MI * (ge[4 + f] * A.Wf[fn]);
#pragma unroll
            for (int s = 0; s < NGF; ++s) gf[s] += w * sJ[f * Nfp + fn][s];
        }
        double *go = gfo + (int64_t)Np * NGF * e + n;
#pragma unroll
        for (int s = 0; s < NGF; ++s) go[Np * s] = gf[s];
    }
    // the second-order flux rows of the face nodes, written straight to the face-major
    // trace: the 25 lanes of one face fill its 25 contiguous slots per row entry
    if (act && (i == 0 || i == Nq - 1 || j == 0 || j == Nq - 1 || k == 0 || k == Nq - 1)) {
        double nu[3];
        viscosity<SMAG>(P, q[0], gf, a.gPhi, ge[10], nu);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int fn;
            const int f = node_face_dir<Nq>(i, j, k, d, fn);
            if (f < 0) continue;
            double F2[5];
            flux2_row<SMAG>(P, q, gf, nu, d, F2);
#pragma unroll
            for (int v = 0; v < TR_F2; ++v) f2tr[tri<Nfp>(e, f, v, TR_F2, fn)] = F2[1 + v];
        }
    }
}

// ------------------------------------------------------------------ tendency
// Numerical flux (first + second order, Rusanov or central) at face node fn of face
// f of element e, seen from e (normal +-e_d): NumericalFluxes.jl:219-340, 670-715,
// walls bc_momentum.jl:24-49 / bc_energy.jl:10-17 (zero second-order flux).
template <int Nq, bool DIFF, bool SMAG>
__device__ __forceinline__ void face_flux_aff(const KParams &P, const AffArrays &A, int f, int fn, int32_t ne,
                                              const double (*sQ)[Nq * Nq * Nq],
                                              const double *__restrict__ Q, const double *__restrict__ f2tr,
                                              const double *f2own, double flux[5]) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    const int d = f >> 1, vn = face_node<Nq>(f, fn);
    const double nd = (f & 1) ? 1.0 : -1.0;
    const double h = nd * 0.5;
    const bool wall = ne < 0;
    const int fo = f ^ 1;
    // the own side at the face's volume node, staged in LDS (state rows 0-4, Phi 5,
    // p_ref 6); the neighbour's state from its volume
    double q[5], qp[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) q[s] = sQ[s][vn];
    const double Phi = sQ[5][vn], ref_p = sQ[6][vn];
    double Php = Phi, refpp = ref_p;
    double f2p[TR_F2] = {};  // the neighbour's second-order row, loaded with its state
    if (!wall) {
        const double *qb = Q + (int64_t)Np * 5 * ne + face_node<Nq>(fo, fn);
#pragma unroll
        for (int s = 0; s < 5; ++s) qp[s] = qb[Np * s];
        if (DIFF)
#pragma unroll
            for (int v = 0; v < TR_F2; ++v) f2p[v] = f2tr[tri<Nfp>(ne, fo, v, TR_F2, fn)];
        if (!A.aux_cont) {
            Php = A.str[tri<Nfp>(ne, fo, 0, TR_S, fn)];
            refpp = A.str[tri<Nfp>(ne, fo, 1, TR_S, fn)];
        }
    } else {
        const double nv[3] = {d == 0 ? nd : 0.0, d == 1 ? nd : 0.0, d == 2 ? nd : 0.0};
        wall_state(q, nv, 2.0, qp);
    }
    const Thermo tm = thermo(P, q, Phi);
    const Thermo tp = thermo(P, qp, Php);
    const double u[3] = {q[1] / q[0], q[2] / q[0], q[3] / q[0]};
    const double up[3] = {qp[1] / qp[0], qp[2] / qp[0], qp[3] / qp[0]};
    double Fm[5], Fp[5];
    flux1_row(q, u, tm.p, P.ref_sub ? (tm.p - ref_p) : tm.p, d, Fm);
    flux1_row(qp, up, tp.p, P.ref_sub ? (tp.p - refpp) : tp.p, d, Fp);
#pragma unroll
    for (int s = 0; s < 5; ++s) flux[s] = (Fm[s] + Fp[s]) * h;
    if (P.rusanov) {
        const double rim = 1.0 / q[0], rip = 1.0 / qp[0];
        const double qmd = pick3(d, q[1], q[2], q[3]);
        const double qpd = pick3(d, qp[1], qp[2], qp[3]);
        const double wm = fabs(nd * (rim * qmd)) + sqrt(P.gammaR * tm.temp);
        const double wp = fabs(nd * (rip * qpd)) + sqrt(P.gammaR * tp.temp);
        const double mw = fmax(wm, wp);
#pragma unroll
        for (int s = 0; s < 5; ++s) flux[s] += (mw * (q[s] - qp[s])) * 0.5;
    }
    if (DIFF && !wall) {
        Fm[0] = 0.0;
        Fp[0] = 0.0;
#pragma unroll
        for (int v = 0; v < TR_F2; ++v) {
            Fm[1 + v] = f2own[v];  // the same bits the gradient kernel put in the trace
            Fp[1 + v] = f2p[v];
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) flux[s] += (Fm[s] + Fp[s]) * h;
    }
}

// MODE 0: tend = alpha*RHS + beta*tend.   MODE 1: LSRK stage (tend is dQ):
//   dQ <- RHS + dQ; Qout = Qin + bdt*dQ; dQ *= rka.  Qout is another buffer than Q:
//   blocks read their neighbours' volume state.  MODE 1 takes alpha = 1 and beta = 1,
//   or beta = 0 when dQ holds zeros (the previous stage scaled it by rka = 0): it is then
//   not read; dq_store = 0 leaves the scaled dQ unstored (this stage's rka is 0 and the
//   next stage is launched with beta = 0).
template <int Nq, bool DIFF, bool SMAG, int MODE>
__global__ void __launch_bounds__(threads_for<Nq>())
k_tendency_aff(KParams P, AffArrays A, ElemSet ES, const double *__restrict__ Q,
               const double *__restrict__ gfi, const double *__restrict__ f2tr, double *__restrict__ aux,
               double *__restrict__ tend, double alpha, double beta, double *__restrict__ Qout, double rka,
               double bdt, int write_aux, int dq_store) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq, NT = threads_for<Nq>();
    constexpr int NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    // M(xi_a F) by direction; once every lane has its volume divergence (in registers),
    // the same space holds the face fluxes (6 Nfp x 5) and the element's state, Phi and
    // p_ref for the face tasks (7 x Np), so that the block needs no more LDS
    __shared__ double sF[3][5][Np];
    static_assert(6 * Nfp * 5 + 7 * Np <= 3 * 5 * Np, "face fluxes and staged state fit in sF");
    double(*sFl)[5] = reinterpret_cast<double(*)[5]>(&sF[0][0][0]);
    double(*sQ)[Np] = reinterpret_cast<double(*)[Np]>(&sF[0][0][0] + 6 * Nfp * 5);
    __shared__ double sF2o[DIFF ? 6 * Nfp : 1][TR_F2];  // this element's rows at its face nodes
    const int64_t e = set_element(ES, blockIdx.x);
    if (e < 0) return;
    __shared__ int32_t nb[6];  // the element's face neighbours
    if (threadIdx.x < 6) nb[threadIdx.x] = A.nbe[6 * e + threadIdx.x];
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double *ge = A.geo + (int64_t)AFF_GEO * e;

    double q[5], gf[10], u[3], nu[3], p = 0.0, pm = 0.0, MI = 0.0;
    Aux a;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        a = load_aux(P, A.sa, Np, n, static_class(A, e));
        a.Delta = ge[10];
        Thermo t = thermo(P, q, a.Phi);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = theta_v(P, t);
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        p = t.p;
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
                int fn;
                const int f = node_face_dir<Nq>(i, j, k, ax, fn);
                if (f >= 0)
#pragma unroll
                    for (int v = 0; v < TR_F2; ++v) sF2o[f * Nfp + fn][v] = F2[1 + v];
#pragma unroll
                for (int s = 0; s < 5; ++s) F[s] += F2[s];
            }
            const double x = ge[ax];
#pragma unroll
            for (int s = 0; s < 5; ++s) sF[ax][s][n] = M * (x * F[s]);
        }
    }
    __syncthreads();
    // volume divergence and source of this node, into registers
    double T[5];
    if (act) {
        double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (P.gravity) {
            // Gravity() source (tendencies_momentum.jl:60-72)
            const double drho = P.ref_src ? (q[0] - a.ref_rho) : q[0];
#pragma unroll
            for (int d = 0; d < 3; ++d) src[1 + d] = (-drho) * a.gPhi[d];
        }
        const double *tb = tend + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            double th = 0.0;
#pragma unroll
            for (int m = 0; m < Nq; ++m) {
                th += MI * sD[m + Nq * i] * sF[0][s][m + Nq * (j + Nq * k)];
                th += MI * sD[m + Nq * j] * sF[1][s][i + Nq * (m + Nq * k)];
            }
            // MODE 1 is called with alpha = beta = 1 (exact products).  The branch on
            // beta also bounds the scheduler: without it every LDS read of the five
            // variables is hoisted ahead of the arithmetic (224 VGPRs against 119,
            // occupancy 2 against 4 waves per SIMD).
            const double out =
                (beta != 0.0) ? alpha * th + beta * (MODE == 1 ? __builtin_nontemporal_load(tb + Np * s) : tb[Np * s])
                              : alpha * th;
            double tv = 0.0;
#pragma unroll
            for (int m = 0; m < Nq; ++m) {
                tv += MI * sD[m + Nq * k] * sF[2][s][i + Nq * (j + Nq * m)];
                if (m == k) tv += src[s];
            }
            T[s] = (MODE == 0 ? alpha * tv : tv) + out;
        }
    }
    __syncthreads();  // every read of sF is done: it now holds the staged state, then face fluxes
    if (act) {
#pragma unroll
        for (int s = 0; s < 5; ++s) sQ[s][n] = q[s];
        sQ[5][n] = a.Phi;
        sQ[6][n] = a.ref_p;
    }
    __syncthreads();
    // numerical fluxes of all face nodes, spread over every lane
    for (int t = n; t < 6 * Nfp; t += NT) {
        double fl5[5];
        face_flux_aff<Nq, DIFF, SMAG>(P, A, t / Nfp, t % Nfp, nb[t / Nfp], sQ, Q, f2tr, sF2o[DIFF ? t : 0],
                                      fl5);
#pragma unroll
        for (int s = 0; s < 5; ++s) sFl[t][s] = fl5[s];
    }
    __syncthreads();

    if (act) {
        // the face fluxes of this node's faces, in face order
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int fn;
            const int f = node_face_dir<Nq>(i, j, k, d, fn);
            if (f < 0) continue;
            const double w = (MODE == 0 ? alpha * MI : MI) * (ge[4 + f] * A.Wf[fn]);
#pragma unroll
            for (int s = 0; s < 5; ++s) T[s] -= w * sFl[f * Nfp + fn][s];
        }
        double *tb = tend + (int64_t)Np * 5 * e + n;
        if (MODE == 0) {
#pragma unroll
            for (int s = 0; s < 5; ++s) tb[Np * s] = T[s];
        } else {
            // the stage's state from its LDS copy (sQ is intact after the face tasks); dQ,
            // Q' and dQ' pass through once, so they go nontemporal as in k_tendency_mix (here
            // without a measurable change: 96^3 plain FP64 41.07 GDOF/s with and without)
            double *qo = Qout + (int64_t)Np * 5 * e + n;
#pragma unroll
            for (int s = 0; s < 5; ++s) {
                __builtin_nontemporal_store(sQ[s][n] + bdt * T[s], qo + Np * s);
                if (dq_store) __builtin_nontemporal_store(T[s] * rka, tb + Np * s);
            }
        }
    }
}

}  // namespace cmdg

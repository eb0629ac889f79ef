// cmdg.hip -- MI355X (gfx950) DG right-hand side and LSRK54 stepper for the dry
// AtmosModel balance law, behind the C ABI declared in include/cmdg.h.
//
// One workgroup owns one hexahedral element (Np = Nq^3 nodes, one lane per node,
// rounded up to whole 64-lane wavefronts).  Per right-hand side:
//
//   k_gradient   (diffusive models only) gradient argument G(q) staged in LDS,
//                tensor-product derivative along r/s/t, lifted central face jumps
//                evaluated by the lanes that own the face nodes, gradient-flux
//                transform; writes the gradient flux (Np, ngf, nelem).
//   k_tendency   first + second order flux at each node, contravariant flux
//                M (xi_a . F) staged in LDS, weak divergence D^T along r/s/t,
//                gravity source, Rusanov/central face fluxes evaluated by the
//                lanes owning the face nodes (periodic neighbours come from vmap+
//                on the device), and an epilogue that either writes
//                alpha*RHS + beta*tendency or performs the whole low-storage
//                Runge-Kutta stage update.
//
// Reference kernels replaced: volume_tendency! (DGModel_kernels.jl:64-548),
// dgsem_interface_tendency! (:588-901), volume_gradients! (:934-1328),
// dgsem_interface_gradients! (:1365-1651), kernel_nodal_update_auxiliary_state!
// (:1769-1825) and the LSRK update! kernel (LowStorageRungeKuttaMethod.jl:146-158).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cmdg.h"

namespace {

constexpr int NVG = 11;  // device geometry fields per node: xi (9), M, MI

struct KParams {
    double R_d, cp_d, cv_d, T_0, grav, MSLP, visc, C_smag, inv_Pr;
    double inv_cv, gammaR, kappa;
    int32_t ref_sub;        // hydrostatic reference pressure subtracted in the momentum flux
    int32_t ref_src;        // hydrostatic reference density subtracted in the gravity source
    int32_t gravity;        // Gravity() source present
    int32_t turbulence;
    int32_t with_div;
    int32_t rusanov;
    int32_t naux;
    int32_t a_Phi, a_gPhi, a_ref_rho, a_ref_p, a_Delta, a_theta_v, a_air_T;
};

struct KArrays {
    const double *vg;    // (Np, NVG, nelem)
    const double *sg;    // (5, Nfp, 6, nelem)
    const int32_t *nbr;  // (Nfp, 6, nelem): e+ * Np + vid+ (0-based)
    const int32_t *bnd;  // (6, nelem)
    const double *D;     // (Nq, Nq) column-major
};

// ------------------------------------------------------------------ physics
// Thermodynamics.jl dry relations as reached from src/Atmos/Model/moisture.jl:32-47
// and thermo_states.jl:67-77.
struct Thermo {
    double e_int, temp, p;
};

__device__ __forceinline__ Thermo thermo(const KParams &P, const double q[5], double Phi) {
    Thermo t;
    double rinv = 1.0 / q[0];
    double rek = rinv * ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]) * 0.5;
    double rep = q[0] * Phi;
    t.e_int = rinv * (q[4] - rek - rep);
    t.temp = P.T_0 + t.e_int / P.cv_d;
    t.p = (P.R_d * q[0]) * t.temp;
    return t;
}

__device__ __forceinline__ double theta_v(const KParams &P, const Thermo &t) {
    return t.temp / pow(t.p / P.MSLP, P.kappa);
}

struct Aux {
    double Phi, gPhi[3], ref_rho, ref_p, Delta;
};

__device__ __forceinline__ Aux load_aux(const KParams &P, const double *__restrict__ aux, int Np, int n, int64_t e) {
    Aux a;
    const double *b = aux + (int64_t)Np * P.naux * e + n;
    a.Phi = P.a_Phi >= 0 ? b[(int64_t)Np * P.a_Phi] : -0.0;
    if (P.a_gPhi >= 0) {
        a.gPhi[0] = b[(int64_t)Np * P.a_gPhi];
        a.gPhi[1] = b[(int64_t)Np * (P.a_gPhi + 1)];
        a.gPhi[2] = b[(int64_t)Np * (P.a_gPhi + 2)];
    } else {
        a.gPhi[0] = a.gPhi[1] = a.gPhi[2] = 0.0;
    }
    a.ref_rho = P.a_ref_rho >= 0 ? b[(int64_t)Np * P.a_ref_rho] : 0.0;
    a.ref_p = P.a_ref_p >= 0 ? b[(int64_t)Np * P.a_ref_p] : 0.0;
    a.Delta = P.a_Delta >= 0 ? b[(int64_t)Np * P.a_Delta] : 0.0;
    return a;
}

// First-order flux F[d][s] (tendencies_mass.jl:7-9, tendencies_momentum.jl:13-29,
// tendencies_energy.jl:7-21).
__device__ __forceinline__ void flux1(const KParams &P, const double q[5], double p, double ref_p, double F[3][5]) {
    double pm = P.ref_sub ? (p - ref_p) : p;
    double u[3] = {q[1] / q[0], q[2] / q[0], q[3] / q[0]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        F[d][0] = q[1 + d];
#pragma unroll
        for (int c = 0; c < 3; ++c) F[d][1 + c] = q[1 + d] * u[c] + (d == c ? pm : 0.0);
        F[d][4] = u[d] * q[4] + u[d] * p;
    }
}

// Julia eps(x) for x >= 0
__device__ __forceinline__ double ulp_above(double x) { return nextafter(x, INFINITY) - x; }

// Second-order flux: turbulence tensors (TurbulenceClosures.jl:366-404, 472-499) then
// ViscousStress / ViscousFlux / DiffEnthalpyFlux (tendencies_momentum.jl:37-42,
// tendencies_energy.jl:27-56).  gf = [grad h_tot (3), S (6), N2].
template <bool SMAG>
__device__ __forceinline__ void flux2(const KParams &P, const double q[5], const double *gf, const Aux &a, double F[3][5]) {
    const double *S = gf + 3;
    double Sm[3][3] = {{S[0], S[1], S[2]}, {S[1], S[3], S[4]}, {S[2], S[4], S[5]}};
    double tau[3][3], Dt[3];
    if (SMAG) {
        double norm2 = ((((S[0] * S[0] + 2.0 * (S[1] * S[1])) + 2.0 * (S[2] * S[2])) + S[3] * S[3]) +
                        2.0 * (S[4] * S[4])) + S[5] * S[5];
        double normS = sqrt(2.0 * norm2);
        double k[3] = {a.gPhi[0] / P.grav, a.gPhi[1] / P.grav, a.gPhi[2] / P.grav};
        double Ri = gf[9] / (normS * normS + ulp_above(normS));
        double fb2 = sqrt(fmin(fmax(1.0 - Ri * P.inv_Pr, 0.0), 1.0));
        double CD = P.C_smag * a.Delta;
        double nu0 = normS * (CD * CD) + 1e-5;
        double dotk = (nu0 * k[0] + nu0 * k[1]) + nu0 * k[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double nuv = k[d] * dotk;
            double nud = (nu0 - nuv) + nuv * fb2;
            Dt[d] = nud * P.inv_Pr;
#pragma unroll
            for (int c = 0; c < 3; ++c) tau[d][c] = (-2.0 * nud) * Sm[d][c];
        }
    } else {
        double nu = (P.turbulence == CMDG_CONSTANT_DYNAMIC_VISCOSITY) ? P.visc / q[0] : P.visc;
        double trS = (Sm[0][0] + Sm[1][1]) + Sm[2][2];
        double dv = ((2.0 * nu) / 3.0) * trS;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Dt[d] = nu * P.inv_Pr;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double v = (-2.0 * nu) * Sm[d][c];
                if (P.with_div && d == c) v = v + dv;
                tau[d][c] = v;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        F[d][0] = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) F[d][1 + c] = tau[d][c] * q[0];
        double vf = (tau[d][0] * q[1] + tau[d][1] * q[2]) + tau[d][2] * q[3];
        F[d][4] = vf + ((-Dt[d]) * gf[d]) * q[0];
    }
}

// gradient argument G = (u, h_tot[, theta_v]) (AtmosModel.jl:622-690, energy.jl:20-31)
template <bool SMAG>
__device__ __forceinline__ void grad_arg(const KParams &P, const double q[5], const Thermo &t, double thv, double G[5]) {
    double rinv = 1.0 / q[0];
    G[0] = rinv * q[1];
    G[1] = rinv * q[2];
    G[2] = rinv * q[3];
    G[3] = q[4] * rinv + P.R_d * t.temp;
    G[4] = SMAG ? thv : 0.0;
}

// compute_gradient_flux! (energy.jl:50-59, TurbulenceClosures.jl:366-376, 455-470)
template <bool SMAG>
__device__ __forceinline__ void grad_flux(const double gG[3][5], const Aux &a, double thv, double *gf) {
    gf[0] = gG[0][3];
    gf[1] = gG[1][3];
    gf[2] = gG[2][3];
    gf[3] = gG[0][0];
    gf[4] = (gG[1][0] + gG[0][1]) * 0.5;
    gf[5] = (gG[2][0] + gG[0][2]) * 0.5;
    gf[6] = gG[1][1];
    gf[7] = (gG[2][1] + gG[1][2]) * 0.5;
    gf[8] = gG[2][2];
    if (SMAG) gf[9] = ((gG[0][4] * a.gPhi[0] + gG[1][4] * a.gPhi[1]) + gG[2][4] * a.gPhi[2]) / thv;
}

// Impenetrable(FreeSlip()) wall state (bc_momentum.jl:24-43)
__device__ __forceinline__ void wall_state(const double q[5], const double nv[3], double f, double qb[5]) {
    double dot = (q[1] * nv[0] + q[2] * nv[1]) + q[3] * nv[2];
    qb[0] = q[0];
    qb[4] = q[4];
#pragma unroll
    for (int d = 0; d < 3; ++d) qb[1 + d] = q[1 + d] - (f * dot) * nv[d];
}

// Which faces hold node (i,j,k): face f and its face-node index fn (Grids.jl:581-588).
template <int Nq>
__device__ __forceinline__ int node_faces(int i, int j, int k, int f[3], int fn[3]) {
    constexpr int N = Nq - 1;
    int c = 0;
    if (i == 0) { f[c] = 0; fn[c] = j + Nq * k; ++c; }
    if (i == N) { f[c] = 1; fn[c] = j + Nq * k; ++c; }
    if (j == 0) { f[c] = 2; fn[c] = i + Nq * k; ++c; }
    if (j == N) { f[c] = 3; fn[c] = i + Nq * k; ++c; }
    if (k == 0) { f[c] = 4; fn[c] = i + Nq * j; ++c; }
    if (k == N) { f[c] = 5; fn[c] = i + Nq * j; ++c; }
    return c;
}

template <int Nq> constexpr int threads_for() { return ((Nq * Nq * Nq + 63) / 64) * 64; }

// ------------------------------------------------------------------ gradient kernel
template <int Nq, bool SMAG>
__global__ void __launch_bounds__(threads_for<Nq>())
k_gradient(KParams P, KArrays A, const double *__restrict__ Q, double *__restrict__ aux,
           double *__restrict__ gfo, double *__restrict__ Qcopy, int write_aux) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NG = SMAG ? 5 : 4, NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sG[NG][Np];
    const int64_t e = blockIdx.x;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    double q[5], thv = 0.0;
    Aux a;
    Thermo t;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        if (Qcopy) {
            double *qc = Qcopy + (int64_t)Np * 5 * e + n;
#pragma unroll
            for (int s = 0; s < 5; ++s) qc[Np * s] = q[s];
        }
        a = load_aux(P, aux, Np, n, e);
        t = thermo(P, q, a.Phi);
        thv = theta_v(P, t);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = thv;
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        double G[5];
        grad_arg<SMAG>(P, q, t, thv, G);
#pragma unroll
        for (int s = 0; s < NG; ++s) sG[s][n] = G[s];
    }
    __syncthreads();
    if (!act) return;

    const double *vg = A.vg + (int64_t)Np * NVG * e + n;
    double xi[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) xi[f] = vg[Np * f];

    // horizontal pass, then vertical pass (DGModel_kernels.jl:1007-1127, 1235-1325):
    // the gradient-flux transform is applied to each pass and the results summed
    double gG[3][5], gV[3][5];
#pragma unroll
    for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int d = 0; d < 3; ++d) gG[d][s] = gV[d][s] = 0.0;
#pragma unroll
    for (int s = 0; s < NG; ++s) {
        double g1 = 0.0, g2 = 0.0, g3 = -0.0;
#pragma unroll
        for (int m = 0; m < Nq; ++m) {
            g1 += sD[i + Nq * m] * sG[s][m + Nq * (j + Nq * k)];
            g2 += sD[j + Nq * m] * sG[s][i + Nq * (m + Nq * k)];
        }
#pragma unroll
        for (int m = 0; m < Nq; ++m) g3 += sD[k + Nq * m] * sG[s][i + Nq * (j + Nq * m)];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double v = -0.0;
            v += xi[0 + 3 * d] * g1;
            v += xi[1 + 3 * d] * g2;
            gG[d][s] = v;
            gV[d][s] = -0.0 + xi[2 + 3 * d] * g3;
        }
    }

    double gf[10], gfv[10];
    grad_flux<SMAG>(gG, a, thv, gf);
    grad_flux<SMAG>(gV, a, thv, gfv);
#pragma unroll
    for (int s = 0; s < NGF; ++s) gf[s] += gfv[s];

    // lifted face jumps: M^-1 Mf (Phi(G*) - Phi(n G-)), CentralNumericalFluxGradient
    double Gm[5];
    grad_arg<SMAG>(P, q, t, thv, Gm);
    int fl[3], fnl[3];
    const int nf = node_faces<Nq>(i, j, k, fl, fnl);
    for (int c = 0; c < nf; ++c) {
        const int f = fl[c];
        const int64_t idx = fnl[c] + (int64_t)Nfp * (f + 6 * e);
        const double *sgp = A.sg + 5 * idx;
        double nv[3] = {sgp[0], sgp[1], sgp[2]};
        const double sM = sgp[3], vMI = sgp[4];
        double Gs[3][5], nG[3][5], Gp[5];
        if (A.bnd[f + 6 * e] == 0) {
            const int32_t nb = A.nbr[idx];
            const int64_t ep = nb / Np;
            const int vp = nb - (int)ep * Np;
            double qp[5];
            const double *qpb = Q + (int64_t)Np * 5 * ep + vp;
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = qpb[Np * s];
            double Php = P.a_Phi >= 0 ? aux[(int64_t)Np * P.naux * ep + vp + (int64_t)Np * P.a_Phi] : -0.0;
            Thermo tp = thermo(P, qp, Php);
            double thp = SMAG ? theta_v(P, tp) : 0.0;
            grad_arg<SMAG>(P, qp, tp, thp, Gp);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 5; ++s) Gs[d][s] = nv[d] * (Gp[s] + Gm[s]) * 0.5;
        } else {
            double qb[5];
            wall_state(q, nv, 1.0, qb);
            Thermo tb = thermo(P, qb, a.Phi);
            double thb = SMAG ? theta_v(P, tb) : 0.0;
            grad_arg<SMAG>(P, qb, tb, thb, Gp);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 5; ++s) Gs[d][s] = nv[d] * Gp[s];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int s = 0; s < 5; ++s) nG[d][s] = nv[d] * Gm[s];
        double gfs[10], gfm[10];
        grad_flux<SMAG>(Gs, a, thv, gfs);
        grad_flux<SMAG>(nG, a, thv, gfm);
        const double w = vMI * sM;
#pragma unroll
        for (int s = 0; s < NGF; ++s) gf[s] += w * (gfs[s] - gfm[s]);
    }
    double *go = gfo + (int64_t)Np * NGF * e + n;
#pragma unroll
    for (int s = 0; s < NGF; ++s) go[Np * s] = gf[s];
}

// ------------------------------------------------------------------ tendency kernel
// MODE 0: tend = alpha*RHS + beta*tend.   MODE 1: LSRK stage, tend is dQ:
//   dQ <- RHS + dQ; Qout = Qin + bdt*dQ; dQ *= rka.
template <int Nq, bool DIFF, bool SMAG, int MODE>
__global__ void __launch_bounds__(threads_for<Nq>())
k_tendency(KParams P, KArrays A, const double *__restrict__ Q, const double *__restrict__ gfi,
           double *__restrict__ aux, double *__restrict__ tend, double alpha, double beta,
           double *__restrict__ Qout, double rka, double bdt, int write_aux) {
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NGF = SMAG ? 10 : 9;
    __shared__ double sD[Nq * Nq];
    __shared__ double sF[3][5][Np];
    const int64_t e = blockIdx.x;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);

    double q[5], gf[10], F1[3][5], F2[3][5], src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    Aux a;
    Thermo t;
    double MI = 0.0;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        a = load_aux(P, aux, Np, n, e);
        t = thermo(P, q, a.Phi);
        if (write_aux) {
            double *ab = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) ab[(int64_t)Np * P.a_theta_v] = theta_v(P, t);
            if (P.a_air_T >= 0) ab[(int64_t)Np * P.a_air_T] = t.temp;
        }
        flux1(P, q, t.p, a.ref_p, F1);
        if (DIFF) {
            const double *gb = gfi + (int64_t)Np * NGF * e + n;
#pragma unroll
            for (int s = 0; s < NGF; ++s) gf[s] = gb[Np * s];
            flux2<SMAG>(P, q, gf, a, F2);
        }
        const double *vg = A.vg + (int64_t)Np * NVG * e + n;
        const double M = vg[Np * 9];
        MI = vg[Np * 10];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const double x1 = vg[Np * ax], x2 = vg[Np * (ax + 3)], x3 = vg[Np * (ax + 6)];
#pragma unroll
            for (int s = 0; s < 5; ++s) {
                double Fx = F1[0][s], Fy = F1[1][s], Fz = F1[2][s];
                if (DIFF) { Fx += F2[0][s]; Fy += F2[1][s]; Fz += F2[2][s]; }
                sF[ax][s][n] = M * ((x1 * Fx + x2 * Fy) + x3 * Fz);
            }
        }
        if (P.gravity) {
            // Gravity() source (tendencies_momentum.jl:60-72)
            double drho = P.ref_src ? (q[0] - a.ref_rho) : q[0];
#pragma unroll
            for (int d = 0; d < 3; ++d) src[1 + d] = (-drho) * a.gPhi[d];
        }
    }
    __syncthreads();
    if (!act) return;

    // horizontal pass (written as alpha*T + beta*tendency), then vertical pass with
    // the source added at its own level (DGModel_kernels.jl:281-306, 453-545)
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

    // face fluxes of the faces this node lies on (NumericalFluxes.jl:219-340, 670-715)
    int fl[3], fnl[3];
    const int nf = node_faces<Nq>(i, j, k, fl, fnl);
    double csm = 0.0;
    if (nf > 0 && P.rusanov) csm = sqrt(P.gammaR * t.temp);
    for (int c = 0; c < nf; ++c) {
        const int f = fl[c];
        const int64_t idx = fnl[c] + (int64_t)Nfp * (f + 6 * e);
        const double *sgp = A.sg + 5 * idx;
        const double nv[3] = {sgp[0], sgp[1], sgp[2]};
        const double w = (MODE == 0 ? alpha * sgp[4] : sgp[4]) * sgp[3];  // alpha * vMI * sM
        const bool wall = A.bnd[f + 6 * e] != 0;
        double qp[5], Fp[3][5], Phip, refpp;
        Aux ap;
        if (!wall) {
            const int32_t nb = A.nbr[idx];
            const int64_t ep = nb / Np;
            const int vp = nb - (int)ep * Np;
            const double *qpb = Q + (int64_t)Np * 5 * ep + vp;
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = qpb[Np * s];
            ap = load_aux(P, aux, Np, vp, ep);
            Phip = ap.Phi;
            refpp = ap.ref_p;
        } else {
            wall_state(q, nv, 2.0, qp);
            ap = a;
            Phip = a.Phi;
            refpp = a.ref_p;
        }
        Thermo tp = thermo(P, qp, Phip);
        flux1(P, qp, tp.p, refpp, Fp);
        const double h[3] = {nv[0] * 0.5, nv[1] * 0.5, nv[2] * 0.5};
        double flux[5];
#pragma unroll
        for (int s = 0; s < 5; ++s)
            flux[s] = -0.0 + (((F1[0][s] + Fp[0][s]) * h[0] + (F1[1][s] + Fp[1][s]) * h[1]) + (F1[2][s] + Fp[2][s]) * h[2]);
        if (P.rusanov) {
            const double rim = 1.0 / q[0], rip = 1.0 / qp[0];
            const double uNm = fabs((nv[0] * (rim * q[1]) + nv[1] * (rim * q[2])) + nv[2] * (rim * q[3]));
            const double uNp = fabs((nv[0] * (rip * qp[1]) + nv[1] * (rip * qp[2])) + nv[2] * (rip * qp[3]));
            const double mw = fmax(uNm + csm, uNp + sqrt(P.gammaR * tp.temp));
#pragma unroll
            for (int s = 0; s < 5; ++s) flux[s] += (mw * (q[s] - qp[s])) * 0.5;
        }
        if (DIFF && !wall) {
            const int32_t nb = A.nbr[idx];
            const int64_t ep = nb / Np;
            const int vp = nb - (int)ep * Np;
            double gfp[10];
            const double *gpb = gfi + (int64_t)Np * NGF * ep + vp;
#pragma unroll
            for (int s = 0; s < NGF; ++s) gfp[s] = gpb[Np * s];
            flux2<SMAG>(P, qp, gfp, ap, Fp);
#pragma unroll
            for (int s = 0; s < 5; ++s)
                flux[s] += ((F2[0][s] + Fp[0][s]) * h[0] + (F2[1][s] + Fp[1][s]) * h[1]) + (F2[2][s] + Fp[2][s]) * h[2];
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) T[s] -= w * flux[s];
    }

    if (MODE == 0) {
#pragma unroll
        for (int s = 0; s < 5; ++s) tb[Np * s] = T[s];
    } else {
        double *qo = Qout + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            qo[Np * s] = q[s] + bdt * T[s];
            tb[Np * s] = T[s] * rka;
        }
    }
}

// update! (LowStorageRungeKuttaMethod.jl:146-158)
__global__ void k_lsrk_update(double *__restrict__ dQ, double *__restrict__ Q, int64_t n, double rka, double bdt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const double d = dQ[i];
        Q[i] += bdt * d;
        dQ[i] = d * rka;
    }
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// LSRK54CarpenterKennedy coefficients (LowStorageRungeKuttaMethod.jl:302-324)
const double RKA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                       -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double RKB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                       1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                       2277821191437.0 / 14882151754819.0};

}  // namespace

// ====================================================================== context
struct cmdg_ctx {
    int N, Nq, Np, Nfp;
    int64_t nelem;
    int ngf;
    bool diffusive, smag;
    KParams P;
    KArrays A;
    hipStream_t stream;
    double *d_vg = nullptr, *d_sg = nullptr, *d_D = nullptr, *d_aux = nullptr, *d_gf = nullptr, *d_Qs = nullptr;
    int32_t *d_nbr = nullptr, *d_bnd = nullptr;
    std::string err;
    // optional per-kernel event timing (cmdg_set_kernel_timing)
    bool timing = false;
    struct Rec {
        int kind;
        hipEvent_t start, stop;
    };
    std::vector<Rec> recs;
};

namespace {

int fail(cmdg_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int check_hip(cmdg_ctx *c, hipError_t e, const char *what) {
    if (e == hipSuccess) return CMDG_OK;
    return fail(c, e == hipErrorOutOfMemory ? CMDG_ENOMEM : CMDG_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

// kernel kinds reported by cmdg_kernel_timing
enum { KIND_GRADIENT = 0, KIND_TENDENCY = 1, KIND_UPDATE = 2, KIND_COPY = 3, NKIND = 4 };

struct TimedLaunch {
    cmdg_ctx *c;
    int kind;
    hipEvent_t stop = nullptr;
    TimedLaunch(cmdg_ctx *c_, int kind_) : c(c_), kind(kind_) {
        if (!c->timing) return;
        hipEvent_t start;
        if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) { c->timing = false; return; }
        (void)hipEventRecord(start, c->stream);
        c->recs.push_back({kind, start, stop});
    }
    ~TimedLaunch() {
        if (stop) (void)hipEventRecord(stop, c->stream);
    }
};

template <class T> int upload(cmdg_ctx *c, T **dst, const T *src, size_t count) {
    hipError_t e = hipMalloc((void **)dst, count * sizeof(T));
    if (e != hipSuccess) return check_hip(c, e, "hipMalloc");
    if (src) {
        e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return check_hip(c, e, "hipMemcpy");
    }
    return CMDG_OK;
}

template <int Nq> int launch_gradient_impl(cmdg_ctx *c, const double *Q, double *Qcopy, int write_aux) {
    constexpr int NT = threads_for<Nq>();
    dim3 grid((unsigned)c->nelem), block(NT);
    TimedLaunch tl(c, KIND_GRADIENT);
    if (c->smag)
        hipLaunchKernelGGL((k_gradient<Nq, true>), grid, block, 0, c->stream, c->P, c->A, Q, c->d_aux, c->d_gf, Qcopy, write_aux);
    else
        hipLaunchKernelGGL((k_gradient<Nq, false>), grid, block, 0, c->stream, c->P, c->A, Q, c->d_aux, c->d_gf, Qcopy, write_aux);
    return check_hip(c, hipGetLastError(), "gradient launch");
}

template <int Nq>
int launch_tendency_impl(cmdg_ctx *c, const double *Q, double *tend, double alpha, double beta, int mode,
                         double *Qout, double rka, double bdt, int write_aux) {
    constexpr int NT = threads_for<Nq>();
    dim3 grid((unsigned)c->nelem), block(NT);
    TimedLaunch tl(c, KIND_TENDENCY);
#define CMDG_TEND(DIFF, SMAG, MODE)                                                                            \
    hipLaunchKernelGGL((k_tendency<Nq, DIFF, SMAG, MODE>), grid, block, 0, c->stream, c->P, c->A, Q, c->d_gf, \
                       c->d_aux, tend, alpha, beta, Qout, rka, bdt, write_aux)
    if (mode == 0) {
        if (!c->diffusive) CMDG_TEND(false, false, 0);
        else if (c->smag) CMDG_TEND(true, true, 0);
        else CMDG_TEND(true, false, 0);
    } else {
        if (!c->diffusive) CMDG_TEND(false, false, 1);
        else if (c->smag) CMDG_TEND(true, true, 1);
        else CMDG_TEND(true, false, 1);
    }
#undef CMDG_TEND
    return check_hip(c, hipGetLastError(), "tendency launch");
}

int launch_gradient(cmdg_ctx *c, const double *Q, double *Qcopy, int write_aux) {
    switch (c->Nq) {
    case 4: return launch_gradient_impl<4>(c, Q, Qcopy, write_aux);
    case 5: return launch_gradient_impl<5>(c, Q, Qcopy, write_aux);
    default: return fail(c, CMDG_EUNSUPPORTED, "polynomial order not compiled");
    }
}

int launch_tendency(cmdg_ctx *c, const double *Q, double *tend, double alpha, double beta, int mode, double *Qout,
                    double rka, double bdt, int write_aux) {
    switch (c->Nq) {
    case 4: return launch_tendency_impl<4>(c, Q, tend, alpha, beta, mode, Qout, rka, bdt, write_aux);
    case 5: return launch_tendency_impl<5>(c, Q, tend, alpha, beta, mode, Qout, rka, bdt, write_aux);
    default: return fail(c, CMDG_EUNSUPPORTED, "polynomial order not compiled");
    }
}

// dQ/dt into tend (mode 0) with the auxiliary refresh of DGModel.jl:110-116
int launch_rhs(cmdg_ctx *c, const double *Q, double *tend, double alpha, double beta) {
    int rc;
    if (c->diffusive && (rc = launch_gradient(c, Q, nullptr, 1))) return rc;
    return launch_tendency(c, Q, tend, alpha, beta, 0, nullptr, 0.0, 0.0, c->diffusive ? 0 : 1);
}

int launch_copy(cmdg_ctx *c, double *dst, const double *src, int64_t n) {
    TimedLaunch tl(c, KIND_COPY);
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, c->stream, dst, src, n);
    return check_hip(c, hipGetLastError(), "copy launch");
}

}  // namespace

extern "C" {

const char *cmdg_version(void) { return "cmdg-mi355x 0.1 (gfx950)"; }

const char *cmdg_last_error(cmdg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cmdg_create(const cmdg_desc *d, void *stream, cmdg_ctx **out) {
    if (!out) return CMDG_EINVAL;
    *out = nullptr;
    if (!d) return CMDG_EINVAL;
    cmdg_ctx *c = new (std::nothrow) cmdg_ctx();
    if (!c) return CMDG_ENOMEM;
    *out = c;
    if (d->N != 3 && d->N != 4) return fail(c, CMDG_EUNSUPPORTED, "only N = 3 and N = 4 are compiled");
    if (d->nelem <= 0 || d->nvgeo < 11 || d->naux < 0) return fail(c, CMDG_EINVAL, "bad sizes");
    if (!d->vgeo || !d->sgeo || !d->vmap_minus || !d->vmap_plus || !d->elemtobndy || !d->D)
        return fail(c, CMDG_EINVAL, "missing geometry or connectivity");
    if (d->precision != CMDG_FP64) return fail(c, CMDG_EUNSUPPORTED, "mixed precision not built in this library");
    if (d->turbulence == CMDG_LEARNED_CLOSURE) return fail(c, CMDG_EUNSUPPORTED, "learned closure not built");
    if (d->numerical_flux != CMDG_RUSANOV && d->numerical_flux != CMDG_CENTRAL)
        return fail(c, CMDG_EUNSUPPORTED, "numerical flux");
    auto in_aux = [&](int32_t off, int width) { return off < 0 || off + width <= d->naux; };
    if (!in_aux(d->aux_coord, 3) || !in_aux(d->aux_Phi, 1) || !in_aux(d->aux_gradPhi, 3) || !in_aux(d->aux_ref_rho, 1) ||
        !in_aux(d->aux_ref_p, 1) || !in_aux(d->aux_Delta, 1) || !in_aux(d->aux_theta_v, 1) || !in_aux(d->aux_air_T, 1))
        return fail(c, CMDG_EINVAL, "auxiliary offset out of range");
    if (d->naux > 0 && !d->aux_init) return fail(c, CMDG_EINVAL, "missing initial auxiliary state");
    if (d->orientation == CMDG_FLAT_ORIENTATION && (d->aux_Phi < 0 || d->aux_gradPhi < 0))
        return fail(c, CMDG_EINVAL, "FlatOrientation needs Phi and grad Phi in aux");
    if (d->ref_state == CMDG_HYDROSTATIC_STATE && (d->aux_ref_rho < 0 || d->aux_ref_p < 0))
        return fail(c, CMDG_EINVAL, "HydrostaticState needs ref rho and p in aux");
    if (d->turbulence == CMDG_SMAGORINSKY_LILLY && (d->aux_Delta < 0 || d->aux_theta_v < 0))
        return fail(c, CMDG_EINVAL, "Smagorinsky-Lilly needs Delta and theta_v in aux");

    c->N = d->N;
    c->Nq = d->N + 1;
    c->Np = c->Nq * c->Nq * c->Nq;
    c->Nfp = c->Nq * c->Nq;
    c->nelem = d->nelem;
    if ((double)c->nelem * c->Np >= 2147483647.0) return fail(c, CMDG_EUNSUPPORTED, "mesh too large for int32 node ids");
    c->smag = d->turbulence == CMDG_SMAGORINSKY_LILLY;
    c->diffusive = c->smag || d->viscosity != 0.0;
    c->ngf = c->smag ? 10 : 9;
    c->stream = (hipStream_t)stream;

    KParams &P = c->P;
    P.R_d = d->R_d; P.cp_d = d->cp_d; P.cv_d = d->cv_d; P.T_0 = d->T_0; P.grav = d->grav; P.MSLP = d->MSLP;
    P.visc = d->viscosity; P.C_smag = d->C_smag; P.inv_Pr = d->inv_Pr_turb;
    P.inv_cv = 1.0 / d->cv_d;
    P.gammaR = (d->cp_d / d->cv_d) * d->R_d;
    P.kappa = d->R_d / d->cp_d;
    const bool hyd = d->ref_state == CMDG_HYDROSTATIC_STATE && d->subtract_off;
    P.ref_sub = hyd;
    P.ref_src = hyd;
    P.gravity = d->gravity_source;
    P.turbulence = d->turbulence;
    P.with_div = d->with_divergence;
    P.rusanov = d->numerical_flux == CMDG_RUSANOV;
    P.naux = d->naux;
    P.a_Phi = d->orientation == CMDG_FLAT_ORIENTATION ? d->aux_Phi : -1;
    P.a_gPhi = d->orientation == CMDG_FLAT_ORIENTATION ? d->aux_gradPhi : -1;
    P.a_ref_rho = d->ref_state == CMDG_HYDROSTATIC_STATE ? d->aux_ref_rho : -1;
    P.a_ref_p = d->ref_state == CMDG_HYDROSTATIC_STATE ? d->aux_ref_p : -1;
    P.a_Delta = d->aux_Delta;
    P.a_theta_v = d->aux_theta_v;
    P.a_air_T = d->aux_air_T;

    const int Np = c->Np, Nfp = c->Nfp;
    const int64_t E = c->nelem;
    // geometry: xi (9), M, MI per node
    std::vector<double> vg((size_t)Np * NVG * E);
    for (int64_t e = 0; e < E; ++e)
        for (int f = 0; f < NVG; ++f)
            std::memcpy(&vg[(size_t)Np * (f + NVG * e)], d->vgeo + (size_t)Np * (f + (int64_t)d->nvgeo * e), Np * sizeof(double));
    // connectivity: minus side must be the element's own face nodes (Grids.jl:615)
    std::vector<int32_t> nbr((size_t)Nfp * 6 * E), bnd((size_t)6 * E);
    const int Nq = c->Nq, NN = c->N;
    for (int64_t e = 0; e < E; ++e)
        for (int f = 0; f < 6; ++f) {
            bnd[f + 6 * e] = (int32_t)d->elemtobndy[f + 6 * e];
            for (int fn = 0; fn < Nfp; ++fn) {
                int a = fn % Nq, b = fn / Nq, vid;
                switch (f) {
                case 0: vid = 0 + Nq * (a + Nq * b); break;
                case 1: vid = NN + Nq * (a + Nq * b); break;
                case 2: vid = a + Nq * (0 + Nq * b); break;
                case 3: vid = a + Nq * (NN + Nq * b); break;
                case 4: vid = a + Nq * (b + Nq * 0); break;
                default: vid = a + Nq * (b + Nq * NN); break;
                }
                const int64_t id = fn + (int64_t)Nfp * (f + 6 * e);
                if (d->vmap_minus[id] - 1 != (int64_t)Np * e + vid)
                    return fail(c, CMDG_EINVAL, "vmap_minus is not the element's own face nodes");
                const int64_t p = d->vmap_plus[id] - 1;
                if (p < 0 || p >= (int64_t)Np * E) return fail(c, CMDG_EINVAL, "vmap_plus out of range");
                nbr[id] = (int32_t)p;
            }
        }
    int rc;
    if ((rc = upload(c, &c->d_vg, vg.data(), vg.size()))) return rc;
    if ((rc = upload(c, &c->d_sg, d->sgeo, (size_t)5 * Nfp * 6 * E))) return rc;
    if ((rc = upload(c, &c->d_nbr, nbr.data(), nbr.size()))) return rc;
    if ((rc = upload(c, &c->d_bnd, bnd.data(), bnd.size()))) return rc;
    if ((rc = upload(c, &c->d_D, d->D, (size_t)Nq * Nq))) return rc;
    if ((rc = upload(c, &c->d_aux, d->aux_init, (size_t)Np * (d->naux > 0 ? d->naux : 1) * E))) return rc;
    if ((rc = upload<double>(c, &c->d_gf, nullptr, (size_t)Np * c->ngf * E))) return rc;
    if ((rc = upload<double>(c, &c->d_Qs, nullptr, (size_t)Np * 5 * E))) return rc;
    if ((rc = check_hip(c, hipMemset(c->d_gf, 0, (size_t)Np * c->ngf * E * sizeof(double)), "hipMemset"))) return rc;
    c->A.vg = c->d_vg;
    c->A.sg = c->d_sg;
    c->A.nbr = c->d_nbr;
    c->A.bnd = c->d_bnd;
    c->A.D = c->d_D;
    return CMDG_OK;
}

int cmdg_destroy(cmdg_ctx *c) {
    if (!c) return CMDG_OK;
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    (void)hipFree(c->d_vg);
    (void)hipFree(c->d_sg);
    (void)hipFree(c->d_nbr);
    (void)hipFree(c->d_bnd);
    (void)hipFree(c->d_D);
    (void)hipFree(c->d_aux);
    (void)hipFree(c->d_gf);
    (void)hipFree(c->d_Qs);
    delete c;
    return CMDG_OK;
}

int cmdg_rhs(cmdg_ctx *c, double *tend, const double *Q, double t, double alpha, double beta) {
    (void)t;
    if (!c || !tend || !Q) return c ? fail(c, CMDG_EINVAL, "null pointer") : CMDG_EINVAL;
    if (tend == Q) return fail(c, CMDG_EINVAL, "tendency aliases the state");
    return launch_rhs(c, Q, tend, alpha, beta);
}

int cmdg_lsrk_update(double *dQ, double *Q, int64_t n, double rka, double rkb, double dt, void *stream) {
    if (!dQ || !Q || n < 0) return CMDG_EINVAL;
    if (n == 0) return CMDG_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_lsrk_update, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dQ, Q, n, rka, rkb * dt);
    return hipGetLastError() == hipSuccess ? CMDG_OK : CMDG_EHIP;
}

int cmdg_lsrk_stage(cmdg_ctx *c, double *Q, double *dQ, double t_stage, double dt, double rka_next, double rkb) {
    (void)t_stage;
    if (!c || !Q || !dQ) return c ? fail(c, CMDG_EINVAL, "null pointer") : CMDG_EINVAL;
    int rc = launch_rhs(c, Q, dQ, 1.0, 1.0);
    if (rc) return rc;
    {
        TimedLaunch tl(c, KIND_UPDATE);
        rc = cmdg_lsrk_update(dQ, Q, (int64_t)c->Np * 5 * c->nelem, rka_next, rkb, dt, c->stream);
    }
    return rc ? fail(c, rc, "update launch") : CMDG_OK;
}

int cmdg_lsrk54_step(cmdg_ctx *c, double *Q, double *dQ, double t, double dt) {
    (void)t;
    if (!c || !Q || !dQ) return c ? fail(c, CMDG_EINVAL, "null pointer") : CMDG_EINVAL;
    // Ping-pong between the caller's Q (C) and the context buffer (S) so that the
    // fused stage update never overwrites a state that neighbouring elements still
    // read: S <- C (done by the first gradient pass, or a copy), then the stages
    // run S->C, C->S, S->C, C->S, S->C and the result lands in the caller's Q.
    double *C = Q, *S = c->d_Qs;
    const int64_t n = (int64_t)c->Np * 5 * c->nelem;
    int rc;
    if (!c->diffusive && (rc = launch_copy(c, S, C, n))) return rc;
    double *src = S, *dst = C;
    for (int s = 0; s < 5; ++s) {
        const double a = RKA[(s + 1) % 5], bdt = RKB[s] * dt;
        const int wa = (s == 4);  // aux keeps the state of the last stage, as after dostep!
        if (c->diffusive && (rc = launch_gradient(c, s == 0 ? C : src, s == 0 ? S : nullptr, wa))) return rc;
        if ((rc = launch_tendency(c, src, dQ, 1.0, 1.0, 1, dst, a, bdt, c->diffusive ? 0 : wa))) return rc;
        double *tmp = src;
        src = dst;
        dst = tmp;
    }
    return CMDG_OK;
}

int cmdg_aux(cmdg_ctx *c, double **aux) {
    if (!c || !aux) return CMDG_EINVAL;
    *aux = c->d_aux;
    return CMDG_OK;
}

int cmdg_gradient_flux(cmdg_ctx *c, double **gf, int32_t *ngf) {
    if (!c || !gf) return CMDG_EINVAL;
    *gf = c->d_gf;
    if (ngf) *ngf = c->ngf;
    return CMDG_OK;
}

int cmdg_mlp_set_weights(cmdg_ctx *c, const uint16_t *, const float *, const uint16_t *, const float *,
                         const uint16_t *, const float *, const float *, const float *) {
    return fail(c, CMDG_EUNSUPPORTED, "learned closure not built");
}

int cmdg_set_kernel_timing(cmdg_ctx *c, int enable) {
    if (!c) return CMDG_EINVAL;
    c->timing = enable != 0;
    return CMDG_OK;
}

int cmdg_kernel_timing(cmdg_ctx *c, double *ms_total, int64_t *launches, int32_t nkinds) {
    if (!c || !ms_total || !launches || nkinds < 1) return CMDG_EINVAL;
    for (int k = 0; k < nkinds; ++k) { ms_total[k] = 0.0; launches[k] = 0; }
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (rc) return rc;
    for (auto &r : c->recs) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess && r.kind < nkinds) {
            ms_total[r.kind] += ms;
            launches[r.kind] += 1;
        }
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    c->recs.clear();
    return CMDG_OK;
}

int cmdg_sync(cmdg_ctx *c) {
    if (!c) return CMDG_EINVAL;
    return check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

}  // extern "C"

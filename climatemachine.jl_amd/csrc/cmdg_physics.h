// cmdg_physics.h -- device-side dry-atmosphere physics shared by the DG kernels.
// Expression trees follow the reference operation by operation (the FP64 path is
// compiled without FMA contraction) so that FP64 results match the CPU oracle to
// rounding in the last place.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/cmdg.h"

namespace cmdg {

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
    int64_t nelem;
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

// x^y for finite x > 0 (a pressure ratio) in FP64, within 2.1 ulp of the exact value
// for x in [1e-3, 4] (checked against quad precision): x = 2^k m, m in [sqrt(1/2),
// sqrt(2)); log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172, by its odd series to
// s^23; y log x carried as a hi + lo pair; exp(y log x) = 2^i e^r, |r| <= ln2 / 2, by the
// Taylor series to r^14.  About 45 FP64 operations where the general pow spends
// several times that on range and special-case handling.
__device__ __forceinline__ double pow_pos(double x, double y) {
    constexpr double LN2_HI = 0.6931471804855391, LN2_LO = 7.440617110012397e-11, INV_LN2 = 1.4426950408889634;
    int k;
    double m = frexp(x, &k);
    if (m < 0.7071067811865476) {
        m *= 2.0;
        k -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double a = 1.0 / 23.0;
    a = fma(a, s2, 1.0 / 21.0);
    a = fma(a, s2, 1.0 / 19.0);
    a = fma(a, s2, 1.0 / 17.0);
    a = fma(a, s2, 1.0 / 15.0);
    a = fma(a, s2, 1.0 / 13.0);
    a = fma(a, s2, 1.0 / 11.0);
    a = fma(a, s2, 1.0 / 9.0);
    a = fma(a, s2, 1.0 / 7.0);
    a = fma(a, s2, 1.0 / 5.0);
    a = fma(a, s2, 1.0 / 3.0);
    const double logm = fma(2.0 * s * s2, a, 2.0 * s);
    const double kd = (double)k;
    const double lhi = fma(kd, LN2_HI, logm);
    const double llo = fma(kd, LN2_LO, kd * LN2_HI - (lhi - logm));
    const double t = y * lhi;
    const double tlo = fma(y, lhi, -t) + y * llo;
    const double id = rint(t * INV_LN2);
    const double r = (fma(-id, LN2_HI, t) - id * LN2_LO) + tlo;
    double e = 1.0 / 87178291200.0;  // 1 / 14!
    e = fma(e, r, 1.0 / 6227020800.0);
    e = fma(e, r, 1.0 / 479001600.0);
    e = fma(e, r, 1.0 / 39916800.0);
    e = fma(e, r, 1.0 / 3628800.0);
    e = fma(e, r, 1.0 / 362880.0);
    e = fma(e, r, 1.0 / 40320.0);
    e = fma(e, r, 1.0 / 5040.0);
    e = fma(e, r, 1.0 / 720.0);
    e = fma(e, r, 1.0 / 120.0);
    e = fma(e, r, 1.0 / 24.0);
    e = fma(e, r, 1.0 / 6.0);
    e = fma(e, r, 0.5);
    e = fma(e, r * r, r) + 1.0;
    return ldexp(e, (int)id);
}

// virtual potential temperature T / (p / MSLP)^(R_d / cp_d) (moisture.jl:58-69)
__device__ __forceinline__ double theta_v(const KParams &P, const Thermo &t) {
    return t.temp / pow_pos(t.p / P.MSLP, P.kappa);
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

// Component d (0, 1, 2) of three values, chosen among values rather than among array
// elements: `d == 0 ? a[0] : ...` on arrays is a choice of address, which keeps the
// array in private (scratch) memory when d is only known at run time.
template <typename T>
__device__ __forceinline__ T pick3(int d, T a, T b, T c) {
    return d == 0 ? a : (d == 1 ? b : c);
}

// The face of direction d (0, 1, 2) that node (i, j, k) lies on -- 2d or 2d + 1 -- or
// -1, with its face-node index fn.  Looping d = 0, 1, 2 (unrolled) visits a node's
// faces in ascending face order, the order of the reference's face loop, with compile-time
// indices: per-direction arrays then stay in registers instead of private memory.
template <int Nq>
__device__ __forceinline__ int node_face_dir(int i, int j, int k, int d, int &fn) {
    constexpr int N = Nq - 1;
    const int id = pick3(d, i, j, k);
    fn = pick3(d, j + Nq * k, i + Nq * k, i + Nq * j);
    return id == 0 ? 2 * d : (id == N ? 2 * d + 1 : -1);
}

template <int Nq> constexpr int threads_for() { return ((Nq * Nq * Nq + 63) / 64) * 64; }

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD one contiguous
// range of elements so that spatial neighbours share that XCD's L2.  The grid has
// 8 * ceil(nelem / 8) workgroups; the caller drops indices >= nelem.
constexpr int NXCD = 8;
__device__ __forceinline__ int64_t xcd_element(int64_t b, int64_t nelem) {
    const int64_t per = (nelem + NXCD - 1) / NXCD;
    return (b % NXCD) * per + b / NXCD;
}
inline int64_t xcd_grid(int64_t nelem) { return ((nelem + NXCD - 1) / NXCD) * NXCD; }

// Row d of the first-order flux (F[d][0..4]); u = q / rho, pm = p or p - p_ref.
__device__ __forceinline__ void flux1_row(const double q[5], const double u[3], double p, double pm, int d,
                                          double F[5]) {
    const double qd = pick3(d, q[1], q[2], q[3]);
    const double ud = pick3(d, u[0], u[1], u[2]);
    F[0] = qd;
    F[1] = qd * u[0] + (d == 0 ? pm : 0.0);
    F[2] = qd * u[1] + (d == 1 ? pm : 0.0);
    F[3] = qd * u[2] + (d == 2 ? pm : 0.0);
    F[4] = ud * q[4] + ud * p;
}

// Diagonal of the kinematic viscosity tensor (TurbulenceClosures.jl:378-404, 472-499).
template <bool SMAG>
__device__ __forceinline__ void viscosity(const KParams &P, double rho, const double *gf, const double gPhi[3],
                                          double Delta, double nu[3]) {
    if (SMAG) {
        const double *S = gf + 3;
        double norm2 = ((((S[0] * S[0] + 2.0 * (S[1] * S[1])) + 2.0 * (S[2] * S[2])) + S[3] * S[3]) +
                        2.0 * (S[4] * S[4])) + S[5] * S[5];
        double normS = sqrt(2.0 * norm2);
        double k[3] = {gPhi[0] / P.grav, gPhi[1] / P.grav, gPhi[2] / P.grav};
        double Ri = gf[9] / (normS * normS + ulp_above(normS));
        double fb2 = sqrt(fmin(fmax(1.0 - Ri * P.inv_Pr, 0.0), 1.0));
        double CD = P.C_smag * Delta;
        double nu0 = normS * (CD * CD) + 1e-5;
        double dotk = (nu0 * k[0] + nu0 * k[1]) + nu0 * k[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double nuv = k[d] * dotk;
            nu[d] = (nu0 - nuv) + nuv * fb2;
        }
    } else {
        double v = (P.turbulence == CMDG_CONSTANT_DYNAMIC_VISCOSITY) ? P.visc / rho : P.visc;
        nu[0] = nu[1] = nu[2] = v;
    }
}

// Row d of the second-order flux (tendencies_momentum.jl:37-42, tendencies_energy.jl:27-56):
// tau[d][c] = -2 nu_d S[d][c] (+ 2 nu / 3 tr S on the diagonal WithDivergence).
template <bool SMAG>
__device__ __forceinline__ void flux2_row(const KParams &P, const double q[5], const double *gf, const double nu[3],
                                          int d, double F[5]) {
    const double *S = gf + 3;
    const double S0 = pick3(d, S[0], S[1], S[2]);
    const double S1 = pick3(d, S[1], S[3], S[4]);
    const double S2 = pick3(d, S[2], S[4], S[5]);
    const double nud = pick3(d, nu[0], nu[1], nu[2]);
    double t0 = (-2.0 * nud) * S0, t1 = (-2.0 * nud) * S1, t2 = (-2.0 * nud) * S2;
    if (!SMAG && P.with_div) {
        const double dv = ((2.0 * nud) / 3.0) * ((S[0] + S[3]) + S[5]);
        if (d == 0) t0 = t0 + dv;
        if (d == 1) t1 = t1 + dv;
        if (d == 2) t2 = t2 + dv;
    }
    F[0] = 0.0;
    F[1] = t0 * q[0];
    F[2] = t1 * q[0];
    F[3] = t2 * q[0];
    const double vf = (t0 * q[1] + t1 * q[2]) + t2 * q[3];
    const double gd = pick3(d, gf[0], gf[1], gf[2]);
    F[4] = vf + ((-(nud * P.inv_Pr)) * gd) * q[0];
}

// Gradient-flux transform of n (x) X with n = +-e_d (only row d of the 3 x 5 argument is non-zero).
template <bool SMAG>
__device__ __forceinline__ void grad_flux_row(int d, const double X[5], const double gPhi[3], double thv, double *gf) {
    double g[3][5];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 5; ++s) g[r][s] = (r == d) ? X[s] : 0.0;
    Aux a;
    a.gPhi[0] = gPhi[0];
    a.gPhi[1] = gPhi[1];
    a.gPhi[2] = gPhi[2];
    grad_flux<SMAG>(g, a, thv, gf);
}

}  // namespace cmdg

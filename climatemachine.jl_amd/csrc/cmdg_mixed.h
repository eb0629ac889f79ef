// cmdg_mixed.h -- mixed-precision DG kernels for affine bricks (CMDG_MIXED).
//
// Arithmetic split: the state Q, the low-storage register dQ and the stage update
// stay FP64; the thermodynamic state (internal energy, temperature, pressure) is
// evaluated in FP64 from the FP64 state, and the hydrostatic-reference
// subtractions p - p_ref (pressure perturbation in the momentum flux) and
// rho - rho_ref (gravity source) are formed in FP64 before rounding.  Without a
// subtracted reference the momentum flux carries p minus an element constant.  Fluxes,
// the gradient argument, tensor-product derivatives, face fluxes, the
// Smagorinsky-Lilly closure and the gradient flux are FP32; the gradient flux is
// stored in FP32 in stress form (the closure folded in: see StressForm).  The Rusanov
// jump q- - q+ is formed in FP64.  Metric terms are computed once in FP64 and rounded
// on use.
#pragma once
#include "cmdg_affine.h"
#include "cmdg_closure.h"

namespace cmdg {

struct MixParams {
    float R_d, gammaR, kappa, inv_MSLP, grav, C_smag, inv_Pr, visc;
    int32_t dyn_visc, with_div, ref_sub, ref_src, gravity, rusanov;
};

// First-order flux row along axis d, every operation rounded (no contraction): the
// tendency kernel evaluates an element's own row at a face node in its volume phase and
// the neighbour evaluates the same row from the trace; both get the same bits.
__device__ __forceinline__ void flux1_row_f(const float q[5], const float u[3], float p, float pm, int d, float F[5]) {
#pragma clang fp contract(off)
    const float qd = pick3(d, q[1], q[2], q[3]);
    const float ud = pick3(d, u[0], u[1], u[2]);
    F[0] = qd;
    F[1] = qd * u[0] + (d == 0 ? pm : 0.0f);
    F[2] = qd * u[1] + (d == 1 ? pm : 0.0f);
    F[3] = qd * u[2] + (d == 2 ? pm : 0.0f);
    F[4] = ud * (q[4] + p);
}

// strain_norm_f and viscosity_f round every operation (no contraction): the viscosity
// of a node does not depend on where the compiler inlines them.
__device__ __forceinline__ float strain_norm_f(const float *gf) {
#pragma clang fp contract(off)
    const float *S = gf + 3;
    const float norm2 = S[0] * S[0] + S[3] * S[3] + S[5] * S[5] + 2.0f * (S[1] * S[1] + S[2] * S[2] + S[4] * S[4]);
    return sqrtf(2.0f * norm2);
}

// Smagorinsky-Lilly viscosity; LEARN takes nu_0 from gf[10] (the learned closure).
// BUOY false: no gravitational potential, so k = 0 and the buoyancy factor drops out:
// nu_d = (nu_0 - 0) + 0 f_b = nu_0 exactly, formed without N^2, Ri or f_b.
template <bool SMAG, bool LEARN = false, bool BUOY = true>
__device__ __forceinline__ void viscosity_f(const MixParams &P, float rho, const float *gf, const float gPhi[3],
                                            float Delta, float nu[3]) {
#pragma clang fp contract(off)
    if (SMAG && !BUOY) {
        const float CD = P.C_smag * Delta;
        const float nu0 = LEARN ? gf[10] : strain_norm_f(gf) * (CD * CD) + 1e-5f;
        nu[0] = nu[1] = nu[2] = nu0;
    } else if (SMAG) {
        const float normS = strain_norm_f(gf);
        const float ig = 1.0f / P.grav;
        const float k[3] = {gPhi[0] * ig, gPhi[1] * ig, gPhi[2] * ig};
        const float Ri = gf[9] / (normS * normS + (nextafterf(normS, INFINITY) - normS));
        const float fb2 = sqrtf(fminf(fmaxf(1.0f - Ri * P.inv_Pr, 0.0f), 1.0f));
        const float CD = P.C_smag * Delta;
        const float nu0 = LEARN ? gf[10] : normS * (CD * CD) + 1e-5f;
        const float dotk = nu0 * (k[0] + k[1] + k[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float nuv = k[d] * dotk;
            nu[d] = (nu0 - nuv) + nuv * fb2;
        }
    } else {
        const float v = P.dyn_visc ? P.visc / rho : P.visc;
        nu[0] = nu[1] = nu[2] = v;
    }
}

// Mixed-precision gradient flux in stress form: what the second-order flux row needs,
// formed once per node by the gradient kernel (the kinematic stress of
// TurbulenceClosures.jl:378-404, 472-499 and the enthalpy diffusivity times the
// enthalpy gradient, tendencies_momentum.jl:37-42, tendencies_energy.jl:27-56).
// Planes (Np, NGF, nelem):
//   tau  row d = -2 nu_d S_d. (plus (2 nu / 3) tr S on the diagonal for a constant
//        viscosity with divergence).  An isotropic viscosity (no buoyancy correction:
//        constant viscosity, or Smagorinsky-Lilly without a gravitational potential)
//        makes it symmetric: 6 planes xx xy xz yy yz zz; otherwise 9 planes, row-major;
//   e    e_d = (nu_d / Pr_t) dh_tot/dx_d, 3 planes;
//   nu   the viscosity: 1 plane (isotropic) or 3 (nu_d); only diagnostics read it.
// Row d of the second-order flux is then (0, tau_d. rho, tau_d . rho u - e_d rho): the
// tendency reads the first NREAD planes, its own for the volume and face rows and the
// face neighbour's (4 values at the shared node) for the other side of a face.
template <bool ISO> struct StressForm {
    static constexpr int NTAU = ISO ? 6 : 9, E = NTAU, NU = NTAU + 3, NREAD = NTAU + 3;
    static constexpr int NGF = NTAU + 3 + (ISO ? 1 : 3);
    // plane of tau[d][c]
    static __device__ __forceinline__ int tau(int d, int c) {
        if (!ISO) return 3 * d + c;
        const int a = d < c ? d : c, b = d < c ? c : d;
        return a == 0 ? b : (a == 1 ? 2 + b : 5);
    }
};

// Stress form of one node from its gradient flux gf = [grad h_tot (3), S (6), N2, ...]
// and viscosity nu[3], plane p stored at out[p * stride] as soon as it is formed; every
// operation rounded (no contraction), in the order the row below and the face
// neighbour's row use them.
template <bool SMAG, bool ISO>
__device__ __forceinline__ void stress_form_f(const MixParams &P, const float *gf, const float nu[3], float *out,
                                              int stride) {
#pragma clang fp contract(off)
    using L = StressForm<ISO>;
    const float *S = gf + 3;
    const float trS = S[0] + S[3] + S[5];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float nud = ISO ? nu[0] : nu[d];  // isotropic: the three are equal
        const float m2 = -2.0f * nud;
        const float dv = (2.0f * nud / 3.0f) * trS;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (ISO && c < d) continue;
            const int a = d < c ? d : c, b = d < c ? c : d;
            float t = m2 * S[a == 0 ? b : (a == 1 ? 2 + b : 5)];
            if (!SMAG && P.with_div && c == d) t += dv;
            out[stride * L::tau(d, c)] = t;
        }
        out[stride * (L::E + d)] = (nud * P.inv_Pr) * gf[d];
    }
    if (ISO) {
        out[stride * L::NU] = nu[0];
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) out[stride * (L::NU + d)] = nu[d];
    }
}

// Row d of the second-order flux from that row of the stress form (t = tau_d., e = e_d)
// and the FP32 state.
__device__ __forceinline__ void flux2_row_stress(const float t[3], float e, const float q[5], float F[5]) {
#pragma clang fp contract(off)
    F[0] = 0.0f;
    F[1] = t[0] * q[0];
    F[2] = t[1] * q[0];
    F[3] = t[2] * q[0];
    F[4] = t[0] * q[1] + t[1] * q[2] + t[2] * q[3] - e * q[0];
}

// gradient flux of the 3 x 5 gradient (energy.jl:50-59, TurbulenceClosures.jl:455-470);
// BUOY false: no gravitational potential, N^2 = 0
template <bool SMAG, bool BUOY = true>
__device__ __forceinline__ void grad_flux_f(const float g[3][5], const float gPhi[3], float thv, float *gf) {
#pragma clang fp contract(fast)
    gf[0] = g[0][3];
    gf[1] = g[1][3];
    gf[2] = g[2][3];
    gf[3] = g[0][0];
    gf[4] = 0.5f * (g[1][0] + g[0][1]);
    gf[5] = 0.5f * (g[2][0] + g[0][2]);
    gf[6] = g[1][1];
    gf[7] = 0.5f * (g[2][1] + g[1][2]);
    gf[8] = g[2][2];
    if (SMAG) gf[9] = BUOY ? (g[0][4] * gPhi[0] + g[1][4] * gPhi[1] + g[2][4] * gPhi[2]) / thv : 0.0f;
}

// FP64 reciprocal of a positive normal x: the hardware estimate refined by two Newton
// steps (an error at most about one FP64 ulp, against a correctly rounded quotient; the
// FP32 values the mixed kernels round from it are the same but where an FP64 result lies
// within that ulp of an FP32 rounding boundary).  Five FP64 instructions where the IEEE
// division takes eleven (its range scaling and fix-ups serve zero, denormal and huge
// operands, which a density or a mass-matrix entry is not).  Every mixed kernel takes
// its 1/rho and M^-1 from here, so both sides of a face get the same bits.
__device__ __forceinline__ double recip_d(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

struct ThermoD {
    double temp, p;
};
__device__ __forceinline__ ThermoD thermo_d(const KParams &P, const double q[5], double Phi) {
#pragma clang fp contract(fast)
    const double rinv = recip_d(q[0]);
    const double e_int = rinv * (q[4] - 0.5 * rinv * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) - q[0] * Phi);
    ThermoD t;
    t.temp = P.T_0 + e_int * P.inv_cv;
    t.p = P.R_d * q[0] * t.temp;
    return t;
}

// gradient argument (u, h_tot, theta_v) in FP32 from the FP64 state and thermodynamics;
// theta_v = T (p / MSLP)^-kappa as T 2^(-kappa log2(p / MSLP)) on the hardware log2 and
// exp2 (a few FP32 ulp: below the FP32 rounding of theta_v that its gradient sees).
// theta_v enters only N^2 = grad theta_v . grad Phi / theta_v: without a gravitational
// potential (buoy false: grad Phi = 0, N^2 = 0 exactly) it is not evaluated and the
// argument holds 1 in its place.
template <bool SMAG>
__device__ __forceinline__ void grad_arg_f(const KParams &P, const MixParams &M, const double q[5], const ThermoD &t,
                                           bool buoy, float G[5]) {
#pragma clang fp contract(fast)
    const double rinv = recip_d(q[0]);
    G[0] = (float)(rinv * q[1]);
    G[1] = (float)(rinv * q[2]);
    G[2] = (float)(rinv * q[3]);
    G[3] = (float)(q[4] * rinv + P.R_d * t.temp);
    G[4] = !SMAG ? 0.0f
           : buoy ? (float)t.temp * __builtin_amdgcn_exp2f(-M.kappa * __builtin_amdgcn_logf((float)t.p * M.inv_MSLP))
                  : 1.0f;
}

// ------------------------------------------------------------------ gradient (mixed)
// Seven waves per SIMD at Nq = 5 (a cap of 72 registers, no spills): the derivative and
// the lifted jumps run a field at a time and the stress form is stored as it is formed,
// which keeps the learned variants at 66-70 registers and the Smagorinsky-Lilly ones at
// 49-57; a whole-gradient-then-lifting order needed 79-94 (5-6 waves) and spilled under an
// 80 cap with a gravitational potential.  Eight waves (a cap of 64, which the LDS would
// allow the learned block) spill 9 registers without and 20 with a potential.  Measured on one box (whole order -> field
// order): 96^3 chosen form 56.8 -> 57.1 GDOF/s (gradient 4036 -> 3975 us), mixed
// Smagorinsky-Lilly 60.6 -> 61.9; 25^3 50.1 -> 50.4 (86.7 -> 84.5 us).  Without a
// gravitational potential the kernel also skips N^2, Ri and the buoyancy factor (exactly
// nu_d = nu_0 there) and theta_v's jumps: 96^3 gradient 3972 -> 3913 us, bit for bit.
// BUOY: the context has a gravitational potential (FlatOrientation), so Smagorinsky-Lilly's
// N^2 = grad theta_v . grad Phi / theta_v needs theta_v and its gradient; without one
// (grad Phi = 0, N^2 = 0 exactly) the kernel evaluates neither.
template <int Nq, bool SMAG, bool LEARN, bool BUOY>
__global__ void __launch_bounds__(threads_for<Nq>()) __attribute__((amdgpu_waves_per_eu(Nq == 5 ? 7 : 5)))
k_gradient_mix(KParams P, MixParams MP, AffArrays A, ElemSet ES, const double *__restrict__ Q,
               double *__restrict__ aux, float *__restrict__ gfo, int write_aux, const MlpDev *__restrict__ mlp) {
#pragma clang fp contract(fast)
    static_assert(!LEARN || SMAG, "the learned closure keeps the Smagorinsky-Lilly structure");
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NG = SMAG ? 5 : 4;
    constexpr bool ISO = !(SMAG && BUOY);
    using L = StressForm<ISO>;
    constexpr int NT = threads_for<Nq>();
    __shared__ float sD[Nq * Nq];
    __shared__ float sG[NG][Np];
    __shared__ float sJ[6 * Nfp][5];
    __shared__ uint16_t sX[LEARN ? (NT / 64) * MLP_LDS_PER_WAVE : 1];
    const int64_t e = set_element(ES, blockIdx.x);
    if (e < 0) return;
    const int64_t ce = static_class(A, e);
    constexpr bool buoy = SMAG && BUOY;
    __shared__ int32_t nb[6];  // the element's face neighbours
    if (threadIdx.x < 6) nb[threadIdx.x] = A.nbe[6 * e + threadIdx.x];
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = (float)A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double *ge = A.geo + (int64_t)AFF_GEO * e;
    double q[5];
    float G[5], gPhi[3] = {0.0f, 0.0f, 0.0f}, thv = 1.0f;
    double Phi = -0.0;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        const double *ab = A.sa + (int64_t)Np * P.naux * ce + n;
        if (P.a_Phi >= 0) Phi = ab[(int64_t)Np * P.a_Phi];
        if (A.gphi32) {
            const float *gp = A.gphi32 + (int64_t)Np * 3 * ce + n;
            gPhi[0] = gp[0];
            gPhi[1] = gp[Np];
            gPhi[2] = gp[2 * Np];
        }
        const ThermoD t = thermo_d(P, q, Phi);
        grad_arg_f<SMAG>(P, MP, q, t, buoy, G);
        thv = SMAG ? G[4] : 1.0f;
        if (write_aux) {
            double *aw = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) aw[(int64_t)Np * P.a_theta_v] = t.temp / pow_pos(t.p / P.MSLP, P.kappa);
            if (P.a_air_T >= 0) aw[(int64_t)Np * P.a_air_T] = t.temp;
        }
#pragma unroll
        for (int s = 0; s < NG; ++s) sG[s][n] = G[s];
    }
    __syncthreads();
    // lifted central jumps n (G* - G-) of all face nodes, spread over every lane
    for (int t = n; t < 6 * Nfp; t += NT) {
        const int f = t / Nfp, fn = t % Nfp, d = f >> 1, vn = face_node<Nq>(f, fn);
        const float nd = (f & 1) ? 1.0f : -1.0f;
        const int32_t ne = nb[f];
        float Gm[5], Gp[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) Gm[s] = s < NG ? sG[s][vn] : 0.0f;
        float jump[5];
        if (ne >= 0) {
            const int fo = f ^ 1;
            // the neighbour's state at the shared node, from its volume (this pass writes no
            // state, so every block reads the stage's state)
            double qp[5];
            const int64_t qb = (int64_t)Np * 5 * ne + face_node<Nq>(fo, fn);
#pragma unroll
            for (int s = 0; s < 5; ++s) qp[s] = Q[qb + Np * s];
            const double Php = A.aux_cont ? own_aux<Nq>(P, A.sa, ce, vn, P.a_Phi, -0.0)
                                          : A.str[tri<Nfp>(ne, fo, 0, TR_S, fn)];
            grad_arg_f<SMAG>(P, MP, qp, thermo_d(P, qp, Php), buoy, Gp);
#pragma unroll
            for (int s = 0; s < 5; ++s) jump[s] = nd * (0.5f * (Gp[s] - Gm[s]));
        } else {
            double qb[5];
#pragma unroll
            for (int s = 0; s < 5; ++s) qb[s] = Q[(int64_t)Np * (s + 5 * e) + vn];
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (c == d) qb[1 + c] = 0.0;  // the gradient wall state removes the normal momentum
            grad_arg_f<SMAG>(P, MP, qb, thermo_d(P, qb, own_aux<Nq>(P, A.sa, ce, vn, P.a_Phi, -0.0)), buoy, Gp);
#pragma unroll
            for (int s = 0; s < 5; ++s) jump[s] = nd * (Gp[s] - Gm[s]);
        }
#pragma unroll
        for (int s = 0; s < 5; ++s)
            if (s < 4 || buoy) sJ[t][s] = jump[s];
    }
    __syncthreads();

    const float xi0 = (float)ge[0], xi1 = (float)ge[1], xi2 = (float)ge[2];
    float g[3][5] = {};
    float gf[11];
    if (act) {
        // per field: the volume derivative along r, s, t, then the lifted central jumps of
        // this node's faces, M^-1 Mf n (G* - G-) (a field at a time keeps fewer values live)
        const float MI = (float)recip_d(ge[3] * A.W[n]);
        int fl[3];  // this node's face slot along each axis (-1: interior along it)
        float wl[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int fn;
            const int f = node_face_dir<Nq>(i, j, k, d, fn);
            fl[d] = f < 0 ? -1 : f * Nfp + fn;
            wl[d] = f < 0 ? 0.0f : MI * (float)(ge[4 + f] * A.Wf[fn]);
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            float g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
            if (s < NG && (s < 4 || buoy)) {
#pragma unroll
                for (int m = 0; m < Nq; ++m) {
                    g1 += sD[i + Nq * m] * sG[s][m + Nq * (j + Nq * k)];
                    g2 += sD[j + Nq * m] * sG[s][i + Nq * (m + Nq * k)];
                    g3 += sD[k + Nq * m] * sG[s][i + Nq * (j + Nq * m)];
                }
            }
            g[0][s] = xi0 * g1;
            g[1][s] = xi1 * g2;
            g[2][s] = xi2 * g3;
            if (s < 4 || buoy)  // theta_v's gradient only where N^2 needs it
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (fl[d] >= 0) g[d][s] += wl[d] * sJ[fl[d]][s];
        }
        grad_flux_f<SMAG, buoy>(g, gPhi, thv, gf);
    }
    if (LEARN) {
        // the network runs on the assembled gradient (volume + lifted face jumps):
        // it is nonlinear, so it cannot be split over the two contributions
        float x[MLP_IN];
        mlp_inputs(mlp, g, (float)ge[10], x);
        if (!act)
#pragma unroll
            for (int c = 0; c < MLP_IN; ++c) x[c] = 0.0f;
        const float y = mlp_wave(mlp, x, sX);
        gf[10] = mlp_nu0(y, (float)ge[10], strain_norm_f(gf), MP.visc > 0.0f ? MP.visc : 1e-5f);
    }
    if (act) {
        // the stress form of the node's second-order flux: the tendency forms its rows,
        // and the face neighbours' rows, from it
        float nu[3];
        viscosity_f<SMAG, LEARN, buoy>(MP, (float)q[0], gf, gPhi, (float)ge[10], nu);
        stress_form_f<SMAG, ISO>(MP, gf, nu, gfo + (int64_t)Np * L::NGF * e + n, Np);
    }
}

// ------------------------------------------------------------------ tendency (mixed)
// Numerical flux at face node fn of face f of element e.  The own side arrives
// precomputed by the volume phase (own[0..3]: entries 1-4 of the first-order flux row
// along the face's axis, own[4..7]: the second-order row, own[8]: |u_d| + c; the row's
// mass entry is the FP32 rounding of rho u_d, taken here from the staged state); the
// neighbour's side is
// evaluated here from its FP64 volume state (thermodynamics, p - p_ref in FP64) and, for
// the second-order row, from its stress form at the shared node (4 of its planes) and
// its FP32-rounded state: the bits the neighbour's own volume phase forms for that row.
// The jump q- - q+ is formed in FP64 from the own state, staged in LDS with Phi and
// p_ref (sQ rows 0-4, 5, 6).
template <int Nq, bool DIFF, bool ISO>
__device__ __forceinline__ void face_flux_mix(const KParams &P, const MixParams &MP, const AffArrays &A, int f, int fn,
                                              int32_t ne, const double (*sQ)[Nq * Nq * Nq],
                                              const double *__restrict__ Q, const float *__restrict__ gfi,
                                              const float own[9], double p_e, float flux[5]) {
#pragma clang fp contract(fast)
    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    const int d = f >> 1, vn = face_node<Nq>(f, fn);
    const float h = (f & 1) ? 0.5f : -0.5f;
    const bool wall = ne < 0;
    const int fo = f ^ 1;
    double q[5], qp[5], Php, refpp;
    float t2p[3] = {0.0f, 0.0f, 0.0f}, e2p = 0.0f;  // the neighbour's stress-form row
#pragma unroll
    for (int s = 0; s < 5; ++s) q[s] = sQ[s][vn];
    if (!wall) {
        // the neighbour's state and stress-form row at the shared node, loaded together
        const double *qb = Q + (int64_t)Np * 5 * ne + face_node<Nq>(fo, fn);
#pragma unroll
        for (int s = 0; s < 5; ++s) qp[s] = qb[Np * s];
        if (DIFF) {
            using L = StressForm<ISO>;
            const float *gn = gfi + (int64_t)Np * L::NGF * ne + face_node<Nq>(fo, fn);
#pragma unroll
            for (int c = 0; c < 3; ++c) t2p[c] = gn[Np * L::tau(d, c)];
            e2p = gn[Np * (L::E + d)];
        }
        if (A.aux_cont) {
            Php = sQ[5][vn];
            refpp = sQ[6][vn];
        } else {
            Php = A.str[tri<Nfp>(ne, fo, 0, TR_S, fn)];
            refpp = A.str[tri<Nfp>(ne, fo, 1, TR_S, fn)];
        }
    } else {
#pragma unroll
        for (int s = 0; s < 5; ++s) qp[s] = q[s];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c == d) qp[1 + c] = -q[1 + c];  // reflect the normal momentum
        Php = sQ[5][vn];
        refpp = sQ[6][vn];
    }
    const ThermoD tp = thermo_d(P, qp, Php);
    float qpf[5], upf[3], jump[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        qpf[s] = (float)qp[s];
        jump[s] = (float)(q[s] - qp[s]);
    }
    const float rip = 1.0f / qpf[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) upf[c] = qpf[1 + c] * rip;
    float Fp[5];
    flux1_row_f(qpf, upf, (float)tp.p, (float)(MP.ref_sub ? tp.p - refpp : tp.p - p_e), d, Fp);
    flux[0] = (pick3(d, (float)q[1], (float)q[2], (float)q[3]) + Fp[0]) * h;
#pragma unroll
    for (int s = 1; s < 5; ++s) flux[s] = (own[s - 1] + Fp[s]) * h;
    if (MP.rusanov) {
        const float upd = pick3(d, upf[0], upf[1], upf[2]);
        const float mw = 0.5f * fmaxf(own[8], fabsf(upd) + sqrtf(MP.gammaR * (float)tp.temp));
#pragma unroll
        for (int s = 0; s < 5; ++s) flux[s] += mw * jump[s];
    }
    if (DIFF && !wall) {
        float F2p[5];
        flux2_row_stress(t2p, e2p, qpf, F2p);
#pragma unroll
        for (int v = 0; v < 4; ++v) flux[1 + v] += (own[4 + v] + F2p[1 + v]) * h;
    }
}

// MODE 0: tend = alpha*RHS + beta*tend.   MODE 1: LSRK stage (tend is dQ):
//   dQ <- RHS + dQ; Qout = Qin + bdt*dQ; dQ *= rka.  Qout is another buffer than Q:
//   blocks read their neighbours' volume state.  In MODE 1 beta = 0 says dQ holds zeros
//   (the previous stage scaled it by rka = 0) and is not read; dq_store = 0 leaves the
//   scaled dQ unstored (this stage's rka is 0 and the next stage gets beta = 0).
// DIFF: second-order fluxes from the gradient flux gfi in stress form (ISO: its
// isotropic layout); the closure (constant viscosity, Smagorinsky-Lilly or the learned
// nu_0) is already folded into it.
// Six waves per SIMD at Nq = 5: the register cap this asks for (80) spills nothing there
// (at Nq = 4 it would, so that order keeps the compiler's choice).
template <int Nq, bool DIFF, bool ISO, int MODE>
__global__ void __launch_bounds__(threads_for<Nq>()) __attribute__((amdgpu_waves_per_eu(Nq == 5 ? 6 : 1)))
k_tendency_mix(KParams P, MixParams MP, AffArrays A, ElemSet ES, const double *__restrict__ Q,
               const float *__restrict__ gfi, double *__restrict__ aux, double *__restrict__ tend, double alpha,
               double beta, double *__restrict__ Qout, double rka, double bdt, int write_aux, int dq_store) {

    constexpr int Np = Nq * Nq * Nq, Nfp = Nq * Nq;
    constexpr int NT = threads_for<Nq>();
    using L = StressForm<ISO>;
    __shared__ float sD[Nq * Nq];
    __shared__ float sF[3][5][Np];
    // per face node: the own side's flux rows and wave speed from the volume phase
    // (first-order row entries 1-4, second-order row 4, |u_d| + c), then the face flux
    // (first 5); 9 floats keep the block at 13 KB of LDS, 12 blocks (6 waves/SIMD) a CU
    __shared__ float sFc[6 * Nfp][9];
    const int64_t e = set_element(ES, blockIdx.x);
    if (e < 0) return;
    const int64_t ce = static_class(A, e);
    __shared__ int32_t nb[6];  // the element's face neighbours
    if (threadIdx.x < 6) nb[threadIdx.x] = A.nbe[6 * e + threadIdx.x];
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = (float)A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double *ge = A.geo + (int64_t)AFF_GEO * e;

    double q[5], Phi = -0.0, ref_p = 0.0, p_e = 0.0;
    if (!MP.ref_sub) {
        // Without a subtracted reference pressure the momentum flux carries p
        // relative to an element constant, the pressure at the element's first
        // node.  A constant flux is annihilated exactly by the weak-form volume term
        // plus this element's face terms (the LGL summation-by-parts identity on an
        // affine element), so the offset only removes rounding: the FP32 flux holds
        // the in-element pressure variation, not p itself.
        const double *q0 = Q + (int64_t)Np * 5 * e;
        const double qe[5] = {q0[0], q0[Np], q0[2 * Np], q0[3 * Np], q0[4 * Np]};
        const double Phe = P.a_Phi >= 0 ? A.sa[(int64_t)Np * P.naux * ce + (int64_t)Np * P.a_Phi] : -0.0;
        p_e = thermo_d(P, qe, Phe).p;
    }
    float qf[5], uf[3], sf[L::NREAD], gPhi[3] = {0.0f, 0.0f, 0.0f}, pf = 0.0f, pmf = 0.0f, cm = 0.0f, drho = 0.0f;
    float MI = 0.0f;
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        const double *ab = A.sa + (int64_t)Np * P.naux * ce + n;
        if (P.a_Phi >= 0) Phi = ab[(int64_t)Np * P.a_Phi];
        if (P.a_ref_p >= 0) ref_p = ab[(int64_t)Np * P.a_ref_p];
        double ref_rho = P.a_ref_rho >= 0 ? ab[(int64_t)Np * P.a_ref_rho] : 0.0;
        if (A.gphi32) {
            const float *gp = A.gphi32 + (int64_t)Np * 3 * ce + n;
            gPhi[0] = gp[0];
            gPhi[1] = gp[Np];
            gPhi[2] = gp[2 * Np];
        }
        const ThermoD t = thermo_d(P, q, Phi);
        if (write_aux) {
            double *aw = aux + (int64_t)Np * P.naux * e + n;
            if (P.a_theta_v >= 0) aw[(int64_t)Np * P.a_theta_v] = t.temp / pow_pos(t.p / P.MSLP, P.kappa);
            if (P.a_air_T >= 0) aw[(int64_t)Np * P.a_air_T] = t.temp;
        }
        pf = (float)t.p;
        pmf = (float)(MP.ref_sub ? t.p - ref_p : t.p - p_e);
        drho = (float)(MP.ref_src ? q[0] - ref_rho : q[0]);
        cm = MP.rusanov ? sqrtf(MP.gammaR * (float)t.temp) : 0.0f;
#pragma unroll
        for (int s = 0; s < 5; ++s) qf[s] = (float)q[s];
        const float ri = 1.0f / qf[0];
        uf[0] = qf[1] * ri;
        uf[1] = qf[2] * ri;
        uf[2] = qf[3] * ri;
        if (DIFF) {
            const float *gb = gfi + (int64_t)Np * L::NGF * e + n;
#pragma unroll
            for (int s = 0; s < L::NREAD; ++s) sf[s] = gb[Np * s];
        }
        const double Md = ge[3] * A.W[n];
        MI = (float)recip_d(Md);
        const float M = (float)Md;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            float F[5], F2[5];
            flux1_row_f(qf, uf, pf, pmf, ax, F);
            int fn;
            const int f = node_face_dir<Nq>(i, j, k, ax, fn);
            if (f >= 0) {
#pragma unroll
                for (int s = 1; s < 5; ++s) sFc[f * Nfp + fn][s - 1] = F[s];
                sFc[f * Nfp + fn][8] = fabsf(uf[ax]) + cm;
            }
            if (DIFF) {
                const float t[3] = {sf[L::tau(ax, 0)], sf[L::tau(ax, 1)], sf[L::tau(ax, 2)]};
                flux2_row_stress(t, sf[L::E + ax], qf, F2);
                if (f >= 0)
#pragma unroll
                    for (int v = 0; v < TR_F2; ++v) sFc[f * Nfp + fn][4 + v] = F2[1 + v];
#pragma unroll
                for (int s = 0; s < 5; ++s) F[s] += F2[s];
            }
            const float x = M * (float)ge[ax];
#pragma unroll
            for (int s = 0; s < 5; ++s) sF[ax][s][n] = x * F[s];
        }
    }
    __syncthreads();
    // volume divergence and gravity source of this node, while sF holds the fluxes
    float T[5];
    if (act) {
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < Nq; ++m) {
                acc += sD[m + Nq * i] * sF[0][s][m + Nq * (j + Nq * k)];
                acc += sD[m + Nq * j] * sF[1][s][i + Nq * (m + Nq * k)];
                acc += sD[m + Nq * k] * sF[2][s][i + Nq * (j + Nq * m)];
            }
            T[s] = MI * acc;
        }
        if (MP.gravity) {
#pragma unroll
            for (int d = 0; d < 3; ++d) T[1 + d] -= drho * gPhi[d];
        }
    }
    __syncthreads();
    // sF is free: it takes the element's state, Phi and p_ref (FP64) for the face tasks,
    // which would otherwise read them again from memory after the cache has let them go
    double(*sQ)[Np] = reinterpret_cast<double(*)[Np]>(&sF[0][0][0]);
    static_assert(7 * Np * sizeof(double) <= sizeof(sF), "state, Phi and p_ref fit in sF");
    if (act) {
#pragma unroll
        for (int s = 0; s < 5; ++s) sQ[s][n] = q[s];
        sQ[5][n] = Phi;
        sQ[6][n] = ref_p;
    }
    __syncthreads();
    // numerical fluxes of all face nodes, spread over every lane
    for (int t = n; t < 6 * Nfp; t += NT) {
        float own[9], fl5[5];
#pragma unroll
        for (int c = 0; c < 9; ++c) own[c] = sFc[t][c];
        face_flux_mix<Nq, DIFF, ISO>(P, MP, A, t / Nfp, t % Nfp, nb[t / Nfp], sQ, Q, gfi, own, p_e, fl5);
#pragma unroll
        for (int s = 0; s < 5; ++s) sFc[t][s] = fl5[s];
    }
    // the stage's register dQ, loaded before the barrier so that the wait overlaps it.
    // dQ and the stage's outputs Q' and dQ' pass through once (the next reader comes a
    // whole grid later), so they go nontemporal and leave the cache to the face
    // neighbours' lines (96^3: 5590 -> 5513 us a launch; nontemporal stress-form stores
    // in the gradient kernel cost it 3815 -> 3924 us and were left out)
    double dq[5] = {};
    double *tb = tend + (int64_t)Np * 5 * e + n;
    if (MODE == 1 && act && beta != 0.0)
#pragma unroll
        for (int s = 0; s < 5; ++s) dq[s] = __builtin_nontemporal_load(tb + Np * s);
    __syncthreads();

    if (act) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int fn;
            const int f = node_face_dir<Nq>(i, j, k, d, fn);
            if (f < 0) continue;
            const float w = MI * (float)(ge[4 + f] * A.Wf[fn]);
#pragma unroll
            for (int s = 0; s < 5; ++s) T[s] -= w * sFc[f * Nfp + fn][s];
        }
        if (MODE == 0) {
#pragma unroll
            for (int s = 0; s < 5; ++s)
                tb[Np * s] = (beta != 0.0) ? alpha * (double)T[s] + beta * tb[Np * s] : alpha * (double)T[s];
        } else {
            // the stage's state from its LDS copy (sQ is intact after the face tasks): no
            // register holds it across them
            double *qo = Qout + (int64_t)Np * 5 * e + n;
#pragma unroll
            for (int s = 0; s < 5; ++s) {
                const double dd = (double)T[s] + dq[s];
                __builtin_nontemporal_store(sQ[s][n] + bdt * dd, qo + Np * s);
                if (dq_store) __builtin_nontemporal_store(dd * rka, tb + Np * s);
            }
        }
    }
}

}  // namespace cmdg

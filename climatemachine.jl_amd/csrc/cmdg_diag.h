// cmdg_diag.h -- driver-loop diagnostics on the device: local Courant numbers and
// their maximum (kernel_min_neighbor_distance!, src/Numerics/Mesh/Grids.jl:1228-1334;
// kernel_local_courant!, src/Numerics/DGMethods/DGModel_kernels.jl:3028-3095;
// courant, src/Numerics/DGMethods/SpaceDiscretization.jl:307-365; the Atmos
// advective/nondiffusive/diffusive Courant numbers, src/Atmos/Model/courant.jl:11-73)
// and the mass-weighted sum of one state (weightedsum,
// src/Arrays/MPIStateArrays.jl:655-674, double-double accumulation).
#pragma once
#include "cmdg_physics.h"
#include "cmdg_mixed.h"

namespace cmdg {

enum { COURANT_ADVECTIVE = 0, COURANT_NONDIFFUSIVE = 1, COURANT_DIFFUSIVE = 2 };
enum { DIR_EVERY = 0, DIR_HORIZONTAL = 1, DIR_VERTICAL = 2 };

// MIX: the gradient flux is the FP32 stress form of a mixed-precision context (gf32, ISO
// its isotropic layout), and the diffusive Courant number takes the viscosity the mixed
// gradient kernel formed (with the learned closure, from the learned nu_0).
template <int Nq, bool SMAG, bool MIX = false, bool ISO = true>
__global__ void __launch_bounds__(threads_for<Nq>())
k_courant(KParams P, MixParams MP, const double *__restrict__ Q, const double *__restrict__ aux,
          const double *__restrict__ gf, const float *__restrict__ gf32, int ngf, int64_t nelem, double dt, int kind,
          int direction, int coord, double *__restrict__ partial) {
    constexpr int Np = Nq * Nq * Nq, NT = threads_for<Nq>();
    __shared__ double sx[3][Np];
    __shared__ double red[NT];
    const int64_t e = blockIdx.x;
    const int n = threadIdx.x;
    const bool act = n < Np;
    const double *ab = aux + (int64_t)Np * P.naux * e + n;
    if (act)
        for (int d = 0; d < 3; ++d) sx[d][n] = ab[(int64_t)Np * (coord + d)];
    __syncthreads();
    double c = -INFINITY;
    if (act) {
        const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
        const bool dirs[3] = {direction != DIR_VERTICAL, direction != DIR_VERTICAL, direction != DIR_HORIZONTAL};
        double md = INFINITY;
        const int ijk[3] = {i, j, k}, stride[3] = {1, Nq, Nq * Nq};
        for (int a = 0; a < 3; ++a) {
            if (!dirs[a]) continue;
            for (int s = -1; s <= 1; s += 2) {
                const int m = ijk[a] + s;
                if (m < 0 || m >= Nq) continue;
                const int o = n + s * stride[a];
                const double dx = sx[0][n] - sx[0][o], dy = sx[1][n] - sx[1][o], dz = sx[2][n] - sx[2][o];
                md = fmin(md, sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        double q[5];
        const double *qb = Q + (int64_t)Np * 5 * e + n;
        for (int s = 0; s < 5; ++s) q[s] = qb[Np * s];
        const Aux a = load_aux(P, aux, Np, n, e);
        const double kh[3] = {a.gPhi[0] / P.grav, a.gPhi[1] / P.grav, a.gPhi[2] / P.grav};
        if (kind == COURANT_DIFFUSIVE) {
            double nu[3];
            if (MIX && SMAG) {
                // the viscosity the gradient kernel formed from the last RHS's gradient
                using L = StressForm<ISO>;
                const float *gb = gf32 + (int64_t)Np * L::NGF * e + n;
                for (int d = 0; d < 3; ++d) nu[d] = (double)gb[(int64_t)Np * (ISO ? L::NU : L::NU + d)];
            } else if (MIX) {
                // a constant viscosity (mu / rho: of this state)
                float nf[3];
                const float g0[1] = {0.0f}, gp[3] = {0.0f, 0.0f, 0.0f};
                viscosity_f<false>(MP, (float)q[0], g0, gp, 0.0f, nf);
                for (int d = 0; d < 3; ++d) nu[d] = (double)nf[d];
            } else {
                double g[11] = {};
                const double *gb = gf + (int64_t)Np * ngf * e + n;
                for (int s = 0; s < ngf && s < 11; ++s) g[s] = gb[(int64_t)Np * s];
                viscosity<SMAG>(P, q[0], g, a.gPhi, a.Delta, nu);
            }
            double nn;
            if (!SMAG) {
                nn = nu[0];  // a scalar viscosity in every direction
            } else {
                const double dk = nu[0] * kh[0] + nu[1] * kh[1] + nu[2] * kh[2];
                if (direction == DIR_VERTICAL) {
                    nn = dk;
                } else {
                    double v[3] = {nu[0], nu[1], nu[2]};
                    if (direction == DIR_HORIZONTAL)
                        for (int d = 0; d < 3; ++d) v[d] -= dk * kh[d];
                    nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                }
            }
            c = dt * nn / (md * md);
        } else {
            double m3[3] = {q[1], q[2], q[3]};
            const double mk = m3[0] * kh[0] + m3[1] * kh[1] + m3[2] * kh[2];
            double nu_;
            if (direction == DIR_VERTICAL) {
                nu_ = fabs(mk) / q[0];
            } else {
                if (direction == DIR_HORIZONTAL)
                    for (int d = 0; d < 3; ++d) m3[d] -= mk * kh[d];
                const double rinv = 1.0 / q[0];
                const double u0 = m3[0] * rinv, u1 = m3[1] * rinv, u2 = m3[2] * rinv;
                nu_ = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            }
            if (kind == COURANT_NONDIFFUSIVE) {
                const Thermo t = thermo(P, q, a.Phi);
                nu_ += sqrt(P.gammaR * t.temp);
            }
            c = dt * nu_ / md;
        }
    }
    red[n] = c;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (n < w) red[n] = fmax(red[n], red[n + w]);
        __syncthreads();
    }
    if (n == 0) partial[e] = red[0];
}

__global__ void __launch_bounds__(256) k_max_reduce(const double *__restrict__ partial, int64_t n, double *__restrict__ out) {
    __shared__ double red[256];
    double m = -INFINITY;
    for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, partial[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// double-double arithmetic (Knuth two-sum, fma two-product)
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    const double s = a.hi + b.hi;
    const double v = s - a.hi;
    const double err = (a.hi - (s - v)) + (b.hi - v);
    const double lo = err + a.lo + b.lo;
    const double hi = s + lo;
    return {hi, lo - (hi - s)};
}
__device__ __forceinline__ DD dd_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

// sum over the nodes of M * Q[:, state, :], M = the mass-matrix diagonal (vg field 9)
// M per node from vg (per-node geometry, field 9) or, when geo is set, from the affine
// element record: M = J * W[n] (J at geo[AFF_GEO e + 3]).
__global__ void __launch_bounds__(256) k_weightedsum(const double *__restrict__ Q, const double *__restrict__ vg,
                                                     int Np, int nvg, const double *__restrict__ geo,
                                                     const double *__restrict__ W, int state, int64_t nelem,
                                                     double2 *__restrict__ partial) {
    __shared__ DD red[256];
    DD acc{0.0, 0.0};
    const int64_t total = (int64_t)Np * nelem;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const int64_t e = t / Np;
        const int n = (int)(t - e * Np);
        const double M = geo ? geo[(int64_t)AFF_GEO * e + 3] * W[n] : vg[(int64_t)Np * (9 + (int64_t)nvg * e) + n];
        acc = dd_add(acc, dd_prod(Q[(int64_t)Np * (state + 5 * e) + n], M));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = dd_add(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(red[0].hi, red[0].lo);
}

// Which of the nstate fields of Q (Np, nstate, nelem) hold a non-finite value:
// bit s of the block's mask is set when some Q[:, s, :] in its share is NaN or Inf
// (the map-reduce of show_not_finite_fields, src/Arrays/MPIStateArrays.jl:884-903).
// Each block writes its mask (as a double, exact) to partial[blockIdx].x; k_or_reduce
// combines them.  Plain vector stores only.
__global__ void __launch_bounds__(256) k_not_finite(const double *__restrict__ Q, int Np, int nstate, int64_t total,
                                                    double2 *__restrict__ partial) {
    __shared__ unsigned red[256];
    unsigned mask = 0u;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
        const double v = Q[i];
        if (!isfinite(v)) mask |= 1u << (int)((i / Np) % nstate);
    }
    red[threadIdx.x] = mask;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] |= red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2((double)red[0], 0.0);
}

__global__ void __launch_bounds__(256) k_or_reduce(const double2 *__restrict__ partial, int n, double *__restrict__ out) {
    __shared__ unsigned red[256];
    unsigned mask = 0u;
    for (int i = threadIdx.x; i < n; i += 256) mask |= (unsigned)partial[i].x;
    red[threadIdx.x] = mask;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] |= red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (double)red[0];
}

__global__ void __launch_bounds__(256) k_dd_reduce(const double2 *__restrict__ partial, int n, double *__restrict__ out) {
    __shared__ DD red[256];
    DD acc{0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 256) acc = dd_add(acc, DD{partial[i].x, partial[i].y});
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = dd_add(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0].hi + red[0].lo;
}

}  // namespace cmdg

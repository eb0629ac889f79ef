// cmdg_learn.h -- kernels that build and train the learned closure and measure it:
// the FP64 velocity gradient of the DG gradient pass, the 2:1 element-agglomeration
// filter, training samples of the subgrid stress, the MLP loss gradient and Adam
// step, and DG -> equispaced interpolation for kinetic-energy spectra.
#pragma once
#include "cmdg_affine.h"
#include "cmdg_closure.h"

namespace cmdg {

// ------------------------------------------------------------------ velocity gradient
// du_c/dx_d at (n, 3c + d, e), FP64, formed as in the gradient pass: the volume
// derivative of u = rho u / rho plus the lifted central jump M^-1 Mf n (u* - u-),
// u* = (u+ + u-)/2 between elements, u* = u(q_wall) (normal momentum removed) on walls
// (DGModel_kernels.jl:1007-1127, 1471-1647; bc_momentum.jl:24-43).
template <int Nq>
__global__ void __launch_bounds__(threads_for<Nq>())
k_velgrad_aff(AffArrays A, int64_t nelem, const double *__restrict__ Q, double *__restrict__ Gu) {
    constexpr int Np = Nq * Nq * Nq;
    __shared__ double sD[Nq * Nq];
    __shared__ double sU[3][Np];
    const int64_t e = xcd_element(blockIdx.x, nelem);
    if (e >= nelem) return;
    const int n = threadIdx.x;
    if (n < Nq * Nq) sD[n] = A.D[n];
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    double u[3] = {0.0, 0.0, 0.0};
    if (act) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
        const double rinv = 1.0 / qb[0];
        for (int c = 0; c < 3; ++c) {
            u[c] = rinv * qb[Np * (1 + c)];
            sU[c][n] = u[c];
        }
    }
    __syncthreads();
    if (!act) return;
    const double *ge = A.geo + (int64_t)AFF_GEO * e;
    double g[3][3];
    for (int c = 0; c < 3; ++c) {
        double g1 = 0.0, g2 = 0.0, g3 = 0.0;
        for (int m = 0; m < Nq; ++m) {
            g1 += sD[i + Nq * m] * sU[c][m + Nq * (j + Nq * k)];
            g2 += sD[j + Nq * m] * sU[c][i + Nq * (m + Nq * k)];
            g3 += sD[k + Nq * m] * sU[c][i + Nq * (j + Nq * m)];
        }
        g[0][c] = ge[0] * g1;
        g[1][c] = ge[1] * g2;
        g[2][c] = ge[2] * g3;
    }
    const double MI = 1.0 / (ge[3] * A.W[n]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int fn;
        const int f = node_face_dir<Nq>(i, j, k, d, fn);
        if (f < 0) continue;
        const double w = MI * (ge[4 + f] * A.Wf[fn]) * ((f & 1) ? 1.0 : -1.0);
        const int32_t ne = A.nbe[6 * e + f];
        double jump[3];
        if (ne >= 0) {
            const double *qp = Q + (int64_t)Np * 5 * ne + face_node<Nq>(f ^ 1, fn);
            const double rp = 1.0 / qp[0];
            for (int s = 0; s < 3; ++s) jump[s] = 0.5 * (rp * qp[Np * (1 + s)] - u[s]);
        } else {
            for (int s = 0; s < 3; ++s) jump[s] = 0.0;
            jump[d] = -u[d];
        }
        for (int s = 0; s < 3; ++s) g[d][s] += w * jump[s];
    }
    double *o = Gu + (int64_t)Np * 9 * e + n;
    for (int cc = 0; cc < 3; ++cc)
        for (int d = 0; d < 3; ++d) o[(int64_t)Np * (3 * cc + d)] = g[d][cc];
}

// ------------------------------------------------------------------ 2:1 filter
// Coarse element ec is the union of 8 fine elements children[8 ec + (hx + 2 hy + 4 hz)],
// h = 0 for the lower half of the coarse element along that axis.  The coarse nodal
// values are the L2 projection (exact mass matrix) of the fine DG field onto the
// coarse polynomial space, a tensor product of the 1-D projections
// proj[h][i][a] (coarse node i, fine node a).  Fields filtered: rho, rho u (3), rho e
// and the second moments rho u_a u_b (11, 12, 13, 22, 23, 33) -> out (Np, 11, ncoarse).
constexpr int NMOM = 11;
struct FilterProj {
    double p[2][5][5];
};

template <int Nq>
__global__ void __launch_bounds__(threads_for<Nq>())
k_box_filter(const double *__restrict__ Qf, const int32_t *__restrict__ children, int64_t ncoarse, FilterProj PR,
             double *__restrict__ out) {
    constexpr int Np = Nq * Nq * Nq;
    __shared__ double sM[NMOM][Np];
    const int64_t ec = blockIdx.x;
    if (ec >= ncoarse) return;
    const int n = threadIdx.x;
    const bool act = n < Np;
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    double acc[NMOM];
    for (int f = 0; f < NMOM; ++f) acc[f] = 0.0;
    for (int ch = 0; ch < 8; ++ch) {
        const int hx = ch & 1, hy = (ch >> 1) & 1, hz = ch >> 2;
        const int64_t ef = children[8 * ec + ch];
        __syncthreads();
        if (act) {
            const double *qb = Qf + (int64_t)Np * 5 * ef + n;
            const double r = qb[0], m1 = qb[Np], m2 = qb[2 * Np], m3 = qb[3 * Np];
            const double ri = 1.0 / r;
            sM[0][n] = r;
            sM[1][n] = m1;
            sM[2][n] = m2;
            sM[3][n] = m3;
            sM[4][n] = qb[4 * Np];
            sM[5][n] = m1 * m1 * ri;
            sM[6][n] = m1 * m2 * ri;
            sM[7][n] = m1 * m3 * ri;
            sM[8][n] = m2 * m2 * ri;
            sM[9][n] = m2 * m3 * ri;
            sM[10][n] = m3 * m3 * ri;
        }
        __syncthreads();
        if (act) {
            for (int c = 0; c < Nq; ++c) {
                const double pz = PR.p[hz][k][c];
                for (int b = 0; b < Nq; ++b) {
                    const double pyz = PR.p[hy][j][b] * pz;
                    for (int a = 0; a < Nq; ++a) {
                        const double w = PR.p[hx][i][a] * pyz;
                        const int m = a + Nq * (b + Nq * c);
                        for (int f = 0; f < NMOM; ++f) acc[f] += w * sM[f][m];
                    }
                }
            }
        }
    }
    if (act)
        for (int f = 0; f < NMOM; ++f) out[(int64_t)Np * (f + NMOM * ec) + n] = acc[f];
}

// ------------------------------------------------------------------ training samples
// One sample per coarse node from the filtered moments M (Np, 11, E) and the coarse
// velocity gradient Gu (Np, 9, E):
//   u = (rho u)/rho,  tau_ab = (rho u_a u_b)/rho - u_a u_b   (subgrid stress),
//   S = sym(Gu), S^d its deviator, |S| = sqrt(2 S:S), k = Delta^2 |S|,
//   T = tau^d - 2 (nu_mol - nu_floor) S^d   (what nu_0 - nu_floor must supply),
//   a = 2 k S^d;  with the network's coefficient c the model's deviatoric stress is
//   -c a - 2 nu_floor S^d, so the squared error is c^2 P + 2 c Q + R with P = a:a,
//   Q = a:T, R = T:T.
// Row layout (SoA, leading dimension ld): rows 0-8 du_c/dx_d at 3c + d, row 9
// Delta, rows 10-12 P, Q, R.
constexpr int SAMPLE_W = 13;

__global__ void k_closure_samples(const double *__restrict__ M, const double *__restrict__ Gu, int64_t nelem, int Np,
                                  double Delta, double nu_mol, double nu_floor, float *__restrict__ out, int64_t ld,
                                  int64_t offset) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nelem * Np) return;
    const int64_t e = t / Np;
    const int n = (int)(t - e * Np);
    const double *mb = M + (int64_t)Np * NMOM * e + n;
    const double *gb = Gu + (int64_t)Np * 9 * e + n;
    const double r = mb[0], ri = 1.0 / r;
    const double u[3] = {mb[Np] * ri, mb[2 * Np] * ri, mb[3 * Np] * ri};
    const int mi[3][3] = {{5, 6, 7}, {6, 8, 9}, {7, 9, 10}};
    double tau[3][3], G[3][3], S[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            tau[a][b] = mb[(int64_t)Np * mi[a][b]] * ri - u[a] * u[b];
            G[a][b] = gb[(int64_t)Np * (3 * a + b)];
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) S[a][b] = 0.5 * (G[a][b] + G[b][a]);
    const double trS = (S[0][0] + S[1][1] + S[2][2]) / 3.0, trT = (tau[0][0] + tau[1][1] + tau[2][2]) / 3.0;
    double SS = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) SS += S[a][b] * S[a][b];
    const double c = Delta * Delta * sqrt(2.0 * SS);
    double P = 0.0, Qv = 0.0, R = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double Sd = S[a][b] - (a == b ? trS : 0.0);
            const double T = (tau[a][b] - (a == b ? trT : 0.0)) - 2.0 * (nu_mol - nu_floor) * Sd;
            const double av = 2.0 * c * Sd;
            P += av * av;
            Qv += av * T;
            R += T * T;
        }
    float *o = out + offset + (int64_t)(t);
    for (int q = 0; q < 9; ++q) o[ld * q] = (float)G[q / 3][q % 3];
    o[ld * 9] = (float)Delta;
    o[ld * 10] = (float)P;
    o[ld * 11] = (float)Qv;
    o[ld * 12] = (float)R;
}

// Replaces P, Q, R of each sample by their mean over the sample's element (the local
// average of the dynamic procedure): the optimum coefficient of a node is then the
// element's least-squares coefficient -sum Q / sum P, not the noisy pointwise one.
__global__ void __launch_bounds__(128) k_element_average_pqr(float *__restrict__ out, int64_t ld, int64_t offset,
                                                             int Np) {
    __shared__ double red[3][128];
    const int64_t e = blockIdx.x;
    const int n = threadIdx.x;
    float *o = out + offset + e * Np;
    for (int r = 0; r < 3; ++r) red[r][n] = n < Np ? (double)o[ld * (10 + r) + n] : 0.0;
    __syncthreads();
    for (int w = 64; w > 0; w >>= 1) {
        if (n < w)
            for (int r = 0; r < 3; ++r) red[r][n] += red[r][n + w];
        __syncthreads();
    }
    if (n < Np)
        for (int r = 0; r < 3; ++r) o[ld * (10 + r) + n] = (float)(red[r][0] / Np);
}

// ------------------------------------------------------------------ training
// FP32 master parameters, flat: W1 [32][10] at 0, b1 at 320, W2 [32][32] at 352,
// b2 at 1376, W3 [32] at 1408, b3 at 1440.
constexpr int MP_W1 = 0, MP_B1 = 320, MP_W2 = 352, MP_B2 = 1376, MP_W3 = 1408, MP_B3 = 1440, MLP_NPARAM = 1441;
constexpr int GRAD_W = MLP_NPARAM + 1;  // + loss
constexpr int TRAIN_TB = 64;            // samples per workgroup

__device__ __forceinline__ float silu_fwd(float a, float &dsilu) {
#pragma clang fp contract(fast)
    const float s = 1.0f / (1.0f + expf(-a));
    dsilu = s * (1.0f + a * (1.0f - s));
    return a * s;
}

// Per workgroup of 64 samples: the loss sum and its gradient with respect to every
// parameter, written to partial[blockIdx][GRAD_W].  Forward as the inference (FP32
// here, with the same input normalisation and output clamp), loss
// sum_s (c^2 P + 2 c Q + R) * inv_norm, c = clamp(y, MLP_CMIN, MLP_CMAX).
__global__ void __launch_bounds__(TRAIN_TB)
k_mlp_grad(const float *__restrict__ params, const float *__restrict__ shift, const float *__restrict__ scale,
           const float *__restrict__ samples, int64_t ld, const int64_t *__restrict__ idx, int64_t batch,
           float inv_norm, float *__restrict__ partial) {
#pragma clang fp contract(fast)
    __shared__ float sW[MLP_NPARAM];
    __shared__ float sX[MLP_IN][TRAIN_TB], sH1[MLP_H][TRAIN_TB], sH2[MLP_H][TRAIN_TB];
    __shared__ float sD1[MLP_H][TRAIN_TB], sD2[MLP_H][TRAIN_TB], sDy[TRAIN_TB], sL[TRAIN_TB];
    const int t = threadIdx.x;
    for (int q = t; q < MLP_NPARAM; q += TRAIN_TB) sW[q] = params[q];
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * TRAIN_TB + t;
    const bool valid = s < batch;
    const int64_t row = valid ? (idx ? idx[s] : s) : 0;
    float x[MLP_IN], P = 0.0f, Qv = 0.0f, R = 0.0f;
    for (int k = 0; k < MLP_IN; ++k) {
        x[k] = valid ? (samples[ld * k + row] - shift[k]) * scale[k] : 0.0f;
        sX[k][t] = x[k];
    }
    if (valid) {
        P = samples[ld * 10 + row];
        Qv = samples[ld * 11 + row];
        R = samples[ld * 12 + row];
    }
    float a1d[MLP_H], h1[MLP_H];
    for (int o = 0; o < MLP_H; ++o) {
        float a = sW[MP_B1 + o];
        for (int k = 0; k < MLP_IN; ++k) a += sW[MP_W1 + MLP_IN * o + k] * x[k];
        h1[o] = silu_fwd(a, a1d[o]);
        sH1[o][t] = h1[o];
    }
    float a2d[MLP_H], y = sW[MP_B3];
    for (int o = 0; o < MLP_H; ++o) {
        float a = sW[MP_B2 + o];
        for (int k = 0; k < MLP_H; ++k) a += sW[MP_W2 + MLP_H * o + k] * h1[k];
        const float h2 = silu_fwd(a, a2d[o]);
        sH2[o][t] = h2;
        y += sW[MP_W3 + o] * h2;
    }
    float dy = 0.0f, L = 0.0f;
    if (valid) {
        const bool clamped = y > MLP_CMAX || y < MLP_CMIN;
        const float c = fminf(fmaxf(y, MLP_CMIN), MLP_CMAX);
        L = (c * c * P + 2.0f * c * Qv + R) * inv_norm;
        dy = clamped ? 0.0f : (2.0f * c * P + 2.0f * Qv) * inv_norm;
    }
    sDy[t] = dy;
    sL[t] = L;
    float dh1[MLP_H];
    for (int k = 0; k < MLP_H; ++k) dh1[k] = 0.0f;
    for (int o = 0; o < MLP_H; ++o) {
        const float d2 = dy * sW[MP_W3 + o] * a2d[o];
        sD2[o][t] = d2;
        for (int k = 0; k < MLP_H; ++k) dh1[k] += sW[MP_W2 + MLP_H * o + k] * d2;
    }
    for (int k = 0; k < MLP_H; ++k) sD1[k][t] = dh1[k] * a1d[k];
    __syncthreads();
    float *pb = partial + (int64_t)blockIdx.x * GRAD_W;
    for (int q = t; q < GRAD_W; q += TRAIN_TB) {
        const float *u, *v = nullptr;
        if (q < MP_B1) { u = sD1[q / MLP_IN]; v = sX[q % MLP_IN]; }
        else if (q < MP_W2) u = sD1[q - MP_B1];
        else if (q < MP_B2) { const int r = q - MP_W2; u = sD2[r / MLP_H]; v = sH1[r % MLP_H]; }
        else if (q < MP_W3) u = sD2[q - MP_B2];
        else if (q < MP_B3) { u = sDy; v = sH2[q - MP_W3]; }
        else if (q == MP_B3) u = sDy;
        else u = sL;
        float acc = 0.0f;
        if (v)
            for (int r = 0; r < TRAIN_TB; ++r) acc += u[r] * v[r];
        else
            for (int r = 0; r < TRAIN_TB; ++r) acc += u[r];
        pb[q] = acc;
    }
}

// grad[q] = sum over workgroups of partial[b][q], in workgroup order (deterministic);
// the last entry is the loss.
__global__ void k_grad_reduce(const float *__restrict__ partial, int64_t nb, float *__restrict__ grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= GRAD_W) return;
    double acc = 0.0;
    for (int64_t b = 0; b < nb; ++b) acc += partial[b * GRAD_W + q];
    grad[q] = (float)acc;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
#pragma clang fp contract(fast)
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// ------------------------------------------------------------------ spectra
// Velocity u = (rho u)/rho of element e interpolated to p^3 equispaced points
// xi_m = -1 + (2m + 1)/p per axis (cell centres), written to
// out[c][iz p + mz][iy p + my][ix p + mx] for the element's brick index (ix, iy, iz)
// (InterpolationBrick semantics, src/Numerics/Mesh/Interpolation.jl:66-571, onto a
// uniform grid).  L[m][i] = l_i(xi_m).
struct InterpMat {
    double L[8][5];
};

template <int Nq>
__global__ void __launch_bounds__(512)
k_interp_equispaced(const double *__restrict__ Q, const int32_t *__restrict__ eijk, int64_t nelem, int p,
                    InterpMat IM, int64_t nxt, int64_t nyt, int64_t nzt, double *__restrict__ out) {
    constexpr int Np = Nq * Nq * Nq;
    __shared__ double sU[3][Np];
    const int64_t e = blockIdx.x;
    if (e >= nelem) return;
    for (int n = threadIdx.x; n < Np; n += blockDim.x) {
        const double *qb = Q + (int64_t)Np * 5 * e + n;
        const double ri = 1.0 / qb[0];
        for (int c = 0; c < 3; ++c) sU[c][n] = qb[Np * (1 + c)] * ri;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= p * p * p) return;
    const int mx = t % p, my = (t / p) % p, mz = t / (p * p);
    double v[3] = {0.0, 0.0, 0.0};
    for (int c3 = 0; c3 < Nq; ++c3)
        for (int b = 0; b < Nq; ++b) {
            const double lyz = IM.L[my][b] * IM.L[mz][c3];
            for (int a = 0; a < Nq; ++a) {
                const double w = IM.L[mx][a] * lyz;
                const int m = a + Nq * (b + Nq * c3);
                for (int c = 0; c < 3; ++c) v[c] += w * sU[c][m];
            }
        }
    const int64_t X = (int64_t)eijk[3 * e] * p + mx, Y = (int64_t)eijk[3 * e + 1] * p + my,
                  Z = (int64_t)eijk[3 * e + 2] * p + mz;
    for (int c = 0; c < 3; ++c) out[((c * nzt + Z) * nyt + Y) * nxt + X] = v[c];
}

}  // namespace cmdg

// cmdg_closure.h -- learned eddy-viscosity closure: a per-node MLP
// 10 -> 32 -> 32 -> 1 with SiLU hidden units, evaluated with BF16 MFMA and FP32
// accumulation inside the second-order-flux (gradient) kernel.
//
// Inputs (raw, per node): the nine velocity-gradient components du_c/dx_d at
// index 3c + d, and the filter width Delta at index 9; the network sees
// x = (raw - shift) * scale.  The output y is a signed eddy-viscosity coefficient
// in units of Delta^2 |S| (|S| = sqrt(2 S:S)):
//     nu_0 = max(nu_floor + clamp(y, MLP_CMIN, MLP_CMAX) * Delta^2 * |S|, 0),
// nu_floor the context's kinematic viscosity (the molecular viscosity of a flow at
// finite Reynolds number), 1e-5 when that is zero.  A negative coefficient takes
// viscosity away (backscatter), never below zero in total.  It takes the place of
// nu_0 = |S| (C_s Delta)^2 + 1e-5 of Smagorinsky-Lilly
// (src/Common/TurbulenceClosures/TurbulenceClosures.jl:478-491); the buoyancy
// correction and the stress/diffusivity that follow are Smagorinsky-Lilly's.  A
// network whose output is the constant C_s^2 reproduces Smagorinsky-Lilly.
//
// Arithmetic: inputs and weights are rounded to BF16 (round to nearest even),
// layers 1 and 2 run as 16x16x16 BF16 MFMA tiles with FP32 accumulation, the
// first hidden layer's SiLU output is rounded to BF16 to feed layer 2, layer 3
// (32 -> 1) is an FP32 dot product.  The hidden layers work in base-2 units: with
// u = -log2(e) a, SiLU(a) = a / (1 + e^-a) = -ln 2 * u / (1 + 2^u), so the device copy
// holds -log2(e) x scale, -log2(e) b1, -log2(e) b2 and -ln 2 w3 (W1 and W2 unchanged):
// the MFMA then yields u directly and a unit costs one exp2, one reciprocal, an add and
// a multiply.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cmdg {

constexpr int MLP_IN = 10, MLP_H = 32, MLP_K1 = 16;  // layer-1 K padded to 16
// the output clamp bounds the network's viscosity by 0.37 Delta^2 |S| either way (a
// Smagorinsky coefficient of at most 0.6, three times the usual 0.21): the closure
// cannot drive the explicit step past its viscous limit wherever the network
// extrapolates
constexpr float MLP_CMAX = 0.36787944f, MLP_CMIN = -0.36787944f;
// base-2 units of the hidden layers
constexpr double MLP_LOG2E = 1.4426950408889634, MLP_LN2 = 0.6931471805599453;

// Device copy of the network (one allocation per context), in the base-2 units of the
// hidden layers: b1, b2 and scale hold -log2(e) times the network's values, w3 holds
// -ln 2 times its BF16 value (each rounded to FP32 once, by mlp_device_copy).
struct MlpDev {
    uint16_t w1[MLP_H * MLP_K1];  // [out][in], in padded with zeros to 16
    uint16_t w2[MLP_H * MLP_H];   // [out][in]
    float b1[MLP_H], b2[MLP_H], w3[MLP_H], b3;
    float shift[MLP_IN], scale[MLP_IN];
};

// The device copy of a network given as BF16 weights (row-major [out][in]) and FP32
// biases, shift and scale.
inline MlpDev mlp_device_copy(const uint16_t *w1, const float *b1, const uint16_t *w2, const float *b2,
                              const uint16_t *w3, const float *b3, const float *shift, const float *scale) {
    MlpDev h{};
    for (int o = 0; o < MLP_H; ++o) {
        for (int k = 0; k < MLP_IN; ++k) h.w1[o * MLP_K1 + k] = w1[o * MLP_IN + k];
        for (int k = 0; k < MLP_H; ++k) h.w2[o * MLP_H + k] = w2[o * MLP_H + k];
        h.b1[o] = (float)(-MLP_LOG2E * (double)b1[o]);
        h.b2[o] = (float)(-MLP_LOG2E * (double)b2[o]);
        const uint32_t bits = (uint32_t)w3[o] << 16;
        float w;
        __builtin_memcpy(&w, &bits, 4);
        h.w3[o] = (float)(-MLP_LN2 * (double)w);
    }
    h.b3 = b3[0];
    for (int k = 0; k < MLP_IN; ++k) {
        h.shift[k] = shift[k];
        h.scale[k] = (float)(-MLP_LOG2E * (double)scale[k]);
    }
    return h;
}

typedef short mlp_s16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t mlp_u32x2 __attribute__((ext_vector_type(2)));

// BF16 rounding (nearest even) of two values by one hardware conversion: lo in the
// low half of the result.
__device__ __forceinline__ uint32_t bf16x2_hw(float lo, float hi) {
    const mlp_f32x2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, mlp_bf16x2));
}

// A hidden unit in base-2 units: u / (1 + 2^u) = -SiLU(a) / ln 2 for u = -log2(e) a,
// with the hardware exp2 and reciprocal (1 ulp each): its outputs are rounded to BF16
// (layer 1) or weighted in an FP32 sum (layer 2).
__device__ __forceinline__ float mlp_silu2(float u) {
#pragma clang fp contract(fast)
    return u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}

// LDS needed per 64-lane wave: 64 rows of 16 BF16 inputs.
constexpr int MLP_LDS_PER_WAVE = 64 * MLP_K1;

// Evaluates the network for the 64 nodes of the calling wave: lane l supplies the
// normalised inputs x of its node and receives that node's output.  Every lane of the
// wave must call this function (the wave exchanges its inputs through LDS); sX points
// to the workgroup's LDS slab of MLP_LDS_PER_WAVE uint16 per wave.
//
// Tiling: the wave's 64 nodes are four column tiles of 16; hidden features are the
// rows.  For a 16x16x16 MFMA, lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)..+3][l%16], so the D fragment of feature rows
// 16m..16m+15 is, element for element, the B fragment of K chunk m of the next layer:
// layer 1's output feeds layer 2 from registers, without a transpose.
__device__ __forceinline__ float mlp_wave(const MlpDev *__restrict__ W, const float x[MLP_IN], uint16_t *sX) {
#pragma clang fp contract(fast)
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    uint16_t *slab = sX + (threadIdx.x >> 6) * MLP_LDS_PER_WAVE;
    uint32_t *row = reinterpret_cast<uint32_t *>(slab + lane * MLP_K1);
#pragma unroll
    for (int k = 0; k < MLP_IN / 2; ++k) row[k] = bf16x2_hw(x[2 * k], x[2 * k + 1]);
#pragma unroll
    for (int k = MLP_IN / 2; k < MLP_K1 / 2; ++k) row[k] = 0u;
    // the slab is the wave's own: LDS serves one wave's accesses in order, so the reads
    // below see these writes once the compiler keeps them in program order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const mlp_s16x4 *w1 = reinterpret_cast<const mlp_s16x4 *>(W->w1);
    const mlp_s16x4 *w2 = reinterpret_cast<const mlp_s16x4 *>(W->w2);
    const mlp_s16x4 *xs = reinterpret_cast<const mlp_s16x4 *>(slab);
    mlp_s16x4 A1[2], A2[2][2];
    mlp_f32x4 B1[2], B2[2], W3[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        A1[m] = w1[(16 * m + r) * (MLP_K1 / 4) + g];
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) A2[m][kc] = w2[(16 * m + r) * (MLP_H / 4) + 4 * kc + g];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            B1[m][i] = W->b1[16 * m + 4 * g + i];
            B2[m][i] = W->b2[16 * m + 4 * g + i];
            W3[m][i] = W->w3[16 * m + 4 * g + i];
        }
    }
    float mine = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const mlp_s16x4 bx = xs[(16 * t + r) * (MLP_K1 / 4) + g];
        mlp_s16x4 hb[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const mlp_f32x4 h = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1[m], bx, B1[m], 0, 0, 0);
            const mlp_u32x2 w = {bf16x2_hw(mlp_silu2(h[0]), mlp_silu2(h[1])),
                                 bf16x2_hw(mlp_silu2(h[2]), mlp_silu2(h[3]))};
            hb[m] = __builtin_bit_cast(mlp_s16x4, w);
        }
        float part = 0.0f;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            mlp_f32x4 a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A2[m][0], hb[0], B2[m], 0, 0, 0);
            a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A2[m][1], hb[1], a, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) part += W3[m][i] * mlp_silu2(a[i]);
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (g == t) mine = part;
    }
    return mine + W->b3;
}

// Learned nu_0 from the 3 x 5 gradient g[d][s] (s < 3: velocity), Delta and |S|.
__device__ __forceinline__ void mlp_inputs(const MlpDev *__restrict__ W, const float g[3][5], float Delta,
                                           float x[MLP_IN]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) x[3 * c + d] = (g[d][c] - W->shift[3 * c + d]) * W->scale[3 * c + d];
    x[9] = (Delta - W->shift[9]) * W->scale[9];
}

__device__ __forceinline__ float mlp_nu0(float y, float Delta, float normS, float nu_floor) {
#pragma clang fp contract(fast)
    const float c = fminf(fmaxf(y, MLP_CMIN), MLP_CMAX);
    return fmaxf(nu_floor + c * (Delta * Delta) * normS, 0.0f);
}

// Stand-alone evaluation of the network on n raw input rows (row-major n x 10):
// y[i] = output for row i.  128 threads per workgroup.
__global__ void __launch_bounds__(128) k_mlp_forward(const MlpDev *__restrict__ W, const float *__restrict__ x_raw,
                                                     float *__restrict__ y, int64_t n) {
#pragma clang fp contract(fast)
    __shared__ uint16_t sX[2 * MLP_LDS_PER_WAVE];
    const int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x;
    float x[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) x[k] = i < n ? (x_raw[i * MLP_IN + k] - W->shift[k]) * W->scale[k] : 0.0f;
    const float v = mlp_wave(W, x, sX);
    if (i < n) y[i] = v;
}

}  // namespace cmdg

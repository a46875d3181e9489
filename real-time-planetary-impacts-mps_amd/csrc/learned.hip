/*
 * learned.hip -- learned Newton-Raphson starting guess for the smoothing
 * length, for gfx950.  No counterpart in the reference: it seeds the
 * iteration of density.metal:73-143, which stays the judge of convergence.
 *
 *   features   12 per particle at the start of a density pass: log h,
 *              (dh/dt) dt / h, log rho, 7 statistics of the particle counts
 *              of the 27 cells around the particle's cell, material one-hot
 *   network    12 -> 64 -> 64 -> 1, SiLU; layers 1 and 2 on the matrix cores
 *              (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulate),
 *              layer 3 and every elementwise step in FP32; predicts
 *              d = log(h_converged) - log(h_start) on normalised features
 *   training   weighted MSE, backward on the matrix cores (dW2, dA1, dW1 are
 *              16x16x32 MFMA products over 32 samples), per-wave partial
 *              gradients summed per workgroup, then over the workgroups, in
 *              a fixed order; Adam in FP32
 *   samples    16-float records (12 features, d, weight, 2 spare) appended
 *              by density passes with dumping on
 *
 * MFMA 16x16x32 operand layout (lane l of 64, c = l % 16, g = l / 16):
 *   A[i][k]: lane i + 16 (k / 8), element k % 8   -> lane holds A[c][8g .. 8g+7]
 *   B[k][j]: lane j + 16 (k / 8), element k % 8   -> lane holds B[8g .. 8g+7][c]
 *   C[i][j]: lane j + 16 (i / 4), element i % 4   -> lane holds C[4g .. 4g+3][c]
 * A wave processes 32 samples per iteration: two 16-row tiles in the forward
 * pass and one K = 32 reduction over samples for each weight gradient.
 * Transposes between the C and A/B layouts go through LDS; all waves of a
 * workgroup run the same number of iterations, so block barriers order them.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sph_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int NI = SPH_MLP_INPUTS;   /* 12 */
constexpr int NH = SPH_MLP_HIDDEN;   /* 64 */
constexpr int NP = SPH_MLP_PARAMS;   /* 5057 */
constexpr int OW1 = 0, OB1 = NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH;
static_assert(OB3 + 1 == NP, "parameter layout");
constexpr int S = 32;    /* samples per wave iteration */
constexpr int P64 = 72;  /* LDS pitch (bf16) of 64-wide rows: +16 B against bank conflicts */
constexpr int P32 = 40;  /* LDS pitch (bf16) of 32-wide rows */
constexpr int WAVES = 4; /* per workgroup */
constexpr int TRAIN_BLOCKS = 256;  /* one workgroup per CU (the training LDS is ~130 KB) */
constexpr int INFER_BLOCKS = 2048; /* inference: ~35 KB LDS, several workgroups per CU */
constexpr float D_CLAMP = 0.5f; /* |log(h_guess / h_start)| cap */

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ float sig(float z) { return 1.f / (1.f + expf(-z)); }
__device__ __forceinline__ f32x4 splat(float v) { return f32x4{v, v, v, v}; }

/* Weights staged per workgroup: BF16 copies for the MFMA operands, FP32 for
 * biases and the last layer. */
struct WeightsLds {
    __bf16 w1[NH * P32];   /* [j][k], k >= 12 zero */
    __bf16 w2[NH * P64];   /* [j][k] */
    __bf16 w2t[NH * P64];  /* [k][j] (training only) */
    float b1[NH], b2[NH], w3[NH];
    float b3;
    float norm[SPH_MLP_NORM];
};

/* Inference needs only the forward operands. */
struct WeightsInf {
    __bf16 w1[NH * P32];
    __bf16 w2[NH * P64];
    float b1[NH], b2[NH], w3[NH];
    float b3;
    float norm[SPH_MLP_NORM];
};
struct WaveInf {
    __bf16 a1[S * P64];
};

struct WaveLds {
    __bf16 a1[S * P64];    /* [s][k] */
    __bf16 a1t[NH * P32];  /* [k][s] */
    __bf16 d2[S * P64];    /* dz2 [s][j] */
    __bf16 d2t[NH * P32];  /* dz2 [j][s] */
    __bf16 d1t[NH * P32];  /* dz1 [j][s] */
    __bf16 xnt[16 * P32];  /* xn [f][s] */
    float tgt[S], wgt[S];
};

template <bool TRAIN, class W>
__device__ void stage_weights(W& w, const float* __restrict__ p, const float* __restrict__ norm) {
    for (int e = threadIdx.x; e < NH * 32; e += blockDim.x) {
        const int j = e / 32, k = e % 32;
        w.w1[j * P32 + k] = (__bf16)(k < NI ? p[OW1 + j * NI + k] : 0.f);
    }
    for (int e = threadIdx.x; e < NH * NH; e += blockDim.x) {
        const int j = e / NH, k = e % NH;
        const float v = p[OW2 + e];
        w.w2[j * P64 + k] = (__bf16)v;
        if constexpr (TRAIN) w.w2t[k * P64 + j] = (__bf16)v;
    }
    for (int e = threadIdx.x; e < NH; e += blockDim.x) {
        w.b1[e] = p[OB1 + e];
        w.b2[e] = p[OB2 + e];
        w.w3[e] = p[OW3 + e];
    }
    for (int e = threadIdx.x; e < SPH_MLP_NORM; e += blockDim.x) w.norm[e] = norm[e];
    if (threadIdx.x == 0) w.b3 = p[OB3];
}

/* Forward pass of 32 samples.  `xa[m]` is this lane's A operand of tile m
 * (normalised features of sample 16 m + c, features 8g .. 8g+7).  Leaves
 * z1, z2 in C layout, a1 in LDS (row-major, and transposed if `train`), and
 * y[m][r] = network output of sample 16 m + 4 g + r in every lane of group g. */
template <bool TRAIN, class W, class L>
__device__ __forceinline__ void forward(const W& w, L& l, const bf16x8 xa[2], int c, int g,
                                        f32x4 z1[2][4], f32x4 z2[2][4], float y[2][4]) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x8 b = ld8(&w.w1[(t * 16 + c) * P32 + g * 8]);
            z1[m][t] = mfma(xa[m], b, splat(w.b1[t * 16 + c]));
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float z = z1[m][t][r];
                const __bf16 a = (__bf16)(z * sig(z));
                const int s = m * 16 + g * 4 + r, k = t * 16 + c;
                l.a1[s * P64 + k] = a;
                if constexpr (TRAIN) l.a1t[k * P32 + s] = a;
            }
        }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = splat(w.b2[t * 16 + c]);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const bf16x8 a = ld8(&l.a1[(m * 16 + c) * P64 + ks * 32 + g * 8]);
                const bf16x8 b = ld8(&w.w2[(t * 16 + c) * P64 + ks * 32 + g * 8]);
                acc = mfma(a, b, acc);
            }
            z2[m][t] = acc;
            const float w3 = w.w3[t * 16 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float z = acc[r];
                part[r] += (z * sig(z)) * w3;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = part[r];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
            y[m][r] = v + w.b3;
        }
    }
}

/* ------------------------------------------------------------- training */
struct TrainArgs {
    const float4* rec;   /* 4 float4 per sample */
    int64_t count;       /* samples available */
    int batch;           /* samples per step (multiple of 32) */
    float inv_norm;      /* 1 / (samples the loss is averaged over) */
    int iters;           /* 32-sample chunks per wave */
    int random;          /* 1: draw samples with a counter hash; 0: sample b = record b */
    uint64_t seed;
    int step;
    const float* params;
    const float* norm;
    float* part;         /* [workgroups][NP] */
    float* loss_part;    /* [workgroups] */
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_train(TrainArgs a) {
    __shared__ WeightsLds w;
    __shared__ WaveLds wl[WAVES];
    stage_weights<true>(w, a.params, a.norm);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int gw = blockIdx.x * WAVES + wave;
    const int total_waves = gridDim.x * WAVES;
    const int chunks = a.batch / S;
    WaveLds& l = wl[wave];
    const float ymu = w.norm[2 * NI], ysig = w.norm[2 * NI + 1];

    f32x4 gW2[4][4], gW1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        gW1[i] = splat(0.f);
#pragma unroll
        for (int j = 0; j < 4; ++j) gW2[i][j] = splat(0.f);
    }
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0};
    float gb3 = 0.f, loss = 0.f;
    const float inv_b = a.inv_norm;

    for (int it = 0; it < a.iters; ++it) {
        const int chunk = it * total_waves + gw;
        const bool live = chunk < chunks;
        bf16x8 xa[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int s = m * 16 + c;
            const int64_t b = (int64_t)chunk * S + s;
            int64_t idx = 0;
            bool ok = live;
            if (a.random) idx = (int64_t)(mix64(a.seed ^ ((uint64_t)a.step << 32) ^ (uint64_t)b) % (uint64_t)a.count);
            else {
                idx = b;
                ok = ok && b < a.count;
            }
            if (!ok) idx = 0;
            float x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = 0.f;
            if (g < 2) {
                const float4 v0 = a.rec[idx * 4 + g * 2];
                const float4 v1 = a.rec[idx * 4 + g * 2 + 1];
                const float raw[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int f = g * 8 + e;
                    if (f < NI) x[e] = (raw[e] - w.norm[f]) / w.norm[NI + f];
                }
            }
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (__bf16)x[e];
            xa[m] = v;
            if (g < 2) {
#pragma unroll
                for (int e = 0; e < 8; ++e) l.xnt[(g * 8 + e) * P32 + s] = v[e];
            }
            if (g == 0) {
                const float4 v3 = a.rec[idx * 4 + 3];
                l.tgt[s] = (v3.x - ymu) / ysig;
                l.wgt[s] = ok ? v3.y : 0.f;
            }
        }
        f32x4 z1[2][4], z2[2][4];
        float y[2][4];
        forward<true>(w, l, xa, c, g, z1, z2, y);
        /* loss, output layer, dz2 */
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            float dy[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s = m * 16 + g * 4 + r;
                const float res = y[m][r] - l.tgt[s];
                const float ws = l.wgt[s];
                dy[r] = 2.f * ws * res * inv_b;
                if (c == 0) {
                    loss += ws * res * res;
                    gb3 += dy[r];
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float w3 = w.w3[t * 16 + c];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float z = z2[m][t][r];
                    const float sg = sig(z);
                    gw3[t] += dy[r] * (z * sg);
                    const float d = dy[r] * w3 * (sg * (1.f + z * (1.f - sg)));
                    gb2[t] += d;
                    const int s = m * 16 + g * 4 + r, j = t * 16 + c;
                    const __bf16 db = (__bf16)d;
                    l.d2[s * P64 + j] = db;
                    l.d2t[j * P32 + s] = db;
                }
            }
        }
        __syncthreads();
        /* dW2 = dz2^T a1 over the 32 samples */
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) {
            const bf16x8 A = ld8(&l.d2t[(ti * 16 + c) * P32 + g * 8]);
#pragma unroll
            for (int tn = 0; tn < 4; ++tn) gW2[ti][tn] = mfma(A, ld8(&l.a1t[(tn * 16 + c) * P32 + g * 8]), gW2[ti][tn]);
        }
        /* dA1 = dz2 W2, dz1 = dA1 * silu'(z1) */
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int tn = 0; tn < 4; ++tn) {
                f32x4 acc = splat(0.f);
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
                    acc = mfma(ld8(&l.d2[(m * 16 + c) * P64 + ks * 32 + g * 8]),
                               ld8(&w.w2t[(tn * 16 + c) * P64 + ks * 32 + g * 8]), acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float z = z1[m][tn][r];
                    const float sg = sig(z);
                    const float d = acc[r] * (sg * (1.f + z * (1.f - sg)));
                    gb1[tn] += d;
                    l.d1t[(tn * 16 + c) * P32 + m * 16 + g * 4 + r] = (__bf16)d;
                }
            }
        __syncthreads();
        /* dW1 = dz1^T xn */
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
            gW1[ti] = mfma(ld8(&l.d1t[(ti * 16 + c) * P32 + g * 8]), ld8(&l.xnt[c * P32 + g * 8]), gW1[ti]);
        __syncthreads();
    }

    /* per-wave partial gradient into LDS (the wave buffers are free now),
     * every parameter written once; then the workgroup's partial, the sum
     * of its waves in wave order */
    static_assert(sizeof(WaveLds) * WAVES >= sizeof(float) * (WAVES * NP + WAVES), "reduction buffer");
    __syncthreads();
    float* red = reinterpret_cast<float*>(&wl[0]);
    float* out = red + wave * NP;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = ti * 16 + g * 4 + r;
            if (c < NI) out[OW1 + j * NI + c] = gW1[ti][r];
#pragma unroll
            for (int tn = 0; tn < 4; ++tn) out[OW2 + j * NH + tn * 16 + c] = gW2[ti][tn][r];
        }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float v1 = gb1[t], v2 = gb2[t], v3 = gw3[t];
#pragma unroll
        for (int off = 16; off < 64; off <<= 1) {
            v1 += __shfl_xor(v1, off, 64);
            v2 += __shfl_xor(v2, off, 64);
            v3 += __shfl_xor(v3, off, 64);
        }
        if (g == 0) {
            out[OB1 + t * 16 + c] = v1;
            out[OB2 + t * 16 + c] = v2;
            out[OW3 + t * 16 + c] = v3;
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off, 64);
        loss += __shfl_xor(loss, off, 64);
    }
    if (lane == 0) {
        out[OB3] = gb3;
        red[WAVES * NP + wave] = loss * inv_b;
    }
    __syncthreads();
    float* bout = a.part + (int64_t)blockIdx.x * NP;
    for (int i = threadIdx.x; i < NP; i += blockDim.x) {
        float v = red[i];
#pragma unroll
        for (int k = 1; k < WAVES; ++k) v += red[k * NP + i];
        bout[i] = v;
    }
    if (threadIdx.x == 0) {
        float L = red[WAVES * NP];
        for (int k = 1; k < WAVES; ++k) L += red[WAVES * NP + k];
        a.loss_part[blockIdx.x] = L;
    }
}

/* Sum of the workgroup partials, then one Adam update (Kingma & Ba 2015,
 * algorithm 1) when `lr` > 0.  A workgroup takes RED_P parameters: its 256
 * threads split the `waves` partials of each into 256 / RED_P interleaved
 * slices, and the slice sums are added in slice order (a fixed order, so
 * results are reproducible).  The last workgroup sums the loss partials
 * (at most 256) with a fixed-order tree. */
constexpr int RED_P = 16;
static_assert(TRAIN_BLOCKS <= 256, "the loss partials are summed by one workgroup of 256");
__global__ __launch_bounds__(256) void k_reduce_adam(const float* __restrict__ part,
                                                     const float* __restrict__ loss_part, int waves,
                                                     float* __restrict__ p, float* __restrict__ mom1,
                                                     float* __restrict__ mom2, float* __restrict__ grad,
                                                     float* __restrict__ loss_out, float lr, float bc1, float bc2) {
    constexpr int NS = 256 / RED_P;
    __shared__ float red[256];
    const int t = threadIdx.x;
    if (blockIdx.x == gridDim.x - 1) {
        red[t] = t < waves ? loss_part[t] : 0.f;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
        }
        if (t == 0) loss_out[0] = red[0];
        return;
    }
    const int c = t % RED_P, sl = t / RED_P;
    const int i = blockIdx.x * RED_P + c;
    float s = 0.f;
    if (i < NP)
        for (int k = sl; k < waves; k += NS) s += part[(int64_t)k * NP + i];
    red[t] = s;
    __syncthreads();
    if (sl != 0 || i >= NP) return;
    float gsum = red[c];
    for (int q = 1; q < NS; ++q) gsum += red[q * RED_P + c];
    grad[i] = gsum;
    if (lr <= 0.f) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * mom1[i] + (1.f - b1) * gsum;
    const float v = b2 * mom2[i] + (1.f - b2) * gsum * gsum;
    mom1[i] = m;
    mom2[i] = v;
    p[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

/* ------------------------------------------------------------ inference */
struct InferArgs {
    const float4* feats;  /* 3 float4 per row */
    int64_t rows;
    const float* params;
    const float* norm;
    float* d_out;         /* raw prediction per row (may be null) */
    /* guess mode: row = sorted slot, or list[row] when `list` is set (then
     * the row count is read from `count` on the device) */
    const int32_t* list;
    const int32_t* count;
    const uint2* sorted;
    const float* h;
    float max_h;
    float* guess;         /* h_start exp(clamp(d)), per sorted slot (may be null) */
};

/* Workgroups loop over 32-sample tiles (one per wave per round, the round
 * count uniform over the workgroup because forward() has block barriers),
 * so the weights are staged once per workgroup. */
__global__ __launch_bounds__(256) void k_infer(InferArgs a) {
    __shared__ WeightsInf w;
    __shared__ WaveInf wl[WAVES];
    stage_weights<false>(w, a.params, a.norm);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    WaveInf& l = wl[wave];
    const float ymu = w.norm[2 * NI], ysig = w.norm[2 * NI + 1];
    const int64_t rows = a.count ? (int64_t)a.count[0] : a.rows;
    const int64_t tiles = (rows + S - 1) / S;
    for (int64_t t0 = (int64_t)blockIdx.x * WAVES; t0 < tiles; t0 += (int64_t)gridDim.x * WAVES) {
        const int64_t base = (t0 + wave) * S;
        bf16x8 xa[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int64_t r0 = base + m * 16 + c;
            const bool ok = r0 < rows;
            const int64_t row = ok && a.list ? (int64_t)a.list[r0] : r0;
            float x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = 0.f;
            if (ok) {
                if (g == 0) {
                    const float4 v0 = a.feats[row * 3], v1 = a.feats[row * 3 + 1];
                    const float raw[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                    for (int e = 0; e < 8; ++e) x[e] = (raw[e] - w.norm[e]) / w.norm[NI + e];
                } else if (g == 1) {
                    const float4 v2 = a.feats[row * 3 + 2];
                    const float raw[4] = {v2.x, v2.y, v2.z, v2.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) x[e] = (raw[e] - w.norm[8 + e]) / w.norm[NI + 8 + e];
                }
            }
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (__bf16)x[e];
            xa[m] = v;
        }
        f32x4 z1[2][4], z2[2][4];
        float y[2][4];
        forward<false>(w, l, xa, c, g, z1, z2, y);
        if (c == 0) {
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t r0 = base + m * 16 + g * 4 + r;
                    if (r0 >= rows) continue;
                    const int64_t row = a.list ? (int64_t)a.list[r0] : r0;
                    const float d = y[m][r] * ysig + ymu;
                    if (a.d_out) a.d_out[row] = d;
                    if (a.guess) {
                        const float h0 = a.h[a.sorted[row].y];
                        a.guess[row] = fminf(h0 * expf(fminf(fmaxf(d, -D_CLAMP), D_CLAMP)), a.max_h);
                    }
                }
        }
        __syncthreads();  /* the next round rewrites the wave's a1 tile */
    }
}

unsigned infer_blocks(int64_t rows) {
    const int64_t tiles = (rows + S - 1) / S;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(INFER_BLOCKS, (tiles + WAVES - 1) / WAVES));
}

/* ------------------------------------------------------------- features */
__device__ __forceinline__ uint32_t compact_bits(uint32_t v) {
    v &= 0x49249249u;
    v = (v ^ (v >> 2)) & 0xC30C30C3u;
    v = (v ^ (v >> 4)) & 0x0F00F00Fu;
    v = (v ^ (v >> 8)) & 0xFF0000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

struct FeatArgs {
    int n;
    const uint2* sorted;
    const uint2* cell_se;
    int mask;
    const float* h;
    const float* dhdt;
    const float* rho;
    const int32_t* mat;
    const int32_t* time;
    float dtmin1;
    float4* feats;        /* 3 per sorted slot */
    const int32_t* list;  /* optional: only these sorted slots, count[0] of them */
    const int32_t* count;
};

__global__ __launch_bounds__(256) void k_features(FeatArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (a.list ? a.count[0] : a.n)) return;
    const int index = a.list ? a.list[t] : t;
    const uint2 kp = a.sorted[index];
    const int cx = (int)compact_bits(kp.x), cy = (int)compact_bits(kp.x >> 1), cz = (int)compact_bits(kp.x >> 2);
    float tot = 0.f, own = 0.f, face = 0.f, edge = 0.f, corner = 0.f, mx = 0.f, mn = 3.0e38f;
    float dx = 0.f, dy = 0.f, dz = 0.f;
    for (int ox = -1; ox <= 1; ++ox) {
        const uint32_t ex = dev_expand_bits((uint32_t)((cx + ox) & a.mask));
        for (int oy = -1; oy <= 1; ++oy) {
            const uint32_t exy = ex | (dev_expand_bits((uint32_t)((cy + oy) & a.mask)) << 1);
            for (int oz = -1; oz <= 1; ++oz) {
                const uint32_t key = exy | (dev_expand_bits((uint32_t)((cz + oz) & a.mask)) << 2);
                const uint2 se = a.cell_se[key];
                const float cnt = se.x == 0xFFFFFFFFu ? 0.f : (float)(se.y - se.x + 1u);
                const int order = abs(ox) + abs(oy) + abs(oz);
                tot += cnt;
                if (order == 0) own += cnt;
                else if (order == 1) face += cnt;
                else if (order == 2) edge += cnt;
                else corner += cnt;
                mx = fmaxf(mx, cnt);
                mn = fminf(mn, cnt);
                dx += cnt * (float)ox;
                dy += cnt * (float)oy;
                dz += cnt * (float)oz;
            }
        }
    }
    const int i = (int)kp.y;
    const float h = a.h[i];
    const float dt = a.dtmin1 * (float)a.time[1];
    const float inv = 1.f / (tot + 1.f);
    const bool m400 = a.mat[i] == 400;
    a.feats[3 * index + 0] = make_float4(logf(h), a.dhdt[i] * dt / h, logf(a.rho[i]), log1pf(tot));
    a.feats[3 * index + 1] = make_float4(log1pf(own), log1pf(face), log1pf(edge), log1pf(corner));
    a.feats[3 * index + 2] =
        make_float4(sqrtf(dx * dx + dy * dy + dz * dz) * inv, (mx - mn) * inv, m400 ? 1.f : 0.f, m400 ? 0.f : 1.f);
}

/* Appends (features, log(h_final) - log(h_start), weight) per sorted slot. */
__global__ void k_dump(int n, const float4* __restrict__ feats, const uint2* __restrict__ sorted,
                       const float* __restrict__ h, const uint8_t* __restrict__ active, float4* __restrict__ rec) {
    const int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= n) return;
    const float4 f0 = feats[3 * index];
    const int i = (int)sorted[index].y;
    rec[4 * (int64_t)index + 0] = f0;
    rec[4 * (int64_t)index + 1] = feats[3 * index + 1];
    rec[4 * (int64_t)index + 2] = feats[3 * index + 2];
    rec[4 * (int64_t)index + 3] = make_float4(logf(h[i]) - f0.x, active[i] ? 1.f : 0.f, 0.f, 0.f);
}

/* FP64 per-block sums of the features, the target and their squares over
 * weighted samples (normalisation statistics). */
__global__ void k_stats(const float4* __restrict__ rec, int64_t count, double* __restrict__ out) {
    __shared__ double red[256];
    double acc[2 * (NI + 1) + 1];
    for (int k = 0; k < 2 * (NI + 1) + 1; ++k) acc[k] = 0.0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < count; s += (int64_t)gridDim.x * blockDim.x) {
        const float4 q3 = rec[4 * s + 3];
        if (q3.y <= 0.f) continue;
        float v[NI + 1];
        const float4 q0 = rec[4 * s], q1 = rec[4 * s + 1], q2 = rec[4 * s + 2];
        v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w;
        v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
        v[8] = q2.x; v[9] = q2.y; v[10] = q2.z; v[11] = q2.w;
        v[12] = q3.x;
        for (int k = 0; k <= NI; ++k) {
            acc[k] += v[k];
            acc[NI + 1 + k] += (double)v[k] * v[k];
        }
        acc[2 * (NI + 1)] += 1.0;
    }
    for (int k = 0; k < 2 * (NI + 1) + 1; ++k) {
        red[threadIdx.x] = acc[k];
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[blockIdx.x * (2 * (NI + 1) + 1) + k] = red[0];
        __syncthreads();
    }
}

}  // namespace

/* ====================================================================== */
struct sph_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;
    float* mom1 = nullptr;
    float* mom2 = nullptr;
    float* norm = nullptr;
    float* grad = nullptr;
    float* part = nullptr;       /* [workgroups][NP] */
    float* loss_part = nullptr;
    float* loss = nullptr;       /* per training step */
    int loss_cap = 0;
    int part_waves = 0;
    int64_t adam_t = 0;
    float4* scratch_rec = nullptr; /* records for explicit-array calls */
    int64_t scratch_cap = 0;
    float4* scratch_feats = nullptr;
    float* scratch_d = nullptr;
    std::string err;
};

struct sph_samples {
    int device = 0;
    float4* rec = nullptr;
    int64_t cap = 0;
    int64_t count = 0;
    hipEvent_t written = nullptr; /* last append (any stream) */
};

namespace {

int mlp_fail(sph_mlp* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    return code;
}

#define MLP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) return mlp_fail(m, (int)_e, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

int ensure_scratch(sph_mlp* m, int64_t rows) {
    if (rows <= m->scratch_cap) return SPH_OK;
    dfree(m->scratch_rec);
    dfree(m->scratch_feats);
    dfree(m->scratch_d);
    const int64_t cap = (rows + 31) / 32 * 32;
    MLP_TRY(dalloc(&m->scratch_rec, (size_t)cap * 4));
    MLP_TRY(dalloc(&m->scratch_feats, (size_t)cap * 3));
    MLP_TRY(dalloc(&m->scratch_d, (size_t)cap));
    m->scratch_cap = cap;
    return SPH_OK;
}

int ensure_loss(sph_mlp* m, int count) {
    if (count <= m->loss_cap) return SPH_OK;
    dfree(m->loss);
    MLP_TRY(dalloc(&m->loss, (size_t)count));
    m->loss_cap = count;
    return SPH_OK;
}

int ensure_part(sph_mlp* m, int parts) {
    if (parts <= m->part_waves) return SPH_OK;
    dfree(m->part);
    dfree(m->loss_part);
    MLP_TRY(dalloc(&m->part, (size_t)parts * NP));
    MLP_TRY(dalloc(&m->loss_part, (size_t)parts));
    m->part_waves = parts;
    return SPH_OK;
}

/* One forward + backward over `batch` samples (and an Adam step if lr > 0). */
int run_step(sph_mlp* m, const float4* rec, int64_t count, int batch, int64_t norm_count, int random,
             uint64_t seed, int step, float lr, float* loss_dst) {
    const int chunks = batch / S;
    const int blocks = std::min(TRAIN_BLOCKS, (chunks + WAVES - 1) / WAVES);
    const int waves = blocks * WAVES;
    int rc = ensure_part(m, blocks);
    if (rc != SPH_OK) return rc;
    TrainArgs a;
    a.rec = rec;
    a.count = count;
    a.batch = batch;
    a.inv_norm = 1.f / (float)norm_count;
    a.iters = (chunks + waves - 1) / waves;
    a.random = random;
    a.seed = seed;
    a.step = step;
    a.params = m->params;
    a.norm = m->norm;
    a.part = m->part;
    a.loss_part = m->loss_part;
    k_train<<<blocks, WAVES * 64, 0, m->stream>>>(a);
    MLP_TRY(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    if (lr > 0.f) {
        m->adam_t += 1;
        bc1 = (float)(1.0 - std::pow(0.9, (double)m->adam_t));
        bc2 = (float)(1.0 - std::pow(0.999, (double)m->adam_t));
    }
    /* one workgroup per RED_P parameters, one more for the loss (blocks <= 256) */
    k_reduce_adam<<<(NP + RED_P - 1) / RED_P + 1, 256, 0, m->stream>>>(m->part, m->loss_part, blocks, m->params,
                                                                     m->mom1, m->mom2, m->grad, loss_dst, lr, bc1,
                                                                     bc2);
    MLP_TRY(hipGetLastError());
    return SPH_OK;
}

/* Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from splitmix64, host side. */
void init_params(uint64_t seed, std::vector<float>& p) {
    p.assign(NP, 0.f);
    uint64_t st = seed;
    auto next = [&st]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    struct Seg { int off, n, fan; };
    const Seg segs[] = {{OW1, NH * NI, NI}, {OB1, NH, NI}, {OW2, NH * NH, NH}, {OB2, NH, NH}, {OW3, NH, NH}, {OB3, 1, NH}};
    for (const Seg& s : segs)
        for (int k = 0; k < s.n; ++k) p[s.off + k] = (float)((2.0 * next() - 1.0) / std::sqrt((double)s.fan));
}

}  // namespace

/* ------------------------------------------------------------- network API */
extern "C" {

int sph_mlp_create(uint64_t seed, sph_mlp** out) {
    if (!out) return SPH_E_ARG;
    *out = nullptr;
    sph_mlp* m = new sph_mlp();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete m;
        return SPH_E_NODEVICE;
    }
    m->device = dev;
    auto fail = [&](hipError_t e) {
        sph_mlp_destroy(m);
        return (int)e;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = dalloc(&m->params, NP)) != hipSuccess) return fail(e);
    if ((e = dalloc(&m->mom1, NP)) != hipSuccess) return fail(e);
    if ((e = dalloc(&m->mom2, NP)) != hipSuccess) return fail(e);
    if ((e = dalloc(&m->grad, NP)) != hipSuccess) return fail(e);
    if ((e = dalloc(&m->norm, SPH_MLP_NORM)) != hipSuccess) return fail(e);
    std::vector<float> p;
    init_params(seed, p);
    std::vector<float> norm(SPH_MLP_NORM, 0.f);
    for (int k = 0; k < NI; ++k) norm[NI + k] = 1.f;
    norm[2 * NI + 1] = 1.f;
    std::vector<float> zero(NP, 0.f);
    if ((e = hipMemcpy(m->params, p.data(), NP * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(m->mom1, zero.data(), NP * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(m->mom2, zero.data(), NP * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(m->norm, norm.data(), SPH_MLP_NORM * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    *out = m;
    return SPH_OK;
}

void sph_mlp_destroy(sph_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    dfree(m->params);
    dfree(m->mom1);
    dfree(m->mom2);
    dfree(m->norm);
    dfree(m->grad);
    dfree(m->part);
    dfree(m->loss_part);
    dfree(m->loss);
    dfree(m->scratch_rec);
    dfree(m->scratch_feats);
    dfree(m->scratch_d);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* sph_mlp_last_error(const sph_mlp* m) { return m ? m->err.c_str() : "null network"; }

int sph_mlp_get(sph_mlp* m, float* params, float* norm) {
    if (!m) return SPH_E_ARG;
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (params) MLP_TRY(hipMemcpy(params, m->params, NP * 4, hipMemcpyDeviceToHost));
    if (norm) MLP_TRY(hipMemcpy(norm, m->norm, SPH_MLP_NORM * 4, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_mlp_set(sph_mlp* m, const float* params, const float* norm) {
    if (!m) return SPH_E_ARG;
    if (norm)
        for (int k = 0; k < NI; ++k)
            if (!(norm[NI + k] > 0.f)) return mlp_fail(m, SPH_E_ARG, "feature scales must be > 0");
    if (norm && !(norm[2 * NI + 1] > 0.f)) return mlp_fail(m, SPH_E_ARG, "target scale must be > 0");
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (params) {
        MLP_TRY(hipMemcpy(m->params, params, NP * 4, hipMemcpyHostToDevice));
        MLP_TRY(hipMemset(m->mom1, 0, NP * 4));
        MLP_TRY(hipMemset(m->mom2, 0, NP * 4));
        m->adam_t = 0;
    }
    if (norm) MLP_TRY(hipMemcpy(m->norm, norm, SPH_MLP_NORM * 4, hipMemcpyHostToDevice));
    return SPH_OK;
}

int sph_mlp_forward(sph_mlp* m, const float* x, int64_t n, float* d_out) {
    if (!m || !x || !d_out || n < 0) return SPH_E_ARG;
    if (n == 0) return SPH_OK;
    int rc = ensure_scratch(m, n);
    if (rc != SPH_OK) return rc;
    std::vector<float> f4((size_t)n * 12);
    std::memcpy(f4.data(), x, (size_t)n * 12 * 4);
    MLP_TRY(hipMemcpyAsync(m->scratch_feats, f4.data(), (size_t)n * 48, hipMemcpyHostToDevice, m->stream));
    InferArgs a{};
    a.feats = m->scratch_feats;
    a.rows = n;
    a.params = m->params;
    a.norm = m->norm;
    a.d_out = m->scratch_d;
    k_infer<<<infer_blocks(n), WAVES * 64, 0, m->stream>>>(a);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpyAsync(d_out, m->scratch_d, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return SPH_OK;
}

int sph_mlp_loss_grad(sph_mlp* m, const float* x, const float* d, const float* w, int64_t n, float* loss,
                      float* grad) {
    if (!m || !x || !d || !loss || !grad || n <= 0 || n > (int64_t)1 << 24) return SPH_E_ARG;
    int rc = ensure_scratch(m, n);
    if (rc != SPH_OK) return rc;
    rc = ensure_loss(m, 1);
    if (rc != SPH_OK) return rc;
    std::vector<float> rec((size_t)n * 16, 0.f);
    for (int64_t s = 0; s < n; ++s) {
        std::memcpy(&rec[s * 16], &x[s * 12], 48);
        rec[s * 16 + 12] = d[s];
        rec[s * 16 + 13] = w ? w[s] : 1.f;
    }
    MLP_TRY(hipMemcpyAsync(m->scratch_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice, m->stream));
    const int batch = (int)((n + S - 1) / S * S);
    rc = run_step(m, m->scratch_rec, n, batch, n, 0, 0, 0, 0.f, m->loss);
    if (rc != SPH_OK) return rc;
    MLP_TRY(hipMemcpyAsync(grad, m->grad, NP * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return SPH_OK;
}

int sph_mlp_fit_norm(sph_mlp* m, const sph_samples* smp) {
    if (!m || !smp || smp->count <= 0) return SPH_E_ARG;
    if (smp->written) MLP_TRY(hipStreamWaitEvent(m->stream, smp->written, 0));
    constexpr int K = 2 * (NI + 1) + 1;
    const int blocks = 240;
    double* part = nullptr;
    MLP_TRY(dalloc(&part, (size_t)blocks * K));
    k_stats<<<blocks, 256, 0, m->stream>>>(smp->rec, smp->count, part);
    hipError_t e = hipGetLastError();
    std::vector<double> host((size_t)blocks * K);
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), part, host.size() * 8, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    dfree(part);
    MLP_TRY(e);
    double sum[K] = {};
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < K; ++k) sum[k] += host[(size_t)b * K + k];
    const double cnt = sum[K - 1];
    if (!(cnt > 0)) return mlp_fail(m, SPH_E_ARG, "no weighted samples");
    std::vector<float> norm(SPH_MLP_NORM);
    for (int k = 0; k <= NI; ++k) {
        const double mu = sum[k] / cnt;
        const double var = std::max(sum[NI + 1 + k] / cnt - mu * mu, 0.0);
        double sd = std::sqrt(var);
        if (!(sd > 1e-12)) sd = 1.0;   /* constant feature (e.g. one material only) */
        if (k < NI) {
            norm[k] = (float)mu;
            norm[NI + k] = (float)sd;
        } else {
            norm[2 * NI] = (float)mu;
            norm[2 * NI + 1] = (float)sd;
        }
    }
    MLP_TRY(hipMemcpy(m->norm, norm.data(), SPH_MLP_NORM * 4, hipMemcpyHostToDevice));
    return SPH_OK;
}

int sph_mlp_train(sph_mlp* m, const sph_samples* smp, int steps, int batch, float lr, uint64_t seed,
                  float* losses) {
    if (!m || !smp || steps < 0 || batch < S || batch % S != 0 || !(lr > 0.f)) return SPH_E_ARG;
    if (smp->count <= 0) return mlp_fail(m, SPH_E_ARG, "no samples");
    if (steps == 0) return SPH_OK;
    if (smp->written) MLP_TRY(hipStreamWaitEvent(m->stream, smp->written, 0));
    int rc = ensure_loss(m, steps);
    if (rc != SPH_OK) return rc;
    for (int st = 0; st < steps; ++st) {
        rc = run_step(m, smp->rec, smp->count, batch, batch, 1, seed, st, lr, m->loss + st);
        if (rc != SPH_OK) return rc;
    }
    if (losses) MLP_TRY(hipMemcpyAsync(losses, m->loss, (size_t)steps * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return SPH_OK;
}

/* ------------------------------------------------------------- samples API */
int sph_samples_create(int64_t capacity, sph_samples** out) {
    if (!out || capacity <= 0) return SPH_E_ARG;
    *out = nullptr;
    sph_samples* s = new sph_samples();
    (void)hipGetDevice(&s->device);
    hipError_t e = dalloc(&s->rec, (size_t)capacity * 4);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->written, hipEventDisableTiming);
    if (e != hipSuccess) {
        sph_samples_destroy(s);
        return (int)e;
    }
    s->cap = capacity;
    *out = s;
    return SPH_OK;
}

void sph_samples_destroy(sph_samples* s) {
    if (!s) return;
    if (s->written) {
        (void)hipEventSynchronize(s->written);
        (void)hipEventDestroy(s->written);
    }
    dfree(s->rec);
    delete s;
}

int64_t sph_samples_count(const sph_samples* s) { return s ? s->count : -1; }

int sph_samples_read(const sph_samples* s, int64_t first, int64_t count, float* dst) {
    if (!s || !dst || first < 0 || count < 0 || first + count > s->count) return SPH_E_ARG;
    if (hipEventSynchronize(s->written) != hipSuccess) return SPH_E_STATE;
    const hipError_t e = hipMemcpy(dst, s->rec + first * 4, (size_t)count * 64, hipMemcpyDeviceToHost);
    return e == hipSuccess ? SPH_OK : (int)e;
}

int sph_samples_append(sph_samples* s, const float* src, int64_t count) {
    if (!s || !src || count < 0 || s->count + count > s->cap) return SPH_E_ARG;
    if (hipEventSynchronize(s->written) != hipSuccess) return SPH_E_STATE;
    hipError_t e = hipMemcpy(s->rec + s->count * 4, src, (size_t)count * 64, hipMemcpyHostToDevice);
    if (e != hipSuccess) return (int)e;
    s->count += count;
    return SPH_OK;
}

int sph_samples_clear(sph_samples* s) {
    if (!s) return SPH_E_ARG;
    if (hipEventSynchronize(s->written) != hipSuccess) return SPH_E_STATE;
    s->count = 0;
    return SPH_OK;
}

}  // extern "C"

/* --------------------------------------------------- density-pass hooks */
int launch_features(sph_ctx* ctx, bool active_only = false) {
    if (!ctx->feats) SPH_HIP_TRY(hipMalloc((void**)&ctx->feats, sizeof(float4) * 3 * (size_t)ctx->n));
    FeatArgs a;
    a.list = nullptr;
    a.count = nullptr;
    if (active_only) {
        const int rc = launch_active_list(ctx);
        if (rc != SPH_OK) return rc;
        a.list = ctx->slot_list;
        a.count = ctx->slot_count;
    }
    a.n = ctx->n;
    a.sorted = ctx->pairs_a;
    a.cell_se = ctx->cell_se;
    a.mask = (1 << ctx->prm.cell_power) - 1;
    a.h = ctx->h;
    a.dhdt = ctx->dhdt;
    a.rho = ctx->rho;
    a.mat = ctx->mat;
    a.time = ctx->time;
    a.dtmin1 = ctx->prm.dt0 / (float)(1 << ctx->prm.r_max);
    a.feats = ctx->feats;
    k_features<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(a);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

/* Before the density kernel: features (when dumping or guessing) and the
 * learned guess per sorted slot. */
int launch_density_prologue(sph_ctx* ctx) {
    if (!ctx->dump && !ctx->learned) return SPH_OK;
    if (ctx->dump && ctx->dump->count + ctx->n > ctx->dump->cap)
        return ctx_fail(ctx, SPH_E_CAPACITY, "sample buffer full");
    /* with individual timesteps only the active particles iterate; without
     * a dump (which records every particle) only their guesses are needed */
    const bool active_only = !ctx->dump && ctx->prm.activity_rule == SPH_ACTIVE_README;
    int rc = launch_features(ctx, active_only);
    if (rc != SPH_OK) return rc;
    if (!ctx->learned) return SPH_OK;
    if (!ctx->guess) SPH_HIP_TRY(hipMalloc((void**)&ctx->guess, sizeof(float) * (size_t)ctx->n));
    sph_mlp* m = ctx->learned;
    InferArgs a{};
    a.feats = ctx->feats;
    a.rows = ctx->n;
    a.params = m->params;
    a.norm = m->norm;
    a.sorted = ctx->pairs_a;
    a.list = active_only ? ctx->slot_list : nullptr;
    a.count = active_only ? ctx->slot_count : nullptr;
    a.h = ctx->h;
    a.max_h = ctx->prm.max_smoothing_length;
    a.guess = ctx->guess;
    k_infer<<<infer_blocks(ctx->n), WAVES * 64, 0, ctx->stream>>>(a);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

/* After the density kernel: append the samples. */
int launch_density_epilogue(sph_ctx* ctx) {
    if (!ctx->dump) return SPH_OK;
    sph_samples* s = ctx->dump;
    k_dump<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, ctx->feats, ctx->pairs_a, ctx->h, ctx->active,
                                                          s->rec + s->count * 4);
    SPH_HIP_TRY(hipGetLastError());
    SPH_HIP_TRY(hipEventRecord(s->written, ctx->stream));
    s->count += ctx->n;
    return SPH_OK;
}

extern "C" {

int sph_set_sample_dump(sph_ctx* ctx, sph_samples* s) {
    if (!ctx) return SPH_E_ARG;
    ctx->dump = s;
    return SPH_OK;
}

int sph_set_learned_guess(sph_ctx* ctx, sph_mlp* m) {
    if (!ctx) return SPH_E_ARG;
    if (m) SPH_HIP_TRY(hipStreamSynchronize(m->stream));
    ctx->learned = m;
    return SPH_OK;
}

int sph_features(sph_ctx* ctx, float* dst) {
    if (!ctx || !dst) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    if (!ctx->have_prev) return ctx_fail(ctx, SPH_E_STATE, "features before any sort");
    int rc = launch_features(ctx);
    if (rc != SPH_OK) return rc;
    std::vector<float> f((size_t)ctx->n * 12);
    std::vector<uint2> pairs(ctx->n);
    SPH_HIP_TRY(hipMemcpyAsync(f.data(), ctx->feats, f.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_TRY(hipMemcpyAsync(pairs.data(), ctx->pairs_a, pairs.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < ctx->n; ++k) std::memcpy(dst + (size_t)pairs[k].y * 12, &f[(size_t)k * 12], 48);
    return SPH_OK;
}

}  // extern "C"

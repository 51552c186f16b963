// Surrogate network of the delayed-acceptance rotor move, and its training kernels.
//
// Network: 4 inputs -> 128 -> 128 -> 128 -> 1, SiLU after each hidden layer.  Inputs are the
// atom-rotor pair geometry in the rotor's body frame (R, ux, uy, uz) normalised by the training
// set's mean and spread; the output, scaled back, is the pair energy in Kelvin.
//
// Master weights are FP32 in one flat vector `theta` (MLP_NW floats, layout MLP_OFF_*); the
// gradient has the same layout.  Training runs batched GEMMs on the matrix cores
// (v_mfma_f32_32x32x8bf16_1k: BF16 operands converted at load, FP32 accumulation) and small
// elementwise kernels; the sampler reads a packed copy (MlpPacked: BF16 matrices in the lane
// order of the matrix instructions, FP32 biases), staged in LDS by the kernels that use it.
//
// Inference layout.  One wave evaluates 32 rows as H^T = W X^T, hidden units on the MFMA's M
// axis and rows on its N axis.  D layout of the 32x32 MFMAs: element i of lane l is
// D[8*(i/4)+4*(l/32)+i%4][l%32].  Layer 1 (K = 4, padded to 8) uses the 32x32x8 MFMA
// (A: lane l holds A[l%32][4*(l/32)+j], j < 4; B: B[4*(l/32)+j][l%32]).  The 128 -> 128 layers
// use gfx950's 32x32x16 MFMA (A: A[l%32][8*(l/32)+j], j < 8; B: B[8*(l/32)+j][l%32]) with the
// contraction index permuted: k-block kb takes, on lane l, elements 8*(kb%2)..8*(kb%2)+7 of the
// previous layer's output tile kb/2, i.e. hidden units 16*kb + 8*(j/4) + 4*(l/32) + j%4.  The
// packed weights follow the same permutation, so activations stay in registers from layer to
// layer and every A operand is one 16-byte LDS read.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

constexpr int MLP_IN = 4;
constexpr int MLP_H = 128;
constexpr int MLP_OFF_W1 = 0;                                // [128][4]
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;      // [128]
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;               // [128][128] (out, in)
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;               // [128][128]
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_H;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_H;               // [128]
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_H;               // [1]
constexpr int MLP_NW = MLP_OFF_B4 + 1;                       // 33793
// FLOP of one evaluation: 2*(4*128 + 128*128 + 128*128 + 128)
constexpr double MLP_FLOP = 2.0 * (MLP_IN * MLP_H + 2 * MLP_H * MLP_H + MLP_H);

struct alignas(16) MlpPacked {
    uint16_t w1[4 * 64 * 4];      // layer 1 A operands [tile][lane][4] (inputs 4..7 zero)
    uint16_t w2[4 * 8 * 64 * 8];  // layer 2 A operands [tile][k-block][lane][8], permuted k
    uint16_t w3[4 * 8 * 64 * 8];
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], w4[MLP_H];
    float b4;
    float xm[MLP_IN], xinv[MLP_IN];  // normalised input = (x - xm) * xinv
    float ym, ys;                    // energy = ys * output + ym
};
static_assert(sizeof(MlpPacked) % 16 == 0, "staged with 16-byte copies");

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short mlp_u16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float x) {  // round to nearest even
    union { float f; uint32_t u; } c;
    c.f = x;
    uint32_t u = c.u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// SiLU z * sigmoid(z) from the hardware exponential and reciprocal (v_exp_f32, v_rcp_f32): the
// activation is the bulk of the vector work per row, and BF16 operands bound the accuracy anyway.
__device__ __forceinline__ float mlp_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float mlp_silu(float z) { return z * mlp_sigmoid(z); }

__device__ __forceinline__ mlp_f32x16 mlp_mfma(mlp_bf16x4 a, mlp_bf16x4 b, mlp_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ mlp_f32x16 mlp_mfma16(mlp_bf16x8 a, mlp_bf16x8 b, mlp_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// unit of the output element i of lane half hi in tile mb
__device__ __forceinline__ int mlp_unit(int mb, int i, int hi) { return mb * 32 + 8 * (i >> 2) + 4 * hi + (i & 3); }

__device__ __forceinline__ mlp_bf16x4 mlp_load4(const uint16_t* p) {  // 4 BF16, 8-byte aligned
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    mlp_bf16x4 r;
    r[0] = (short)(v.x & 0xffffu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffffu);
    r[3] = (short)(v.y >> 16);
    return r;
}

// One hidden 128 -> 128 layer on D-layout activations h (in place): bias and SiLU applied.
// w: the layer's packed A operands (in LDS).
__device__ __forceinline__ void mlp_hidden(const uint16_t* __restrict__ w, const float* __restrict__ b,
                                           mlp_f32x16 h[4], int l) {
    const int hi = l >> 5;
    mlp_bf16x8 bin[8];
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        mlp_u16x8 t;
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = f2bf(h[kb >> 1][8 * (kb & 1) + j]);
        bin[kb] = __builtin_bit_cast(mlp_bf16x8, t);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        mlp_f32x16 acc;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const uint4 a = *reinterpret_cast<const uint4*>(w + (((mb * 8 + kb) * 64 + l) << 3));
            acc = mlp_mfma16(__builtin_bit_cast(mlp_bf16x8, a), bin[kb], acc);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = mlp_silu(acc[i] + b[mlp_unit(mb, i, hi)]);
        h[mb] = acc;
    }
}

// The whole wave evaluates 32 rows.  Lane l < 32 passes the normalised inputs of row l in
// f[0..3] (lanes 32..63 pass anything).  Every lane returns the raw network output of row l%32.
// W: the packed network staged in LDS.
__device__ __forceinline__ float mlp_rows32(const MlpPacked* __restrict__ W, const float f[MLP_IN]) {
    const int l = threadIdx.x & 63, hi = l >> 5;
    mlp_bf16x4 bx;
#pragma unroll
    for (int j = 0; j < 4; ++j) bx[j] = hi ? (short)0 : (short)f2bf(f[j]);
    mlp_f32x16 h[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        mlp_f32x16 acc;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
        acc = mlp_mfma(mlp_load4(W->w1 + ((mb * 64 + l) << 2)), bx, acc);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = mlp_silu(acc[i] + W->b1[mlp_unit(mb, i, hi)]);
        h[mb] = acc;
    }
    mlp_hidden(W->w2, W->b2, h, l);
    mlp_hidden(W->w3, W->b3, h, l);
    float part = 0.0f;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 16; ++i) part += W->w4[mlp_unit(mb, i, hi)] * h[mb][i];
    part += __shfl_xor(part, 32);
    return part + W->b4;
}

// Copies the packed network into LDS (all lanes of the block take part; caller synchronises).
__device__ __forceinline__ void mlp_stage(const MlpPacked* __restrict__ g, MlpPacked* s) {
    const uint4* src = reinterpret_cast<const uint4*>(g);
    uint4* dst = reinterpret_cast<uint4*>(s);
    for (int i = threadIdx.x; i < (int)(sizeof(MlpPacked) / 16); i += blockDim.x) dst[i] = src[i];
}

// Pair geometry in the rotor's body frame: distance and the direction cosines of the partner
// on the body axes (the axes vcord reads from the rotation matrix; z is normalised by its length).
__device__ __forceinline__ void mlp_pair_features(const double m[9], double zn, double cx, double cy, double cz,
                                                  double px, double py, double pz, float out[MLP_IN]) {
    const double rx = px - cx, ry = py - cy, rz = pz - cz;
    const double rn = sqrt(rx * rx + ry * ry + rz * rz);
    const double inv = rn > 0.0 ? 1.0 / rn : 0.0;
    out[0] = (float)rn;
    out[1] = (float)((m[0] * rx + m[1] * ry + m[2] * rz) * inv);
    out[2] = (float)((m[3] * rx + m[4] * ry + m[5] * rz) * inv);
    out[3] = (float)((m[6] * rx + m[7] * ry + m[8] * rz) * inv / zn);
}

// ------------------------------------------------------------------------------ training
// C[M x N] = op(A) op(B) (+ C when accumulate), FP32 in memory, BF16 operands, FP32 sums.
// op(A)[m][k] = ta ? A[k*lda+m] : A[m*lda+k];  op(B)[k][n] = tb ? B[n*ldb+k] : B[k*ldb+n].
// One wave per 32x32 tile of C; grid (ceil(N/32), ceil(M/32)); out-of-range operands read 0.
__global__ __launch_bounds__(64) void k_mlp_gemm(int M, int N, int K, const float* __restrict__ A, int lda, int ta,
                                                 const float* __restrict__ B, int ldb, int tb, float* __restrict__ C,
                                                 int ldc, int accumulate) {
    const int l = threadIdx.x, hi = l >> 5, r = l & 31;
    const int m = blockIdx.y * 32 + r, n = blockIdx.x * 32 + r;
    mlp_f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 8) {
        mlp_bf16x4 a, b;
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * hi + j;
            float av = 0.0f, bv = 0.0f;
            if (m < M && k < K) av = ta ? A[(long long)k * lda + m] : A[(long long)m * lda + k];
            if (n < N && k < K) bv = tb ? B[(long long)n * ldb + k] : B[(long long)k * ldb + n];
            a[j] = (short)f2bf(av);
            b[j] = (short)f2bf(bv);
        }
        acc = mlp_mfma(a, b, acc);
    }
    for (int i = 0; i < 16; ++i) {
        const int mm = blockIdx.y * 32 + 8 * (i >> 2) + 4 * hi + (i & 3);
        if (mm < M && n < N) {
            float* c = C + (long long)mm * ldc + n;
            *c = accumulate ? *c + acc[i] : acc[i];
        }
    }
}

// Z[rows][cols] += bias; H = SiLU(Z) (act) or Z.
__global__ void k_mlp_bias_act(float* __restrict__ Z, float* __restrict__ H, const float* __restrict__ b, int rows,
                               int cols, int act) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const float z = Z[i] + b[i % cols];
    Z[i] = z;
    H[i] = act ? mlp_silu(z) : z;
}

// dZ = dH * SiLU'(Z), in place on dH.
__global__ void k_mlp_act_back(float* __restrict__ dH, const float* __restrict__ Z, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i], s = mlp_sigmoid(z);
    dH[i] *= s * (1.0f + z * (1.0f - s));
}

// out[c] = sum over rows of X[r][c]; one block per column.
__global__ __launch_bounds__(256) void k_mlp_colsum(const float* __restrict__ X, int rows, int cols,
                                                    float* __restrict__ out) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += X[(long long)r * cols + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// Mean-squared-error loss L = (1/2n) sum (out - y)^2: dout = (out - y)/n, loss += partial.
__global__ __launch_bounds__(256) void k_mlp_mse(const float* __restrict__ out, const float* __restrict__ y, int n,
                                                 float* __restrict__ dout, double* __restrict__ loss) {
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const float d = out[i] - y[i];
        dout[i] = d / (float)n;
        e = 0.5 * (double)d * (double)d / (double)n;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// theta -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void k_mlp_adam(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int n, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// Gathers a batch: X[b] = data[perm[start+b]] (4 inputs), y[b] likewise.
__global__ void k_mlp_gather(const float* __restrict__ X, const float* __restrict__ y, const int* __restrict__ perm,
                             int start, int nb, float* __restrict__ bx, float* __restrict__ by) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int s = perm[start + b];
    for (int j = 0; j < MLP_IN; ++j) bx[b * MLP_IN + j] = X[(long long)s * MLP_IN + j];
    by[b] = y[s];
}

// Packs theta and the normalisation into the sampler's copy, weights in the lane order of the
// matrix instructions (see the layout note at the top of this file).
__global__ void k_mlp_pack(const float* __restrict__ th, const float* __restrict__ norm, MlpPacked* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * 64 * 4) {  // [mb][lane][j]: W1[mb*32 + lane%32][4*(lane/32) + j]
        const int j = i & 3, lane = (i >> 2) & 63, mb = i >> 8;
        const int o = mb * 32 + (lane & 31), k = 4 * (lane >> 5) + j;
        p->w1[i] = k < MLP_IN ? f2bf(th[MLP_OFF_W1 + o * MLP_IN + k]) : (uint16_t)0;
    }
    if (i < MLP_H * MLP_H) {  // [mb][kb][lane][j]: W[mb*32 + lane%32][16*kb + 8*(j/4) + 4*(lane/32) + j%4]
        const int j = i & 7, lane = (i >> 3) & 63, kb = (i >> 9) & 7, mb = i >> 12;
        const int o = mb * 32 + (lane & 31), u = 16 * kb + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        p->w2[i] = f2bf(th[MLP_OFF_W2 + o * MLP_H + u]);
        p->w3[i] = f2bf(th[MLP_OFF_W3 + o * MLP_H + u]);
    }
    if (i < MLP_H) {
        p->b1[i] = th[MLP_OFF_B1 + i];
        p->b2[i] = th[MLP_OFF_B2 + i];
        p->b3[i] = th[MLP_OFF_B3 + i];
        p->w4[i] = th[MLP_OFF_W4 + i];
    }
    if (i == 0) {
        p->b4 = th[MLP_OFF_B4];
        for (int j = 0; j < MLP_IN; ++j) {
            p->xm[j] = norm[j];
            p->xinv[j] = 1.0f / norm[MLP_IN + j];
        }
        p->ym = norm[2 * MLP_IN];
        p->ys = norm[2 * MLP_IN + 1];
    }
}

// Batched inference on raw inputs (the sampler's code path): out = ys * net((x - xm)/xs) + ym.
// One wave per 32 rows; the network is staged in LDS (dynamic, sizeof(MlpPacked) bytes).
__global__ __launch_bounds__(256) void k_mlp_infer(const MlpPacked* __restrict__ Wg, const float* __restrict__ X,
                                                   int n, float* __restrict__ out) {
    extern __shared__ uint4 mlp_lds[];
    MlpPacked* W = reinterpret_cast<MlpPacked*>(mlp_lds);
    mlp_stage(Wg, W);
    __syncthreads();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    const int row = wave * 32 + (l & 31);
    if (wave * 32 >= n) return;  // whole waves leave together
    float f[MLP_IN];
    for (int j = 0; j < MLP_IN; ++j) f[j] = row < n ? (X[(long long)row * MLP_IN + j] - W->xm[j]) * W->xinv[j] : 0.0f;
    const float o = mlp_rows32(W, f);
    if (l < 32 && row < n) out[row] = W->ys * o + W->ym;
}

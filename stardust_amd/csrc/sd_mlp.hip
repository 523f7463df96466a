// sd_mlp.hip -- learned sub-cell force correction (no reference counterpart).
//
// MLP 30 -> 128 -> 128 -> 128 -> 3 with GELU (tanh form), BF16 weights, FP32
// accumulation on the gfx950 matrix cores. Training runs BF16 activations
// (v_mfma_f32_32x32x16_bf16); the fused inference kernel holds the same BF16
// weight values as FP16 (exact) with FP16 activations
// (v_mfma_f32_32x32x16_f16). Inputs per particle: the 27 CIC-node densities
// around it (log1p of particles per cell) and its offset from the nearest node
// (x2); output: the direct-sum minus PM acceleration in units of
// a_unit * outScale, a_unit = gravity * G / N (one particle at one cell).
//
// MFMA operand layout used throughout (32x32x16): lane l holds row (l % 32) of
// A and column (l % 32) of B with 8 consecutive K values starting at
// 8 * (l / 32); accumulator element j of lane l is row 8*(j/4) + 4*(l/32) +
// (j % 4), column l % 32.
#include <climits>

#include "sd_device.h"
#include "sd_internal.h"
#include "sd_mlp.h"

namespace sd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// GELU, tanh form: 0.5 u (1 + T(u)), T(u) = tanh(sqrt(2/pi) (u + 0.044715 u^3)).
// Evaluated without transcendental instructions (quarter rate on CDNA): for
// |u| <= 4, T(u) = u R(u^2) with R a degree-6 polynomial (weighted least
// squares fit to T(u)/u, constrained to T(4) = 1); the clamped input makes
// T = sign(u) beyond. Written as u (0.5 + uc R(uc^2) / 2): |error| <= 2.2e-4 for
// every u (FP32 evaluation), below the BF16 rounding of the activations it
// feeds. 10 full-rate VALU ops.
__device__ __forceinline__ float gelu_f(float u) {
    const float uc = __builtin_amdgcn_fmed3f(u, -4.0f, 4.0f);
    const float t = uc * uc;
    float r = 2.3469173271450927e-08f;  // coefficients of R / 2
    r = fmaf(r, t, -1.6407460634582094e-06f);
    r = fmaf(r, t, 4.893512959824875e-05f);
    r = fmaf(r, t, -0.0008244823548011482f);
    r = fmaf(r, t, 0.00881875865161419f);
    r = fmaf(r, t, -0.06459682434797287f);
    r = fmaf(r, t, 0.39771363139152527f);
    return u * fmaf(uc, r, 0.5f);
}
__device__ __forceinline__ float gelu_d(float u) {
    const float c = 0.7978845608028654f;
    const float x = c * fmaf(0.044715f * u, u * u, u);
    const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * x));  // 0.5 (1 + tanh x)
    const float t = 2.0f * sg - 1.0f;
    return sg + 0.5f * u * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * u * u);
}

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---- inference element type --------------------------------------------------
// The fused inference kernel runs its matrix products on FP16 operands: the
// BF16 weights convert to FP16 exactly (8 significant bits fit in 11; every
// |w| in [2^-17, 65504] keeps its value), the activations gain 3 mantissa bits
// over BF16, and GELU evaluates on packed FP16 pairs (2 values per VALU lane
// op), which is what bounds this kernel. SD_MLP_F16=0 keeps BF16 activations
// and the FP32 GELU.
#ifndef SD_MLP_F16
#define SD_MLP_F16 1
#endif
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
#if SD_MLP_F16
typedef _Float16 inf_t;
typedef f16x8 infx8;
__device__ __forceinline__ f32x16 mfma_inf(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
#else
typedef __bf16 inf_t;
typedef bf16x8 infx8;
__device__ __forceinline__ f32x16 mfma_inf(bf16x8 a, bf16x8 b, f32x16 c) { return mfma32(a, b, c); }
#endif

// GELU on FP16 values, 8 at a time, on halved inputs: the weights and biases
// that feed an activation are stored halved (exact: powers of two), so the
// accumulator holds v = u / 2, and the result is gelu(u) / 2 = v phi(u), which
// the next layer's unscaled weights (output layer: doubled) take as is. With
// x = v^2 - 2 (in [-2, 2] for |u| <= 4), phi = clamp(0.5 + v R(x), 0, 1), R of
// degree 6 (a minimax-weighted least-squares fit of
// (0.5 tanh(sqrt(2/pi) (u + 0.044715 u^3))) / v with R(2) = 1/4 and a positive
// leading coefficient: beyond |u| = 4, R(x) >= 1/4, so the clamp, which the
// compiler folds into the packed FMA, saturates phi at exactly 1 and 0). Over
// every finite FP16 input the result is within 2.4e-3 of the tanh-form GELU
// (mean 6e-5 for |u| <= 4); 9 packed instructions per pair of values. Each
// vector operation is four independent packed instructions, which keeps
// dependent packed operations apart (gfx950 needs a wait state between a packed
// FP16 result and its next use). Inputs are finite: the accumulators convert
// with round-toward-zero, which saturates at 65504.
__device__ __forceinline__ f16x8 gelu_half_h8(f16x8 v) {
    const f16x8 x = __builtin_elementwise_fma(v, v, (f16x8)(-2.0f16));
    f16x8 r = (f16x8)(0.00023937225341796875f16);
    r = __builtin_elementwise_fma(r, x, (f16x8)(-0.0011377334594726562f16));
    r = __builtin_elementwise_fma(r, x, (f16x8)(0.0028095245361328125f16));
    r = __builtin_elementwise_fma(r, x, (f16x8)(-0.008453369140625f16));
    r = __builtin_elementwise_fma(r, x, (f16x8)(0.02777099609375f16));
    r = __builtin_elementwise_fma(r, x, (f16x8)(-0.084716796875f16));
    r = __builtin_elementwise_fma(r, x, (f16x8)(0.35205078125f16));
    const f16x8 ph = __builtin_elementwise_min(
        __builtin_elementwise_max(__builtin_elementwise_fma(v, r, (f16x8)(0.5f16)), (f16x8)(0.0f16)), (f16x8)(1.0f16));
    return v * ph;
}

// Scale of the stored inference weights: layer 1 weights and the biases of
// layers 1-3 halved, the output layer's weights doubled (FP16 path only).
#if SD_MLP_F16
constexpr float INF_IN_SCALE = 0.5f;
constexpr float INF_OUT_SCALE = 2.0f;
#else
constexpr float INF_IN_SCALE = 1.0f;
constexpr float INF_OUT_SCALE = 1.0f;
#endif

// Eight FP32 accumulators -> FP16 (round toward zero, 2 per instruction).
__device__ __forceinline__ f16x8 to_f16x8(const float* v) {
    f16x8 h;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        const f16x2 p = __builtin_bit_cast(f16x2, __builtin_amdgcn_cvt_pkrtz(v[e], v[e + 1]));
        h[e] = p[0];
        h[e + 1] = p[1];
    }
    return h;
}

// ---- generic NT GEMM: C[M x N] = A[M x K] * B^T, A row-major (lda), Bt = B^T
// row-major [N x K] (ldb); K % 16 == 0, M % 32 == 0, N % 32 == 0. A block is
// 2 x 2 waves of 32 x 32 tiles; grid.z splits K into equal slices.
struct GemmArgs {
    const __bf16* A;
    const __bf16* Bt;
    int M, N, K, lda, ldb;
    int ksplit;          // K slice per grid.z
    // epilogue
    float* C;            // fp32 out [M x N] (+ z * M * N for split-K partials)
    const float* bias;   // [N]
    float* U;            // pre-activation out (fp32) [M x N]
    __bf16* H;           // activation out (bf16) [M x ldh]
    __bf16* Ht;          // activation out transposed (bf16) [N x ldht]
    int ldh, ldht;
    const float* Uin;    // pre-activation in for the GELU derivative
    float* Df;           // fp32 delta out [M x N]
};

enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_DGELU = 3 };

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nt(GemmArgs g) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    if (m0 >= g.M || n0 >= g.N) return;  // whole wave (no block-level sync below)
    const int k0 = blockIdx.z * g.ksplit, k1 = min(g.K, k0 + g.ksplit);
    f32x16 acc = {};
    const __bf16* ap = g.A + (size_t)(m0 + (l & 31)) * g.lda + 8 * (l >> 5);
    const __bf16* bp = g.Bt + (size_t)(n0 + (l & 31)) * g.ldb + 8 * (l >> 5);
    for (int k = k0; k < k1; k += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(ap + k);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(bp + k);
        acc = mfma32(a, b, acc);
    }
    const int col = n0 + (l & 31);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int row = m0 + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
        const size_t o = (size_t)row * g.N + col;
        if (EPI == EPI_STORE) {
            g.C[(size_t)blockIdx.z * g.M * g.N + o] = acc[j];
        } else if (EPI == EPI_BIAS) {
            g.C[o] = acc[j] + g.bias[col];
        } else if (EPI == EPI_BIAS_GELU) {
            const float u = acc[j] + g.bias[col];
            g.U[o] = u;
            const __bf16 h = (__bf16)gelu_f(u);
            g.H[(size_t)row * g.ldh + col] = h;
            g.Ht[(size_t)col * g.ldht + row] = h;
        } else {  // EPI_DGELU: delta = acc * gelu'(U)
            const float d = acc[j] * gelu_d(g.Uin[o]);
            g.Df[o] = d;
            g.H[(size_t)row * g.ldh + col] = (__bf16)d;
            g.Ht[(size_t)col * g.ldht + row] = (__bf16)d;
        }
    }
}

static void gemm(const GemmArgs& g, int epi, int splits, hipStream_t st) {
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, splits);
    switch (epi) {
        case EPI_STORE: hipLaunchKernelGGL(k_gemm_nt<EPI_STORE>, grid, dim3(256), 0, st, g); break;
        case EPI_BIAS_GELU: hipLaunchKernelGGL(k_gemm_nt<EPI_BIAS_GELU>, grid, dim3(256), 0, st, g); break;
        case EPI_BIAS: hipLaunchKernelGGL(k_gemm_nt<EPI_BIAS>, grid, dim3(256), 0, st, g); break;
        default: hipLaunchKernelGGL(k_gemm_nt<EPI_DGELU>, grid, dim3(256), 0, st, g); break;
    }
}

// Neuron held at position q = 16 s + 8 g + e of the register-resident
// activation vector: accumulator element j = 8 (s % 2) + e of out-tile s / 2
// on lane group g is row 8 (j / 4) + 4 g + (j % 4) of that tile.
__host__ __device__ __forceinline__ int perm_k(int q) {
    const int s = q >> 4, g = (q >> 3) & 1, e = q & 7;
    const int t = s >> 1, h = s & 1;
    return 32 * t + 8 * (2 * h + (e >> 2)) + 4 * g + (e & 3);
}

// ---- small elementwise kernels ------------------------------------------------
// FP32 master parameters -> padded BF16 copies used by the matrix cores:
// W1p [128 x 32] (inputs padded), W2, W3 [128 x 128], W4p [32 x 128] (outputs
// padded with zero rows) and transposes W2t, W3t, W4pt [128 x 32].
__global__ void k_mlp_pack(const float* __restrict__ W, MlpDev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = W;
    const float* W2 = W1 + MLP_H * MLP_IN;
    const float* W3 = W2 + MLP_H * MLP_H;
    const float* W4 = W3 + MLP_H * MLP_H;
    if (i < MLP_H * MLP_INP) {  // W1p
        const int o = i / MLP_INP, k = i % MLP_INP;
        d.W1p[i] = (__bf16)(k < MLP_IN ? W1[o * MLP_IN + k] : 0.f);
    }
    if (i < MLP_H * MLP_H) {
        const int o = i / MLP_H, k = i % MLP_H;
        d.W2[i] = (__bf16)W2[i];
        d.W3[i] = (__bf16)W3[i];
        d.W2t[k * MLP_H + o] = (__bf16)W2[i];
        d.W3t[k * MLP_H + o] = (__bf16)W3[i];
    }
    if (i < MLP_OUTP * MLP_H) {  // W4p [32 x 128], W4pt [128 x 32]
        const int o = i / MLP_H, k = i % MLP_H;
        const float v = o < MLP_OUT ? W4[o * MLP_H + k] : 0.f;
        d.W4p[i] = (__bf16)v;
        d.W4pt[k * MLP_OUTP + o] = (__bf16)v;
    }
    // Inference copies with the input (column) order of layers 2-4 permuted to
    // the accumulator order of the layer before: column q = 16 s + 8 g + e holds
    // neuron perm_k(q) (see k_mlp_infer); biases of layers 1-3 in that order.
    if (i < MLP_H * MLP_H) {
        const int o = i / MLP_H, q = i % MLP_H;
        const int src = perm_k(q);
        d.W2q[i] = (__bf16)W2[o * MLP_H + src];
        d.W3q[i] = (__bf16)W3[o * MLP_H + src];
    }
    if (i < MLP_OUTP * MLP_H) {
        const int o = i / MLP_H, q = i % MLP_H;
        d.W4q[i] = (__bf16)(o < MLP_OUT ? W4[o * MLP_H + perm_k(q)] : 0.f);
    }
    if (i < 3 * MLP_H) {
        const int l = i / MLP_H, q = i % MLP_H;
        d.bq[i] = d.b[l * MLP_H + perm_k(q)];
    }
}

// X [n x 30] fp32 -> Xb [nPad x 32] bf16 and Xt [32 x nPad]; rows >= n are zero.
__global__ void k_mlp_input(const float* __restrict__ X, int n, int nPad, __bf16* Xb, __bf16* Xt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPad * MLP_INP) return;
    const int r = i / MLP_INP, k = i % MLP_INP;
    const float v = (r < n && k < MLP_IN) ? X[(size_t)r * MLP_IN + k] : 0.f;
    Xb[i] = (__bf16)v;
    Xt[(size_t)k * nPad + r] = (__bf16)v;
}

// dY = (Y - T) / n on the 3 live outputs (padded to 32 columns), loss terms.
__global__ void k_mlp_dloss(const float* __restrict__ Y, const float* __restrict__ T, int n, int nPad, float invn,
                            __bf16* dY, __bf16* dYt, float* dYf, double* lossPart) {
    __shared__ double s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double lsum = 0.0;
    if (i < nPad * MLP_OUTP) {
        const int r = i / MLP_OUTP, c = i % MLP_OUTP;
        float d = 0.f;
        if (r < n && c < MLP_OUT) {
            const float e = Y[(size_t)r * MLP_OUTP + c] - T[(size_t)r * MLP_OUT + c];
            lsum = 0.5 * (double)e * (double)e * (double)invn;
            d = e * invn;
        }
        dY[i] = (__bf16)d;
        dYt[(size_t)c * nPad + r] = (__bf16)d;
        dYf[i] = d;
    }
    s[threadIdx.x] = lsum;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossPart[blockIdx.x] = s[0];
}

// Column sums of an fp32 [rows x cols] matrix (bias gradient), fixed order.
__global__ void k_colsum(const float* __restrict__ D, int rows, int cols, float* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
    for (int r = 0; r < rows; ++r) s += D[(size_t)r * cols + c];
    out[c] = (float)s;
}

// Sum of split-K partials [splits x M x N] into out (fixed order); optional
// row / column crop to the unpadded weight shape (out is [Mo x No], row-major,
// taking columns < No of partial rows < Mo).
__global__ void k_reduce_splits(const float* __restrict__ part, int splits, int M, int N, int Mo, int No,
                                float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Mo * No) return;
    const int r = i / No, c = i % No;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += part[((size_t)z * M + r) * N + c];
    out[i] = s;
}

// Adam on the FP32 master parameters.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- features -----------------------------------------------------------------
// Nearest node c = clamp(floor(p G), 0, G-1) (the node closest to p G - 0.5);
// features = log(1 + particles per cell) at c + {-1,0,1}^3 (x fastest, edge
// clamp / periodic wrap) and 2 (p G - 0.5 - c).
__device__ __forceinline__ void particle_features(float4 p, const long long* __restrict__ rho, int G,
                                                  int periodic, float* f) {
    const float fG = (float)G;
    const int cx = min(max((int)floorf(p.x * fG), 0), G - 1);
    const int cy = min(max((int)floorf(p.y * fG), 0), G - 1);
    const int cz = min(max((int)floorf(p.z * fG), 0), G - 1);
    int q = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = wrap_or_clamp(cx + dx, G, periodic), y = wrap_or_clamp(cy + dy, G, periodic),
                          z = wrap_or_clamp(cz + dz, G, periodic);
                const long long r = rho[((size_t)z * G + y) * G + x];
                f[q++] = __logf(1.0f + (float)r * (1.0f / (float)SD_FIX_ONE));
            }
    f[27] = 2.0f * (p.x * fG - 0.5f - (float)cx);
    f[28] = 2.0f * (p.y * fG - 0.5f - (float)cy);
    f[29] = 2.0f * (p.z * fG - 0.5f - (float)cz);
}

__global__ void k_features(const float4* __restrict__ pos, const int32_t* __restrict__ targets, int nt,
                           const long long* __restrict__ rho, int G, int periodic, float* __restrict__ X) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    float f[MLP_IN];
    particle_features(pos[targets[t]], rho, G, periodic, f);
    for (int k = 0; k < MLP_IN; ++k) X[(size_t)t * MLP_IN + k] = f[k];
}

// ---- fused inference ------------------------------------------------------------
// Persistent blocks (one per CU); the weights (W1p and the permuted W2q, W3q,
// W4q; padded rows, as FP16) and biases live in LDS for the block's lifetime.
// Each wave advances tiles of 32 particles: Y^T = W X^T with the
// weights as the A operand (from LDS) and
// the activations as the B operand, which never leave registers: accumulator
// elements 8h .. 8h+7 of out-tile t are, after bias and GELU, exactly the B
// fragment of k-step 2t + h of the next layer because that layer's weight
// columns are stored in the same (permuted) order. Output: the correction as
// a float4 per particle (inactive particles get 0).
// Biases of layers 1-3 inside the matrix products (FP16 inference): each bias
// is split into FP16 halves bh = fp16(b), bl = fp16(b - bh) that multiply inputs
// fixed at 1 -- layer 1's two padding inputs (30, 31), and one more k-step of
// 16 for layers 2 and 3 (columns 128, 129 of their rows) -- so the accumulators
// start from zero and the biases cost no LDS reads (48 of the tile's 128 reads).
#ifndef SD_MLP_BIAS_MFMA
#define SD_MLP_BIAS_MFMA 1
#endif
constexpr bool BIAS_MM = SD_MLP_F16 && SD_MLP_BIAS_MFMA;
constexpr int KH = MLP_H + (BIAS_MM ? 16 : 0);  // k extent of layers 2 and 3
constexpr int W1_LD = MLP_INP + 8;  // LDS row pitches (2-byte elements); +8 spreads rows over banks
constexpr int WH_LD = KH + 8;
#ifndef SD_INF_WAVES
#define SD_INF_WAVES 16
#endif
// Waves per block. The weights take 90 KB of LDS, so one block runs per CU;
// 16 waves are the 4 per SIMD that 124 registers allow (12: 5% slower; two
// tiles per wave sharing each weight fragment read, in 8 waves: 12% slower).
constexpr int INF_WAVES = SD_INF_WAVES;

// One layer, Y^T = W X^T: the accumulators start from the layer's biases, so
// no add is needed afterwards (BIAS false starts from zero). The biases
// are staged per (row tile t, lane half g) as 16 contiguous floats in
// accumulator order: element j of acc[t] is neuron 32 t + 8 (j / 4) + 4 g + (j % 4),
// i.e. bq[32 t + 16 (j / 8) + 8 g + (j % 8)], so one aligned 64-byte read fills
// the accumulator registers.
template <int K, int LDW, int NT, bool BIAS>
__device__ __forceinline__ void mlp_layer(const inf_t* __restrict__ Wl, const infx8* B, f32x16* acc, int l,
                                          const float* __restrict__ bias) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        if constexpr (BIAS)
            acc[t] = *reinterpret_cast<const f32x16*>(bias + (2 * t + (l >> 5)) * 16);
        else
            acc[t] = f32x16{};
    }
#pragma unroll
    for (int s = 0; s < K / 16; ++s) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const infx8 a = *reinterpret_cast<const infx8*>(Wl + (t * 32 + (l & 31)) * LDW + 16 * s + 8 * (l >> 5));
            acc[t] = mfma_inf(a, B[s], acc[t]);
        }
    }
}

// GELU of a layer's accumulators (biases already in them) into the next
// layer's B fragments: activation (s, e) is accumulator (s / 2, 8 (s % 2) + e).
__device__ __forceinline__ void mlp_activate(const f32x16* acc, infx8* B) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        infx8 v;
#if SD_MLP_F16
        float a8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) a8[e] = acc[s >> 1][8 * (s & 1) + e];
        v = gelu_half_h8(to_f16x8(a8));
#else
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)gelu_f(acc[s >> 1][8 * (s & 1) + e]);
#endif
        B[s] = v;
    }
}

// log(1 + count), count = rho / 2^24, as the inference features take it.
__device__ __forceinline__ float log1p_count(long long r) {
    // count from the two 32-bit halves (a generic int64 -> float conversion
    // costs several more instructions); within an ulp of (float)rho / 2^24
    const float c = fmaf((float)(int)(r >> 32), 256.0f, (float)(unsigned int)r * (1.0f / (float)SD_FIX_ONE));
    // log2 by the hardware instruction (its argument is >= 1: no denormal path), times ln 2
    return __builtin_amdgcn_logf(1.0f + c) * 0.69314718055994531f;
}

// The density features of every node, once per correction: log1p_count of the
// node's fixed-point density as an inference element (G^3 x 2 bytes). Each
// particle reads 27 of these instead of converting 27 densities itself.
__global__ void __launch_bounds__(256) k_log_density(const long long* __restrict__ rho, long long cells,
                                                     inf_t* __restrict__ out) {
    const long long i0 = ((long long)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i0 + 4 <= cells) {
        const longlong2 r0 = *reinterpret_cast<const longlong2*>(rho + i0);
        const longlong2 r1 = *reinterpret_cast<const longlong2*>(rho + i0 + 2);
        inf_t v[4] = {(inf_t)log1p_count(r0.x), (inf_t)log1p_count(r0.y), (inf_t)log1p_count(r1.x),
                      (inf_t)log1p_count(r1.y)};
        *reinterpret_cast<uint2*>(out + i0) = __builtin_bit_cast(uint2, v);
    } else {
        for (long long i = i0; i < cells; ++i) out[i] = (inf_t)log1p_count(rho[i]);
    }
}

// Input order of the inference kernel. Lane half g holds k = 16 s + 8 g + e
// (s = 0, 1; e = 0..7) of a particle's inputs, and the order puts the same
// (dy, dx) neighbour in both halves at a given (s, e) wherever it can, so a
// lane's node index is one select and one add. With j = 3 dy + dx (dy, dx in
// 0..2 for offsets -1..1) and density feature 9 dz + j:
//   s = 0:                 density (dz = 2 g, j = e)
//   s = 1, e = 0:          density (dz = 2 g, j = 8)
//   s = 1, g = 0, e >= 1:  density (dz = 1, j = e - 1)
//   s = 1, g = 1:          e = 1, 2 density (dz = 1, j = e + 6); e = 3..5 the
//                          offsets (features 27-29); e = 6, 7 padding (30, 31)
// Returns the feature index (particle_features order) of inference input k.
__host__ __device__ constexpr int infer_input_feature(int k) {
    const int s = k >> 4, g = (k >> 3) & 1, e = k & 7;
    if (s == 0) return 18 * g + e;
    if (e == 0) return 8 + 18 * g;
    if (g == 0) return 8 + e;
    return e <= 2 ? 15 + e : 24 + e;
}

constexpr bool infer_order_ok() {
    bool seen[MLP_INP] = {};
    for (int k = 0; k < MLP_INP; ++k) {
        const int q = infer_input_feature(k);
        if (q < 0 || q >= MLP_INP || seen[q]) return false;
        seen[q] = true;
    }
    return infer_input_feature(30) == 30 && infer_input_feature(31) == 31;  // the bias inputs stay in place
}
static_assert(MLP_INP == 32 && MLP_IN == 30 && infer_order_ok(), "inference input order must be a permutation");

__device__ __forceinline__ inf_t ld_lrho(const inf_t* base, unsigned byteOff) {
    return *reinterpret_cast<const inf_t*>(reinterpret_cast<const char*>(base) + byteOff);
}

// A lane's 16 inputs (its half g) in the order above, from the log-density
// grid; the offsets are those of particle_features. Inactive particles (x < 0)
// get zero offsets and the densities around node 0 (finite; their output is
// scaled by zero).
__device__ __forceinline__ void infer_feature_frags(float4 p, const inf_t* __restrict__ lrho, int G, int periodic,
                                                    int g, infx8* B) {
    const float fG = (float)G;
    const bool alive = p.x >= 0.f;
    const int cx = min(max((int)floorf(p.x * fG), 0), G - 1);
    const int cy = min(max((int)floorf(p.y * fG), 0), G - 1);
    const int cz = min(max((int)floorf(p.z * fG), 0), G - 1);
    // byte offsets of the neighbour planes, rows and columns
    unsigned xs[3], ys[3], zs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        xs[d] = (unsigned)wrap_or_clamp1(cx + d - 1, G, periodic) * (unsigned)sizeof(inf_t);
        ys[d] = (unsigned)(wrap_or_clamp1(cy + d - 1, G, periodic) * G) * (unsigned)sizeof(inf_t);
        zs[d] = (unsigned)(wrap_or_clamp1(cz + d - 1, G, periodic) * G * G) * (unsigned)sizeof(inf_t);
    }
    unsigned yx[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) yx[j] = ys[j / 3] + xs[j % 3];
    const unsigned z02 = g ? zs[2] : zs[0];
    infx8 v0, v1;
#pragma unroll
    for (int e = 0; e < 8; ++e) v0[e] = ld_lrho(lrho, z02 + yx[e]);
    v1[0] = ld_lrho(lrho, z02 + yx[8]);
#pragma unroll
    for (int e = 1; e < 3; ++e) v1[e] = ld_lrho(lrho, zs[1] + (g ? yx[e + 6] : yx[e - 1]));
    const float off[3] = {alive ? 2.0f * (p.x * fG - 0.5f - (float)cx) : 0.f,
                          alive ? 2.0f * (p.y * fG - 0.5f - (float)cy) : 0.f,
                          alive ? 2.0f * (p.z * fG - 0.5f - (float)cz) : 0.f};
#pragma unroll
    for (int e = 3; e < 8; ++e) {
        const inf_t d = ld_lrho(lrho, zs[1] + yx[e - 1]);  // half 0's density (half 1 loads it and drops it)
        const inf_t h1 = (inf_t)(e < 6 ? off[e - 3] : (BIAS_MM ? 1.f : 0.f));
        v1[e] = g ? h1 : d;
    }
    B[0] = v0;
    B[1] = v1;
}

// Eight BF16 weights (as stored) -> eight inference elements, times a power
// of two `scale` (FP16: exact, except weights below 2^-16 whose halving drops
// their last bit into the subnormal range).
__device__ __forceinline__ uint4 to_inf(uint4 v, float scale = 1.0f) {
#if SD_MLP_F16
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const f16x2 h = {(_Float16)(__uint_as_float(w[k] << 16) * scale),
                         (_Float16)(__uint_as_float(w[k] & 0xffff0000u) * scale)};
        w[k] = __builtin_bit_cast(uint32_t, h);
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
#else
    (void)scale;
    return v;
#endif
}

// FP32 bias -> its FP16 halves (bh, bl) as one packed word.
__device__ __forceinline__ uint32_t bias_halves(float b) {
    const _Float16 bh = (_Float16)b;
    const f16x2 h = {bh, (_Float16)(b - (float)bh)};
    return __builtin_bit_cast(uint32_t, h);
}

__global__ void __launch_bounds__(64 * INF_WAVES) k_mlp_infer(MlpInferArgs a) {
    extern __shared__ unsigned char smem[];
    inf_t* sW1 = reinterpret_cast<inf_t*>(smem);
    inf_t* sW2 = sW1 + MLP_H * W1_LD;
    inf_t* sW3 = sW2 + MLP_H * WH_LD;
    inf_t* sW4 = sW3 + MLP_H * WH_LD;
    float* sB = reinterpret_cast<float*>(sW4 + MLP_OUTP * WH_LD);  // b1q | b2q | b3q | b4 (padded to 32)
    const unsigned short* W1s = reinterpret_cast<const unsigned short*>(a.W1p);
    for (int i = threadIdx.x; i < MLP_H * (MLP_INP / 8); i += blockDim.x) {
        const int r = i / (MLP_INP / 8), c = i % (MLP_INP / 8);
        // columns in the inference input order (padding columns 30, 31 stay in place)
        uint32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            w[k] = (uint32_t)W1s[r * MLP_INP + infer_input_feature(8 * c + 2 * k)] |
                   (uint32_t)W1s[r * MLP_INP + infer_input_feature(8 * c + 2 * k + 1)] << 16;
        uint4 v = to_inf(make_uint4(w[0], w[1], w[2], w[3]), INF_IN_SCALE);
        if (BIAS_MM && c == MLP_INP / 8 - 1) v.w = bias_halves(a.b[r] * INF_IN_SCALE);  // columns 30, 31 (zero weights)
        *reinterpret_cast<uint4*>(sW1 + r * W1_LD + 8 * c) = v;
    }
    for (int i = threadIdx.x; i < MLP_H * (KH / 8); i += blockDim.x) {
        const int r = i / (KH / 8), c = i % (KH / 8);
        if (c < MLP_H / 8) {
            const int q = r * (MLP_H / 8) + c;
            *reinterpret_cast<uint4*>(sW2 + r * WH_LD + 8 * c) = to_inf(reinterpret_cast<const uint4*>(a.W2q)[q]);
            *reinterpret_cast<uint4*>(sW3 + r * WH_LD + 8 * c) = to_inf(reinterpret_cast<const uint4*>(a.W3q)[q]);
        } else {  // the bias k-step: columns 128, 129 = (bh, bl), the rest zero
            const bool first = c == MLP_H / 8;
            *reinterpret_cast<uint4*>(sW2 + r * WH_LD + 8 * c) =
                make_uint4(first ? bias_halves(a.b[MLP_H + r] * INF_IN_SCALE) : 0u, 0u, 0u, 0u);
            *reinterpret_cast<uint4*>(sW3 + r * WH_LD + 8 * c) =
                make_uint4(first ? bias_halves(a.b[2 * MLP_H + r] * INF_IN_SCALE) : 0u, 0u, 0u, 0u);
        }
    }
    for (int i = threadIdx.x; i < MLP_OUTP * (MLP_H / 8); i += blockDim.x) {
        const int r = i / (MLP_H / 8), c = i % (MLP_H / 8);
        *reinterpret_cast<uint4*>(sW4 + r * WH_LD + 8 * c) =
            to_inf(reinterpret_cast<const uint4*>(a.W4q)[i], INF_OUT_SCALE);
    }
    for (int i = threadIdx.x; i < 3 * MLP_H + MLP_OUTP; i += blockDim.x) {
        if (i < 3 * MLP_H) {
            // accumulator order of layer L: [t][g][j] <- bq[L H + 32 t + 16 (j / 8) + 8 g + (j % 8)]
            const int L = i / MLP_H, r = i % MLP_H, t = r / 32, g = (r / 16) & 1, j = r & 15;
            sB[i] = a.bq[L * MLP_H + 32 * t + 16 * (j >> 3) + 8 * g + (j & 7)] * INF_IN_SCALE;
        } else {
            sB[i] = i - 3 * MLP_H < MLP_OUT ? a.b[i] : 0.f;
        }
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 5;
    const int ntiles = (a.n + 31) / 32;
    for (int tile = blockIdx.x * INF_WAVES + w; tile < ntiles; tile += gridDim.x * INF_WAVES) {
        // keep the weight-fragment LDS reads inside the loop: hoisted out of it
        // they would need 320 registers per lane
        __asm__ volatile("" ::: "memory");
        const int i = tile * 32 + (l & 31);
        float4 p = make_float4(-1.f, -1.f, -1.f, 0.f);
        if (i < a.n) p = a.pos[a.order ? a.order[i] : i];
        infx8 B[KH / 16];
        if constexpr (BIAS_MM) {
            // inputs of the bias k-step: k = 128, 129 are 1 (lanes 0-31 hold k = 128..135)
            infx8 one{};
            if (g == 0) one[0] = one[1] = (inf_t)1.0f;
            B[KH / 16 - 1] = one;
        }
        f32x16 acc[4];
        infer_feature_frags(p, reinterpret_cast<const inf_t*>(a.lrho), a.G, a.periodic, g, B);
        mlp_layer<MLP_INP, W1_LD, 4, !BIAS_MM>(sW1, B, acc, l, sB);
        mlp_activate(acc, B);
        mlp_layer<KH, WH_LD, 4, !BIAS_MM>(sW2, B, acc, l, sB + MLP_H);
        mlp_activate(acc, B);
        mlp_layer<KH, WH_LD, 4, !BIAS_MM>(sW3, B, acc, l, sB + 2 * MLP_H);
        mlp_activate(acc, B);
        mlp_layer<MLP_H, WH_LD, 1, false>(sW4, B, acc, l, nullptr);
        // outputs 0..2 are accumulator rows 0..2 (elements 0..2 on lanes 0..31)
        if (g == 0 && i < a.n) {
            const float* b4 = sB + 3 * MLP_H;
            const float sc = p.x >= 0.f ? a.scale : 0.f;
            a.out[i] = make_float4((acc[0][0] + b4[0]) * sc, (acc[0][1] + b4[1]) * sc, (acc[0][2] + b4[2]) * sc, 0.f);
        }
    }
}

size_t mlp_infer_smem() {
    return sizeof(inf_t) * (MLP_H * W1_LD + 2 * MLP_H * WH_LD + MLP_OUTP * WH_LD) +
           sizeof(float) * (3 * MLP_H + MLP_OUTP);
}

void launch_log_density(const long long* rho, long long cells, uint16_t* out, hipStream_t st) {
    const long long blocks = (cells + 1023) / 1024;
    hipLaunchKernelGGL(k_log_density, dim3((unsigned)blocks), dim3(256), 0, st, rho, cells,
                       reinterpret_cast<inf_t*>(out));
}

void launch_mlp_infer(const MlpInferArgs& a, int blocks, hipStream_t st) {
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_mlp_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)mlp_infer_smem());
        attr = true;
    }
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * INF_WAVES), mlp_infer_smem(), st, a);
}

__global__ void k_gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx, int n, int width,
                              float* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * width) return;
    const int r = i / width, c = i % width;
    dst[i] = src[(size_t)idx[r] * width + c];
}

void launch_gather_rows(const float* src, const int32_t* idx, int n, int width, float* dst, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gather_rows, dim3((n * width + 255) / 256), dim3(256), 0, st, src, idx, n, width, dst);
}

void launch_features(const float4* pos, const int32_t* targets, int nt, const long long* rho, int G, int periodic,
                     float* X, hipStream_t st) {
    if (nt > 0)
        hipLaunchKernelGGL(k_features, dim3((nt + 255) / 256), dim3(256), 0, st, pos, targets, nt, rho, G, periodic, X);
}

// ---- training step ----------------------------------------------------------------
// One batch: forward (storing U and H, H^T), loss gradient, backward with the
// GELU derivative fused into the data-gradient GEMMs, weight gradients as
// split-K GEMMs over the batch, bias gradients as column sums. Gradients land
// in gW / gb in the oracle's layout (FP32).
int mlp_pad_rows(int n) { return (n + 255) / 256 * 256; }

void mlp_forward(MlpDev& d, const float* X, int n, hipStream_t st) {
    const int nPad = mlp_pad_rows(n);
    hipLaunchKernelGGL(k_mlp_input, dim3((nPad * MLP_INP + 255) / 256), dim3(256), 0, st, X, n, nPad, d.Xb, d.Xt);
    // forward
    GemmArgs g{};
    g.M = nPad;
    // layer 1
    g.A = d.Xb; g.lda = MLP_INP; g.Bt = d.W1p; g.ldb = MLP_INP; g.N = MLP_H; g.K = MLP_INP; g.ksplit = MLP_INP;
    g.bias = d.b; g.U = d.U[0]; g.H = d.Hb[0]; g.Ht = d.Ht[0]; g.ldh = MLP_H; g.ldht = nPad;
    gemm(g, EPI_BIAS_GELU, 1, st);
    // layers 2, 3
    for (int l = 1; l < 3; ++l) {
        g.A = d.Hb[l - 1]; g.lda = MLP_H; g.Bt = l == 1 ? d.W2 : d.W3; g.ldb = MLP_H; g.K = MLP_H; g.ksplit = MLP_H;
        g.bias = d.b + l * MLP_H; g.U = d.U[l]; g.H = d.Hb[l]; g.Ht = d.Ht[l];
        gemm(g, EPI_BIAS_GELU, 1, st);
    }
    // layer 4 -> Y [nPad x 32]
    {
        GemmArgs o{};
        o.M = nPad; o.A = d.Hb[2]; o.lda = MLP_H; o.Bt = d.W4p; o.ldb = MLP_H; o.N = MLP_OUTP; o.K = MLP_H;
        o.ksplit = MLP_H; o.bias = d.b4p; o.C = d.Y;
        gemm(o, EPI_BIAS, 1, st);
    }
}

void mlp_batch_grad(MlpDev& d, const float* X, const float* T, int n, hipStream_t st) {
    const int nPad = mlp_pad_rows(n);
    const float invn = 1.0f / (float)(n > 0 ? n : 1);
    mlp_forward(d, X, n, st);
    const int lb = (nPad * MLP_OUTP + 255) / 256;
    hipLaunchKernelGGL(k_mlp_dloss, dim3(lb), dim3(256), 0, st, d.Y, T, n, nPad, invn, d.dYb, d.dYt, d.dYf,
                       d.lossPart);
    d.lossBlocks = lb;
    // backward. delta of layer 4 = dY; dH3 = dY * W4p, then dU3 = dH3 * gelu'(U3)
    const int splits = nPad >= 4096 ? 16 : 1;
    const int ks = nPad / splits;
    // weight gradient of layer 4: dW4 = dY^T [32 x nPad] * H3 [nPad x 128]
    {
        GemmArgs w{};
        w.A = d.dYt; w.lda = nPad; w.Bt = d.Ht[2]; w.ldb = nPad; w.M = MLP_OUTP; w.N = MLP_H; w.K = nPad;
        w.ksplit = ks; w.C = d.part;
        gemm(w, EPI_STORE, splits, st);
        hipLaunchKernelGGL(k_reduce_splits, dim3((MLP_OUT * MLP_H + 255) / 256), dim3(256), 0, st, d.part, splits,
                           MLP_OUTP, MLP_H, MLP_OUT, MLP_H, d.gW + MLP_H * MLP_IN + 2 * MLP_H * MLP_H);
        hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, st, d.dYf, nPad, MLP_OUTP, d.gb4p);
    }
    // dU3 = (dY * W4p) .* gelu'(U3)  (A = dY [nPad x 32], Bt = W4pt [128 x 32])
    {
        GemmArgs b{};
        b.M = nPad; b.A = d.dYb; b.lda = MLP_OUTP; b.Bt = d.W4pt; b.ldb = MLP_OUTP; b.N = MLP_H; b.K = MLP_OUTP;
        b.ksplit = MLP_OUTP; b.Uin = d.U[2]; b.Df = d.Df[2]; b.H = d.Db[2]; b.Ht = d.Dt[2]; b.ldh = MLP_H;
        b.ldht = nPad;
        gemm(b, EPI_DGELU, 1, st);
    }
    for (int l = 2; l >= 0; --l) {
        // weight gradient of layer l: dW = dU_l^T [128 x nPad] * H_{l-1} [nPad x in]
        const bool first = l == 0;
        const int in = first ? MLP_INP : MLP_H;
        GemmArgs w{};
        w.A = d.Dt[l]; w.lda = nPad; w.Bt = first ? d.Xt : d.Ht[l - 1]; w.ldb = nPad; w.M = MLP_H; w.N = in;
        w.K = nPad; w.ksplit = ks; w.C = d.part;
        gemm(w, EPI_STORE, splits, st);
        float* gWl = d.gW + (l == 0 ? 0 : MLP_H * MLP_IN + (l - 1) * MLP_H * MLP_H);
        const int inLive = first ? MLP_IN : MLP_H;
        hipLaunchKernelGGL(k_reduce_splits, dim3((MLP_H * inLive + 255) / 256), dim3(256), 0, st, d.part, splits,
                           MLP_H, in, MLP_H, inLive, gWl);
        hipLaunchKernelGGL(k_colsum, dim3((MLP_H + 63) / 64), dim3(64), 0, st, d.Df[l], nPad, MLP_H,
                           d.gb + l * MLP_H);
        if (l == 0) break;
        // dU_{l-1} = (dU_l * W_l) .* gelu'(U_{l-1})  (Bt = W_l^T [in x out])
        GemmArgs b{};
        b.M = nPad; b.A = d.Db[l]; b.lda = MLP_H; b.Bt = l == 1 ? d.W2t : d.W3t; b.ldb = MLP_H; b.N = MLP_H;
        b.K = MLP_H; b.ksplit = MLP_H; b.Uin = d.U[l - 1]; b.Df = d.Df[l - 1]; b.H = d.Db[l - 1];
        b.Ht = d.Dt[l - 1]; b.ldh = MLP_H; b.ldht = nPad;
        gemm(b, EPI_DGELU, 1, st);
    }
}

void mlp_pack(MlpDev& d, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, st, d.W, d);
    // padded output bias b4p
    (void)hipMemcpyAsync(d.b4p, d.b + 3 * MLP_H, sizeof(float) * MLP_OUT, hipMemcpyDeviceToDevice, st);
}

void mlp_adam(MlpDev& d, float lr, int t, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    // the output-layer bias gradient lives in gb4p (padded); copy its 3 values
    (void)hipMemcpyAsync(d.gb + 3 * MLP_H, d.gb4p, sizeof(float) * MLP_OUT, hipMemcpyDeviceToDevice, st);
    hipLaunchKernelGGL(k_adam, dim3((MLP_NW + 255) / 256), dim3(256), 0, st, d.W, d.gW, d.mW, d.vW, MLP_NW, lr, b1,
                       b2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(MLP_NB), 0, st, d.b, d.gb, d.mb, d.vb, MLP_NB, lr, b1, b2, eps, c1, c2);
}

}  // namespace sd

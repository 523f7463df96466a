// sd_mlp.h -- device buffers and launchers of the learned force correction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sd {

constexpr int MLP_IN = 30, MLP_INP = 32, MLP_H = 128, MLP_OUT = 3, MLP_OUTP = 32;
constexpr int MLP_NW = MLP_H * MLP_IN + 2 * MLP_H * MLP_H + MLP_OUT * MLP_H;  // 36,992
constexpr int MLP_NB = 3 * MLP_H + MLP_OUT;                                     // 387

// Parameters (FP32 master, oracle layout: W1 [H x 30] | W2 | W3 | W4 [3 x H],
// b1 | b2 | b3 | b4), Adam moments, BF16 copies and per-batch work arrays.
struct MlpDev {
    float *W, *b, *gW, *gb, *mW, *vW, *mb, *vb;
    float *b4p, *gb4p;  // output bias / gradient padded to 32
    __bf16 *W1p, *W2, *W3, *W4p, *W2t, *W3t, *W4pt;
    __bf16 *W2q, *W3q, *W4q;  // inference copies, input order permuted (perm_k)
    float* bq;                // biases of layers 1-3 in that order
    int maxBatch;       // padded rows the work arrays hold
    __bf16 *Xb, *Xt;
    float* U[3];
    __bf16 *Hb[3], *Ht[3];
    float* Df[3];
    __bf16 *Db[3], *Dt[3];
    float* Y;
    __bf16 *dYb, *dYt;
    float* dYf;
    double* lossPart;
    int lossBlocks;
    float* part;        // split-K partials
};

struct MlpInferArgs {
    const float4* pos;
    const int32_t* order;  // nullptr, or the sort's permutation: particle i is pos[order[i]]
    int n;
    const uint16_t* lrho;  // log-density grid of the current positions (launch_log_density)
    int G;
    int periodic;
    const __bf16 *W1p, *W2q, *W3q, *W4q;
    const float* b;        // FP32 biases (oracle layout; the output bias is read from here)
    const float* bq;       // biases of layers 1-3 in permuted order
    float scale;           // output scale: outScale * a_unit
    float4* out;           // correction (box units), w = 0
};

int mlp_pad_rows(int n);
void mlp_pack(MlpDev& d, hipStream_t st);
void mlp_forward(MlpDev& d, const float* X, int n, hipStream_t st);  // -> d.Y
void mlp_batch_grad(MlpDev& d, const float* X, const float* T, int n, hipStream_t st);
void mlp_adam(MlpDev& d, float lr, int t, hipStream_t st);
size_t mlp_infer_smem();
// The inference kernel's density features: one 2-byte inference element per
// node (G^3 of them) from the fixed-point density.
void launch_log_density(const long long* rho, long long cells, uint16_t* out, hipStream_t st);
void launch_mlp_infer(const MlpInferArgs& a, int blocks, hipStream_t st);
void launch_features(const float4* pos, const int32_t* targets, int nt, const long long* rho, int G, int periodic,
                     float* X, hipStream_t st);
void launch_gather_rows(const float* src, const int32_t* idx, int n, int width, float* dst, hipStream_t st);

}  // namespace sd

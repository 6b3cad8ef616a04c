// mlp.h — the learned relocation proposal: per-Gaussian statistics and features, and a
// 16 -> 64 -> 64 -> 4 MLP (GELU) evaluated and trained with bf16 MFMA (fp32 accumulation).
// No counterpart in the reference, whose MCMC relocation samples sources by opacity only
// (Model::relocate, core/src/model.cpp:336-416); the plain sampler stays the fallback.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace msk {

constexpr int MLP_IN = 16, MLP_HID = 64, MLP_OUT = 4;
// fp32 master parameters, in this order: W1 [64 16], b1 [64], W2 [64 64], b2 [64],
// W3 [4 64], b3 [4] (row-major [out][in]).
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = 1024, MLP_OFF_W2 = 1088, MLP_OFF_B2 = 5184, MLP_OFF_W3 = 5248,
              MLP_OFF_B3 = 5504, MLP_NPARAM = 5508;

// The operand block the kernels read: bf16 weights, fp32 biases.
struct alignas(16) MlpPacked {
    uint16_t W1[MLP_HID * MLP_IN];
    uint16_t W2[MLP_HID * MLP_HID];
    uint16_t W3[MLP_OUT * MLP_HID];
    float b1[MLP_HID], b2[MLP_HID], b3[MLP_OUT];
};

// Per-Gaussian statistics gathered between relocation events (struct of arrays, [N] each).
struct ReloStats {
    float* vis;     // views in which the Gaussian was visible (radius > 0)
    float* gsum;    // sum of |d loss / d mean2d|
    float* maxr;    // largest screen radius (pixels)
    float* dsum;    // sum of view depths
    float* dsq;     // sum of squared view depths
};

// One view's contribution to the statistics (after its raster backward).
void relo_stats_accum(hipStream_t s, int N, const int32_t* radii, const float* depths, const float* g2d,
                      ReloStats st);
// Features [N 16] (bf16) of the current state and statistics over `views` views.
void relo_features(hipStream_t s, const Params& P, const ReloStats& st, int views, uint16_t* X);

// theta (fp32 masters) -> the packed operand block.
void mlp_pack(hipStream_t s, const float* theta, MlpPacked* w);
// out [N 4] = MLP(X [N 16]).
void mlp_forward(hipStream_t s, int N, const uint16_t* X, const MlpPacked* w, float* out);
// Training targets, [N 8] per sample: y (0 / 1), weight of its logit term, offset target
// (3), weight of its offset term, unused (2). Loss = sum_i wl_i BCE(z0_i, y_i)
// + wo_i |z_i[1:4] - t_i|^2. Adds d loss / d theta into grad [MLP_NPARAM] and the two loss
// terms into loss2 [2] (both must be cleared by the caller).
void mlp_backward(hipStream_t s, int N, const uint16_t* X, const float* targets, const MlpPacked* w, float* grad,
                  double* loss2);
// Adam on theta (fp32), then re-packs w.
void mlp_adam(hipStream_t s, float* theta, float* m, float* v, const float* grad, float lr, float beta1, float beta2,
              float eps, float bc1, float bc2_sqrt, MlpPacked* w);

// Training records of one relocation event: for each relocated row j, the features of its
// source (before the event), its destination row, and the destination's state right after
// the event (opacity logit, source mean, largest scale).
struct ReloRecords {
    uint16_t* X;      // [n 16]
    uint32_t* dst;    // [n]
    float* o;         // [n]
    float* mu;        // [n 3]
    float* s;         // [n]
};
void relo_record(hipStream_t s, int n, const uint32_t* pair_src, const uint32_t* pair_dst, const uint16_t* X,
                 const Params& P, ReloRecords R);
// Outcome of recorded relocations, n samples written to X_out [n 16] / T_out [n 8]:
// y = 1 if the copy's opacity logit grew since the event; offset target = (mean now -
// source mean) / scale, clamped to [-3, 3]; weights wl, wo.
void relo_label(hipStream_t s, int n, const ReloRecords R, const Params& P, float wl, float wo, uint16_t* X_out,
                float* T_out);
// Sets the loss weights of samples [0, n).
void relo_weights(hipStream_t s, int n, float* T, float wl, float wo);

}  // namespace msk

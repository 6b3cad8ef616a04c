// eval.h — evaluation sums and display conversion kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace msk {

// Device scratch needed by eval_sums for a W x H image.
size_t eval_partials_bytes(int W, int H);
// out3 = {sum (r - g)^2, sum |r - g|, sum SSIM} over all H*W*3 values (FP64), where g is
// the uint8 ground truth / 255 and SSIM uses the replicate-padded 11-tap window.
void eval_sums(hipStream_t s, const float* img, const uint8_t* gt, int W, int H, const float window[11],
               double* partials, double* out3);
// RGB f32 [n, 3] -> RGBA8 [n, 4], clamped and truncated, alpha 255.
void to_rgba8(hipStream_t s, const float* src, int n, uint8_t* dst);

}  // namespace msk

// eval.hip — held-out evaluation metrics and display conversion on the device.
//
// The reference evaluates on the CPU after copying each render back (psnr / l1_loss at
// core/src/model.cpp:34-49, ssim_eval at core/src/ssim.cpp:61-123). Here one kernel per
// 16x16 tile stages a replicate-padded 26x26 halo of render and ground truth in LDS,
// runs the separable 11-tap blur of (x, y, x^2, y^2, xy) per channel, and reduces
// squared error, absolute error and the SSIM map in FP64 per workgroup; a second kernel
// sums the partials. Per-pixel arithmetic is the reference's float expression order with
// the blur accumulated by fused multiply-add (what clang's default contraction makes of
// `sum += in * k` on the reference's arm64 targets), so only the FP64 summation order
// differs from a sequential host loop.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eval.h"

namespace msk {

constexpr int ET = 16, EH = ET + 10;

struct Win11 {
    float w[11];
};

__device__ __forceinline__ float blur_row(const float* row, const Win11& k) {
    float s = 0.0f;
#pragma unroll
    for (int t = 0; t < 11; t++) s = fmaf(row[t], k.w[t], s);
    return s;
}

__global__ __launch_bounds__(256) void k_eval_tiles(const float* __restrict__ img, const uint8_t* __restrict__ gt, int W,
                                                    int H, Win11 win, double* __restrict__ partials) {
    __shared__ float sx[EH][EH], sy[EH][EH];
    __shared__ float hb[5][EH][ET];
    __shared__ double red[3][256];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x0 = blockIdx.x * ET, y0 = blockIdx.y * ET;
    const int x = x0 + tx, y = y0 + ty;
    const bool inside = x < W && y < H;
    double acc_sq = 0.0, acc_abs = 0.0, acc_ssim = 0.0;
    const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
    for (int c = 0; c < 3; c++) {
        for (int e = threadIdx.x; e < EH * EH; e += 256) {
            const int j = e / EH, i = e - j * EH;
            const int gy = min(max(y0 - 5 + j, 0), H - 1), gx = min(max(x0 - 5 + i, 0), W - 1);
            const size_t p = ((size_t)gy * W + gx) * 3 + c;
            sx[j][i] = img[p];
            sy[j][i] = (float)gt[p] * (1.0f / 255.0f);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < EH * ET; e += 256) {
            const int j = e / ET, i = e - j * ET;
            float a[11], b[11], aa[11], bb[11], ab[11];
#pragma unroll
            for (int t = 0; t < 11; t++) {
                a[t] = sx[j][i + t];
                b[t] = sy[j][i + t];
                aa[t] = a[t] * a[t];
                bb[t] = b[t] * b[t];
                ab[t] = a[t] * b[t];
            }
            hb[0][j][i] = blur_row(a, win);
            hb[1][j][i] = blur_row(b, win);
            hb[2][j][i] = blur_row(aa, win);
            hb[3][j][i] = blur_row(bb, win);
            hb[4][j][i] = blur_row(ab, win);
        }
        __syncthreads();
        if (inside) {
            float v[5];
#pragma unroll
            for (int q = 0; q < 5; q++) {
                float s = 0.0f;
#pragma unroll
                for (int t = 0; t < 11; t++) s = fmaf(hb[q][ty + t][tx], win.w[t], s);
                v[q] = s;
            }
            const float m1sq = v[0] * v[0], m2sq = v[1] * v[1], m12 = v[0] * v[1];
            const float sig1sq = v[2] - m1sq, sig2sq = v[3] - m2sq, sig12 = v[4] - m12;
            const float num = (2.0f * m12 + C1) * (2.0f * sig12 + C2);
            const float den = (m1sq + m2sq + C1) * (sig1sq + sig2sq + C2);
            acc_ssim += (double)(num / den);
            const float r = sx[ty + 5][tx + 5], g = sy[ty + 5][tx + 5];
            const double d = (double)(r - g);
            acc_sq += d * d;
            acc_abs += (double)fabsf(r - g);
        }
        __syncthreads();
    }
    red[0][threadIdx.x] = acc_sq;
    red[1][threadIdx.x] = acc_abs;
    red[2][threadIdx.x] = acc_ssim;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) partials[(size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_eval_sum(const double* __restrict__ partials, int nblk, double* __restrict__ out) {
    __shared__ double red[3][256];
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += 256)
        for (int q = 0; q < 3; q++) a[q] += partials[(size_t)b * 3 + q];
    for (int q = 0; q < 3; q++) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = red[threadIdx.x][0];
}

size_t eval_partials_bytes(int W, int H) {
    return (size_t)((W + ET - 1) / ET) * ((H + ET - 1) / ET) * 3 * sizeof(double);
}

void eval_sums(hipStream_t s, const float* img, const uint8_t* gt, int W, int H, const float window[11],
               double* partials, double* out3) {
    Win11 w;
    for (int t = 0; t < 11; t++) w.w[t] = window[t];
    const dim3 grid((W + ET - 1) / ET, (H + ET - 1) / ET);
    hipLaunchKernelGGL(k_eval_tiles, grid, dim3(256), 0, s, img, gt, W, H, w, partials);
    hipLaunchKernelGGL(k_eval_sum, dim3(1), dim3(256), 0, s, partials, (int)(grid.x * grid.y), out3);
}

// RGB float -> RGBA8 as the reference's display path truncates it (msplat_api.mm:228-233).
__global__ void k_rgba8(const float* __restrict__ src, int n, uint8_t* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uchar4 o;
    o.x = (uint8_t)(fminf(fmaxf(src[(size_t)i * 3], 0.f), 1.f) * 255.f);
    o.y = (uint8_t)(fminf(fmaxf(src[(size_t)i * 3 + 1], 0.f), 1.f) * 255.f);
    o.z = (uint8_t)(fminf(fmaxf(src[(size_t)i * 3 + 2], 0.f), 1.f) * 255.f);
    o.w = 255;
    reinterpret_cast<uchar4*>(dst)[i] = o;
}

void to_rgba8(hipStream_t s, const float* src, int n, uint8_t* dst) {
    if (n > 0) hipLaunchKernelGGL(k_rgba8, dim3((n + 255) / 256), dim3(256), 0, s, src, n, dst);
}

}  // namespace msk

// kernels.h — host-side launchers of the gfx950 kernels (internal interface of the engine).
// All pointers are device pointers; every launcher enqueues on the given stream and
// returns without synchronising.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/msplat_types.h"

namespace msk {

constexpr int TILE = 16;            // binning tile (reference msplat.hpp:4-5)
constexpr int CHUNK = 512;          // depth chunk of the K > 1 raster path
constexpr int SORT_SMALL_CAP = 2048;

// Per-Gaussian record written by the projection and copied into the per-intersection
// packed lists by the sort: A = (x, y, sigmoid(opacity), conic.a),
// B = (conic.b, conic.c, r, g), C = (b, gaussian id bits). Colours are raw SH output.
struct Records {
    float4* A;
    float4* B;
    float2* C;
};

struct ProjOut {
    int32_t* radii;
    int32_t* tiles_hit;
    float* depths;
    ushort4* rect;      // tile rectangle [x0, y0, x1, y1)
    float2* aabb;
    Records rec;
};

struct ListBufs {
    uint32_t* tile_counts;     // [T]
    uint32_t* write_ctr;       // [T]
    uint32_t* exact_off;       // [T]
    uint32_t* sort_off;        // [T+1]
    int2* bins;                // [T]
    uint64_t* keys;            // [key_capacity]
    Records pk;                // [capacity]
    uint32_t* totals;          // [8]: exact total, padded total, overflow flag, max count, #large tiles
    int64_t capacity, key_capacity;
    uint32_t* big_list;        // [T] tiles with more than SORT_SMALL_CAP entries
};

struct RasterBufs {
    float* img;       // [H W 3]
    float* final_T;   // [H W]
    int32_t* final_idx;
    // chunked path
    float* chunk_T; float* chunk_C; int32_t* chunk_idx; float* prefix_T; float* after_C;
    uint32_t K;
};

struct GradBufs {
    float2* v_xy;
    float* v_conic;   // [N 3]
    float* v_rgb;     // [N 3]
    float* v_opac;    // [N]
};

struct Params {
    float *means, *scales, *quats, *f_dc, *f_rest, *opac;
    float* m[6];
    float* v[6];
    int N, degree;
};

void project_forward(hipStream_t s, const Params& P, const MsplatCamera& cam, int degrees_to_use, ProjOut& o,
                     uint32_t* nan_counts);
void build_lists(hipStream_t s, int N, const ProjOut& pj, int tbx, int tby, ListBufs& L, uint32_t* nan_counts);
void raster_forward(hipStream_t s, const ListBufs& L, int W, int H, int tbx, int tby, const float bg[3], RasterBufs& R);
void ssim_loss(hipStream_t s, int W, int H, const float* img, const uint8_t* gt, float ssim_w, float inv_n,
               float* hbuf /*[H W 15]*/, float* dbuf /*[H W 9]*/, float* v_rend, float* loss_partials,
               double* loss_out, uint32_t* nan_counts);
void raster_backward(hipStream_t s, const ListBufs& L, const RasterBufs& R, int W, int H, int tbx, int tby,
                     const float bg[3], const float* v_rend, GradBufs& G);
void sh_rest_adam(hipStream_t s, const Params& P, const MsplatCamera& cam, const int32_t* radii, const float* v_rgb,
                  const MsplatStepParams& hp);
void gaussian_update(hipStream_t s, const Params& P, const MsplatCamera& cam, const ProjOut& pj, const GradBufs& G,
                     const MsplatStepParams& hp, uint32_t* nan_counts, float* dbg_v_mean, float* dbg_v_scale,
                     float* dbg_v_quat, float* dbg_v_opac);
void fill_zero(hipStream_t s, void* p, size_t bytes);

}  // namespace msk

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

// Tile-list entries as rasterised: Gaussian ids into per-Gaussian records (row N is an
// all-zero, fully transparent record).
struct EntryView {
    const uint32_t* ids;
    Records rec;
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
    uint32_t* exact_off;       // [T]
    int2* bins;                // [T] (start, end) of each tile's list
    uint32_t* keys;            // [capacity] depth ranks, bucketed per tile
    uint32_t* ids;             // [capacity] sorted per-tile Gaussian ids (the packed lists)
    EntryView view;            // ids + the projection's records
    uint32_t* hist;            // [chunks x T] per-chunk tile histograms -> bucket offsets
    uint32_t* totals;          // [8]: total, padded total (reference layout), overflow flag, max count
    int64_t capacity;
    int num_tiles;
    // global depth order
    uint32_t *dkeys, *dkeys_out, *dvals, *sorted_ids;  // [N]
    void* temp;                // radix-sort scratch
    size_t temp_bytes;
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
    float* m[6];   // groups 3 and 4 hold bf16 data when sh_bf16 is set
    float* v[6];
    int N, degree;
    int sh_bf16 = 0;
    uint16_t* f_dc_h = nullptr;    // bf16 shadows of f_dc / f_rest (sh_bf16 only)
    uint16_t* f_rest_h = nullptr;
};

void project_forward(hipStream_t s, const Params& P, const MsplatCamera& cam, int degrees_to_use, ProjOut& o,
                     uint32_t* nan_counts);
size_t lists_temp_bytes(int N);
size_t lists_hist_bytes(int N, int T);
bool lists_supported(int N, int T);
void lists_rank(hipStream_t s, int N, const ProjOut& pj, ListBufs& L);
void lists_count(hipStream_t s, int N, const ProjOut& pj, int tbx, int tby, ListBufs& L);
void lists_scatter(hipStream_t s, int N, const ProjOut& pj, int tbx, ListBufs& L);
void lists_sort_pack(hipStream_t s, int N, const ProjOut& pj, int tbx, int tby, ListBufs& L);
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
void copy16(hipStream_t s, const void* src, void* dst, size_t bytes);
// bf16 storage helpers: shadows of f_dc / f_rest from the fp32 masters (rows [0, N)), and
// element-wise conversions.
void sh_refresh_shadows(hipStream_t s, const Params& P);
void f32_to_bf16(hipStream_t s, const float* src, uint16_t* dst, size_t n);
void bf16_to_f32(hipStream_t s, const uint16_t* src, float* dst, size_t n);

}  // namespace msk

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
    uint32_t* ids;             // [capacity] sorted per-tile Gaussian ids (the packed lists)
    EntryView view;            // ids + the projection's records
    uint32_t* hist;            // [chunks x T] per-chunk tile histograms, then [chunks x blocks] block offsets
    uint32_t* l1_ids;          // [capacity] block lists: Gaussian ids
    uint16_t* l1_mask;         // [capacity] block lists: masks of the block's tiles the Gaussian covers
    uint32_t* blocks;          // [2 blocks + 2] block list lengths, then their starts and the total
    uint32_t* totals;          // [8]: total, padded total (reference layout), overflow flag, max count
    int64_t capacity;
    int num_tiles;
    // global depth order
    uint32_t *dkeys, *dkeys_out, *dvals, *sorted_ids;  // [N]
    ushort4* srect;            // [N] tile rectangles in depth-rank order
    void* temp;                // radix-sort scratch
    size_t temp_bytes;
    // launch choices (the lists are the same for every choice)
    int chunk_target = 0;      // rank chunks of the count / bucket kernels (0: the default 768)
    int bucket_waves = 0;      // waves per bucket workgroup (0: the process-wide choice)
};

// Constants of one view of a step, read by the kernels from device memory, so that a
// recorded launch sequence (a captured graph) replays for any camera and target whose
// extents fit its launch geometry.
struct ViewConsts {
    MsplatCamera cam;
    int W, H, tbx, tby, T;
    float bg[3];
    float ssim_w, inv_n;
    const uint8_t* gt;   // [H W 3] target, uint8
    const uint8_t* mask; // [H W] training mask (< 128: the pixel owns no loss) or null
};

// Launch geometry: a view's own extents (eager launches) or the largest extents of the
// views a recorded sequence serves (workgroups past a view's extents exit at once).
struct Geom {
    int W, H, tbx, tby, T;
};
Geom geom_of(int W, int H);

struct RasterBufs {
    float* img;       // [H W 3]
    float* final_T;   // [H W]
    int32_t* final_idx;
    // chunked path
    float* chunk_T; float* chunk_C; int32_t* chunk_idx; float* prefix_T; float* after_C;
    uint32_t K;
};

// Per-Gaussian screen-space gradients of the raster backward, interleaved: one record of G2D
// floats per Gaussian (colour 3, conic 3, mean2d 2, opacity logit 1), so that the backward's
// float atomics for one (tile, Gaussian) pair land in one or two cache lines, one lane per
// field (scattered single-field atomics cost several times more).
constexpr int G2D = 9, G2D_RGB = 0, G2D_CONIC = 3, G2D_XY = 6, G2D_OPAC = 8;
struct GradBufs {
    float* g;  // [N G2D]
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

// One view of the persistent raster kernels: its lists, records, outputs and gradients.
struct RasterView {
    const ViewConsts* vc;
    const int2* bins;
    EntryView pk;
    float* img;
    float* final_T;
    int32_t* final_idx;
    const float* v_rend;
    GradBufs G;
    int chunked;   // rasterised by the chunked kernels, not through the work queue
    int32_t* n_contrib = nullptr;  // [H W] entries each pixel took (persistent forward; optional)
};

// Work queue of the persistent raster kernels over B views of at most Tmax tiles each.
struct RasterWork {
    const RasterView* views;   // device [B]
    int B, Tmax;
    uint32_t* queue;           // [B Tmax] work items, longest tile lists first
    uint32_t* n_work;          // [1]
    uint32_t* ctr;             // [2] forward / backward queue heads
    int grid;                  // resident workgroups
    int ppl = 1;               // pixels per lane: 1 (256-lane workgroup per tile) or 4 (one wave)
};

// zero (training steps): the raster backward's gradient accumulators, zeroed row by row by
// the projection (one pass instead of a separate fill)
void project_forward(hipStream_t s, const Params& P, const ViewConsts* vc, int degrees_to_use, ProjOut& o,
                     uint32_t* nan_counts, const GradBufs* zero = nullptr);
size_t lists_temp_bytes(int N);
size_t lists_hist_bytes(int N, int T, int chunk_target = 0);
size_t lists_block_bytes(const Geom& g);
bool lists_supported(int N, int T);
void lists_rank(hipStream_t s, int N, const ProjOut& pj, ListBufs& L);
// wq (a single unchunked view, B = 1): the raster work queue of view 0 is written with the
// tile offsets, and raster_work_build is not needed for it
void lists_count(hipStream_t s, int N, const ProjOut& pj, const ViewConsts* vc, const Geom& g, ListBufs& L,
                 const RasterWork* wq = nullptr);
// Writes every tile hit's Gaussian id into its tile's bucket, in depth-rank order.
void lists_bucket(hipStream_t s, int N, const ViewConsts* vc, const Geom& g, ListBufs& L);
// Waves per bucket workgroup: 0 = 2 where the tile count's LDS allows, else 1; 1, 2 or 4 on
// request where the LDS allows it (process-wide; tests use it to cover every variant).
void set_bucket_waves(int waves);
// Waves per workgroup of the bucket kernel for T tiles under an LDS limit of `limit` bytes
// (request: 0 by LDS, 1, 2 or 4); 0 when the tile count does not fit.
int bucket_waves(int T, size_t limit, int request);
// Fills the persistent raster work queue (ordered by decreasing tile list length) and
// resets its heads; the views' lists must be built.
void raster_work_build(hipStream_t s, const RasterWork& w);
// Monolithic views (K <= 1) rasterise through the persistent kernel over `work`; chunked
// views through the per-(tile, chunk) kernels.
void raster_forward(hipStream_t s, const ListBufs& L, const ViewConsts* vc, const Geom& g, RasterBufs& R,
                    const RasterWork* work);
void raster_forward_views(hipStream_t s, const RasterWork& w);
void ssim_loss(hipStream_t s, const ViewConsts* vc, const Geom& g, const float* img, float* dbuf /*[H W 9]*/,
               float* v_rend, float* loss_partials, double* loss_out, uint32_t* nan_counts);
void raster_backward(hipStream_t s, const ListBufs& L, const RasterBufs& R, const ViewConsts* vc, const Geom& g,
                     const float* v_rend, GradBufs& G, const RasterWork* work);
void raster_backward_views(hipStream_t s, const RasterWork& w);
// g2d: the [N G2D] gradient records (their colour fields are read)
void sh_rest_adam(hipStream_t s, const Params& P, const MsplatCamera& cam, const int32_t* radii, const float* g2d,
                  const MsplatStepParams& hp);
void gaussian_update(hipStream_t s, const Params& P, const MsplatCamera& cam, const ProjOut& pj, const GradBufs& G,
                     const MsplatStepParams& hp, uint32_t* nan_counts, float* dbg_v_mean, float* dbg_v_scale,
                     float* dbg_v_quat, float* dbg_v_opac);
// Gradients of a multi-view batched step, summed over its views (cap-sized, row-major).
struct BatchGrads {
    float* v_mean;    // [N 3]
    float* v_scale;   // [N 3]
    float* v_quat;    // [N 4]
    float* v_opac;    // [N]
    float* v_dc;      // [N 3]
    float* v_rest;    // [N nb3]
    uint32_t* vis;    // [N] 1 = a view gave a finite colour gradient
};
// One view's projection VJP and SH colour gradients added into A (views run in order on one
// stream, so the sums need no atomics).
void view_accumulate(hipStream_t s, const Params& P, const ViewConsts* vc, int degrees_to_use, const ProjOut& pj,
                     const GradBufs& G, BatchGrads& A, uint32_t* nan_counts);
// The batch's single update: SH Adam for Gaussians with vis, regulariser, Adam on the other
// groups, SGLD.
void batch_update(hipStream_t s, const Params& P, const BatchGrads& A, const MsplatStepParams* hpd, int degrees_to_use,
                  uint32_t* nan_counts);
void fill_zero(hipStream_t s, void* p, size_t bytes);
// Four ranges (any may be null or empty) zeroed by one launch.
void fill_zero4(hipStream_t s, void* p0, size_t b0, void* p1, size_t b1, void* p2, size_t b2, void* p3, size_t b3);
void copy16(hipStream_t s, const void* src, void* dst, size_t bytes);
// bf16 storage helpers: shadows of f_dc / f_rest from the fp32 masters (rows [0, N)), and
// element-wise conversions.
void sh_refresh_shadows(hipStream_t s, const Params& P);
void f32_to_bf16(hipStream_t s, const float* src, uint16_t* dst, size_t n);
void bf16_to_f32(hipStream_t s, const uint16_t* src, float* dst, size_t n);

}  // namespace msk

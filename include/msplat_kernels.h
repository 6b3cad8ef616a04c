/* msplat_kernels.h — kernel-level C ABI of the MI355X engine.
 *
 * This is the internal boundary the reference calls bindings.h (core/metal/bindings.h:20-112):
 * the host loop hands device-resident parameter and optimiser arrays to one fused training
 * step or one forward render. Here it is exported as plain C over device pointers so that
 * callers that own their memory (tests, the Python bindings, other frameworks) can drive
 * the same kernels the C API (msplat_c_api.h) drives.
 *
 * All pointers in MsplatDeviceState are device pointers to row-major float32 arrays of
 * N rows: means [N,3], scales [N,3] (log), quats [N,4] (w,x,y,z), f_dc [N,3],
 * f_rest [N,(deg+1)^2-1,3], opac [N] (logit); m[g], v[g] are the Adam moments of group g
 * in the order means, scales, quats, f_dc, f_rest, opacity (same shapes). The work is
 * enqueued on the engine's own HIP stream; msplat_engine_sync() waits for it.
 */
#ifndef MSPLAT_KERNELS_H
#define MSPLAT_KERNELS_H

#include <stddef.h>
#include <stdint.h>
#include "msplat_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t N, degree;
    float *means, *scales, *quats, *f_dc, *f_rest, *opac;
    float* m[6];
    float* v[6];
    /* bfloat16 SH storage (0 = everything fp32). When set, f_dc_bf16 / f_rest_bf16 hold the
     * bf16 shadows of the fp32 masters f_dc / f_rest (the forward reads the shadows; the
     * Adam update refreshes them), and m[3], v[3], m[4], v[4] point to bf16 arrays. */
    int32_t sh_bf16;
    uint16_t* f_dc_bf16;
    uint16_t* f_rest_bf16;
} MsplatDeviceState;

typedef struct MsplatEngine MsplatEngine;

/* Device memory helpers (hipMalloc / hipFree / synchronous copies on the current device),
 * so that callers need no second GPU runtime to own the arrays they pass in. */
int msplat_dev_count(void);
void* msplat_dev_alloc(size_t bytes);
void msplat_dev_free(void* p);
int msplat_dev_upload(void* dst, const void* src, size_t bytes);
int msplat_dev_download(void* dst, const void* src, size_t bytes);
int msplat_dev_memset(void* dst, int value, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (bytes rounded down to 16); the
 * calibration point for HBM traffic counters. Waits for completion. */
int msplat_dev_copy_kernel(void* dst, const void* src, size_t bytes);

/* Waves per workgroup of the block-list (level-1) bucket kernel of the tile-list build (one
 * chunk of ranks per wave, 4 bytes of LDS per 4x4-tile block per wave): 0 (default) = 4
 * where the LDS allows, else 2, else 1; 1, 2 or 4 on request, falling back to fewer where
 * the LDS does not allow it. Process-wide; the lists are the same for every
 * choice, so tests use it to cover each variant. No counterpart in the reference (its list
 * build has no such variants). Returns -1 for any other value. */
int msplat_set_list_bucket_waves(int waves);
/* The waves per workgroup the bucket kernel launches with for `bins` blocks under a device
 * LDS limit of `lds_limit` bytes per workgroup (request as above); 0 when even one wave's
 * LDS does not fit. Host-only arithmetic (no device is touched); no counterpart in the
 * reference. */
int msplat_list_bucket_waves_for(int bins, int64_t lds_limit, int request);

/* Creates an engine bound to the current HIP device (buffers grow on demand and are
 * reused across steps). Replaces the reference's global MetalContext + FusedTensorCache
 * (core/metal/msplat_metal.mm:41-183, 492-625). NULL if no HIP device is usable. */
MsplatEngine* msplat_engine_create(void);
void msplat_engine_destroy(MsplatEngine* e);

/* One training step: msplat_train_step (reference bindings.h:71-95, msplat_metal.mm:1118-1740)
 * followed by the SGLD update of msplat_sgld_noise_gen / msplat_sgld_noise
 * (bindings.h:98-105) when hp->sgld_enable. gt is a device uint8 [H,W,3] image.
 * debug != 0 additionally keeps the per-Gaussian gradients fed to Adam.
 * Returns 0, or a negative error code (-1 bad arguments, -2 HIP error). */
int msplat_engine_train_step(MsplatEngine* e, const MsplatDeviceState* st, const MsplatCamera* cam,
                             const uint8_t* gt, const MsplatStepParams* hp, int debug);

/* msplat_engine_train_step with a per-pixel training mask: the mask / has_mask arguments of
 * msplat_train_step (reference bindings.h:71-95, msplat_metal.mm:1118-1740; kernels
 * msplat_metal.metal:3602-3653 and 3815-3834). mask is a device uint8 [H,W] image or NULL
 * (the unmasked step). A pixel whose mask value is below 128 owns no loss: its SSIM
 * derivative fields and its SSIM and L1 loss terms are zero and its own L1 gradient is
 * dropped, while SSIM gradients of unmasked neighbours' windows still reach it; the loss
 * value still adds the SSIM weight for it. The caller sets hp->loss_inv_n to
 * 1 / (3 * unmasked pixels), as the reference's model does (model.cpp:929-934). */
int msplat_engine_train_step_masked(MsplatEngine* e, const MsplatDeviceState* st, const MsplatCamera* cam,
                                    const uint8_t* gt, const uint8_t* mask, const MsplatStepParams* hp, int debug);

/* Forward render only: msplat_render (reference bindings.h:58-67). The image is left in
 * the engine buffer "img" (device float [H,W,3]). */
int msplat_engine_render(MsplatEngine* e, const MsplatDeviceState* st, const MsplatCamera* cam,
                         int degrees_to_use, const float background[3]);

/* msplat_engine_render for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in
 * the reference. It allocates nothing: every intermediate buffer (projection records, depth
 * rank, tile lists, raster work queue, depth-chunk buffers) lives in a 64 MiB array in
 * device memory that the library holds on each GPU it runs on until the process ends. The
 * caller passes the state and device outputs img [H,W,3] f32, final_T [H,W] f32 and
 * final_idx [H,W] i32, identical to the engine's buffers of the same name after
 * msplat_engine_render; the work is enqueued on `stream` (a hipStream_t, NULL: the null
 * stream) and nothing waits for it. A render whose buffers exceed the array (roughly
 * 70k Gaussians at 1920x1080 on the monolithic path; depth-chunked views of fewer than 400
 * tiles need about 3.4 KB per Gaussian, so about 16k) is rejected with -3 and a message on
 * stderr. Calls on one GPU share its array and overlapping calls on one GPU (two streams,
 * two threads) race. Returns 0, -1 bad arguments, -2 HIP error, -3 too large. */
int msplat_render_noalloc(const MsplatDeviceState* st, const MsplatCamera* cam, int degrees_to_use,
                          const float background[3], float* img, float* final_T, int32_t* final_idx, void* stream);

/* Device pointer and size of an internal buffer of the last step/render: "radii",
 * "tiles_hit", "depths", "rect", "aabb", "recA", "recB", "recC", "tile_counts",
 * "tile_bins", "pkA", "pkB", "pkC", "img", "final_T", "final_idx", "v_rendered", "grad2d"
 * (float [N 9] per Gaussian: colour 0-2, conic 3-5, mean2d 6-7, opacity logit 8, the raster
 * backward's screen-space gradients), "dbg_v_mean", "dbg_v_scale", "dbg_v_quat", "dbg_v_opac",
 * "nan_counts". Returns 0 on success, -1 for an unknown name. Replaces the reference's
 * diagnostics msplat_last_tile_counts / msplat_dump_step_buffers (msplat_metal.mm:633-736). */
int msplat_engine_buffer(MsplatEngine* e, const char* name, void** dptr, size_t* bytes);

/* Copies min(bytes, size of the buffer) bytes of an internal buffer into host memory,
 * synchronising with the engine stream. Returns the number of bytes copied, or -1. */
int64_t msplat_engine_read(MsplatEngine* e, const char* name, void* host, size_t bytes);

/* Waits for the engine's stream. Returns 0 or -2 on a HIP error. */
int msplat_engine_sync(MsplatEngine* e);

/* After msplat_engine_sync: out[0] = intersections I, out[1] = padded key slots,
 * out[2] = key-buffer overflow flag, out[3] = largest tile count, out[4] = K (depth chunks),
 * out[5] = number of tiles. */
int msplat_engine_stats(MsplatEngine* e, int64_t out[6]);

/* Loss of the last training step (sum of the per-pixel losses / (H W)), synchronising. */
double msplat_engine_loss(MsplatEngine* e);

/* The engine's HIP stream (hipStream_t as void*). */
void* msplat_engine_stream(MsplatEngine* e);

/* Time of the stream's kernels between the two most recent msplat_engine_mark() calls,
 * in milliseconds (HIP events recorded on the engine stream). */
void msplat_engine_mark(MsplatEngine* e);
float msplat_engine_elapsed_ms(MsplatEngine* e);

/* Per-kernel timing: when enabled, HIP events bracket each stage of each step on the
 * engine stream; msplat_engine_stage_times fills ms[k] with the accumulated time of stage
 * k since the last reset and returns the number of stages (names via
 * msplat_engine_stage_name). */
void msplat_engine_set_stage_timing(MsplatEngine* e, int enable);
int msplat_engine_stage_times(MsplatEngine* e, double* ms, int max_stages, int64_t* launches);
const char* msplat_engine_stage_name(int k);
void msplat_engine_reset_stage_times(MsplatEngine* e);

/* ---- Learned relocation proposal: the network's kernels on device buffers (no
 * counterpart in the reference; used by the trainer's learned relocation and by tests).
 * Parameters theta: 5508 floats, W1 [64 16], b1 [64], W2 [64 64], b2 [64], W3 [4 64],
 * b3 [4] (row-major [out][in]); inputs X: [N 16] bfloat16 bit patterns; targets: [N 8]
 * floats (y, logit weight, offset target x3, offset weight, 2 unused). All pointers are
 * device pointers; every call waits for its kernels. Return 0, -1 (bad argument) or -2. */
size_t msplat_mlp_packed_bytes(void);
int msplat_mlp_pack(const float* theta, void* packed);
/* out [N 4] = (logit, offset x, y, z). */
int msplat_mlp_forward(int n, const uint16_t* X, const void* packed, float* out);
/* grad [5508] += d loss / d theta, loss2 [2] += (logit term, offset term). */
int msplat_mlp_backward(int n, const uint16_t* X, const float* targets, const void* packed, float* grad,
                        double* loss2);

#ifdef __cplusplus
}
#endif
#endif

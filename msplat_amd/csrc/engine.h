// engine.h — device buffer cache and per-step kernel sequencing (one HIP stream).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "kernels.h"
#include "../../include/msplat_kernels.h"

namespace msplat_hip {

void hip_check(hipError_t e, const char* what);

struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    void ensure(size_t bytes);
    void release();
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

enum Stage {
    ST_ZERO = 0, ST_PROJECT, ST_RANK, ST_COUNTS, ST_BUCKET, ST_RASTER_FWD, ST_SSIM, ST_RASTER_BWD,
    ST_SH_ADAM, ST_UPDATE, ST_COUNT
};

// Pinned staging ring for the small per-step constants (view constants, step scalars):
// slot k is refilled only after the event recorded behind its previous copy has completed,
// so the host never overwrites bytes a copy has not read yet.
class ConstRing {
public:
    ConstRing();
    ~ConstRing();
    void upload(hipStream_t s, void* dst, const void* src, size_t bytes);
    static constexpr int kSlots = 64;
    static constexpr size_t kSlotBytes = 8192;

private:
    uint8_t* host_ = nullptr;
    hipEvent_t ev_[kSlots] = {};
    bool armed_[kSlots] = {};
    int next_ = 0;
};

class Engine {
public:
    // shared: run on this stream (owned by the caller) instead of a stream of its own.
    explicit Engine(hipStream_t shared = nullptr);
    ~Engine();
    // mask: device uint8 [H W] training mask (pixels below 128 own no loss) or null.
    int train_step(const MsplatDeviceState& st, const MsplatCamera& cam, const uint8_t* gt, const MsplatStepParams& hp,
                   bool debug, const uint8_t* mask = nullptr);
    // The view-dependent part of a step: gradient accumulators cleared, forward, loss (into
    // loss slot `slot`), raster backward. Leaves the projection, records and raster
    // gradients of this view in the engine's buffers.
    int forward_backward(const MsplatDeviceState& st, const MsplatCamera& cam, const uint8_t* gt,
                         const MsplatStepParams& hp, int slot, const uint8_t* mask = nullptr);
    // Adds this view's projection VJP and SH colour gradients (after its backward).
    void accumulate(const MsplatDeviceState& st, int degrees_to_use, msk::BatchGrads& acc);
    // Loss slot k as the mean over its view's pixels (waits for the stream).
    double loss_slot(int k, int W, int H);
    static constexpr int kLossSlots = 64;
    uint32_t* nan_counts() { return nanc_.as<uint32_t>(); }
    int render(const MsplatDeviceState& st, const MsplatCamera& cam, int degrees_to_use, const float bg[3]);
    int buffer(const std::string& name, void** p, size_t* bytes);
    int sync();
    void stats(int64_t out[6]);
    double loss();
    hipStream_t stream() const { return stream_; }
    void mark();
    float elapsed_ms();
    void set_stage_timing(bool on);
    int stage_times(double* ms, int max_stages, int64_t* launches);
    void reset_stage_times();
    const float* image() const { return img_.as<float>(); }
    // Releases every size-dependent buffer; the next step or render re-allocates them.
    void trim();
    int last_W = 0, last_H = 0;

    // ---- one view of a batched step (the batch runner drives these per view) ----
    // Buffers for N Gaussians and launch geometry `launch`; `view` is this view's own.
    void prepare(int N, const msk::Geom& launch, const msk::Geom& view, bool train);
    static msk::ViewConsts make_view(const MsplatCamera& cam, const uint8_t* gt, const float bg[3], float ssim_w,
                                     float inv_n, const uint8_t* mask = nullptr);
    // The view's constants live at `d` (already on the device, written by the caller).
    void bind_view(const msk::ViewConsts& h, const msk::ViewConsts* d);
    // The view's constants in the engine's own block (with its raster view table), so that
    // raster_fwd_own / raster_bwd_own run this view alone.
    void own_view(const msk::ViewConsts& h) { upload_view(h, true); }
    // Gradient accumulators cleared (train), projection, tile lists.
    // own_queue: also write this view's raster work queue (false when a batch builds one
    // queue for all its views)
    void front(const msk::Params& P, int degrees_to_use, bool train, bool own_queue = true);
    void loss_to(int slot);
    bool chunked() const { return K_ > 1; }
    int chunks() const { return (int)K_; }
    // Raster forward / backward of this view alone (chunked views always go this way).
    void raster_fwd_own();
    void raster_bwd_own();
    msk::RasterView raster_view(bool train) const;
    int64_t generation() const { return gen_; }
    // This view's projection outputs and raster gradients (valid after its backward).
    const int32_t* radii_dev() const { return radii_.as<int32_t>(); }
    const float* depths_dev() const { return depths_.as<float>(); }
    const float* g2d_dev() const { return g2d_.as<float>(); }
    void set_persist_grid(int g) { persist_grid_ = g; }
    int persist_grid() const { return persist_grid_; }
    // Pixels per lane of the persistent raster kernels (1 or 4; the image is the same).
    void set_raster_ppl(int p) { ppl_ = p == 4 ? 4 : 1; }
    // Width of the tile-list keys: 8 carries the Gaussian id with the depth rank, 4 looks
    // it up after the per-tile sort (the lists are the same).
    int raster_ppl() const { return ppl_; }
    // Tile-list launch choices (the lists are the same for every choice): rank chunks of the
    // count and bucket kernels (0: default), waves per bucket workgroup (0: process-wide
    // default), and one bucket launch per tile-count class (for recorded sequences that serve
    // views smaller than their launch geometry).
    void set_list_chunks(int c) { chunk_target_ = c > 0 ? c : 0; }
    void set_list_bucket_waves(int w) { bucket_waves_ = (w == 1 || w == 2 || w == 4) ? w : 0; }

private:
    void upload_view(const msk::ViewConsts& h, bool train);
    msk::ListBufs list_bufs();
    msk::RasterWork own_work();
    void stage_event(int stage_end);
    void flush_events();
    void ensure(DBuf& b, size_t bytes);

    hipStream_t stream_ = nullptr;
    bool owns_stream_ = true;
    int N_ = 0, W_ = 0, H_ = 0, tbx_ = 0, tby_ = 0;
    msk::Geom launch_{}, view_{};
    int64_t capacity_ = 0;
    uint32_t K_ = 1;
    int64_t gen_ = 0;   // bumps whenever a buffer is re-allocated
    int persist_grid_ = 0;
    int ppl_ = 1;
    int chunk_target_ = 0, bucket_waves_ = 0;
    // view constants: host copy, device copy in use, own device block [ViewConsts | RasterView]
    msk::ViewConsts vch_{};
    const msk::ViewConsts* vcd_ = nullptr;
    DBuf consts_;
    ConstRing ring_;
    // per-Gaussian
    DBuf radii_, tiles_, depths_, rect_, aabb_, recA_, recB_, recC_;
    bool queue_built_ = false;  // front() wrote the raster work queue of the current lists
    DBuf g2d_;  // [N G2D] screen-space gradient records
    DBuf dbg_vm_, dbg_vs_, dbg_vq_, dbg_vo_;
    // per-tile
    DBuf counts_, exact_, bins_, totals_;
    DBuf dkeys_, dkeys_out_, dvals_, sorted_ids_, temp_, srect_;
    DBuf wq_, wmisc_;   // own persistent-raster work queue
    // per-intersection
    DBuf ids_, l1_, l1m_, blocks_, hist_;
    // per-pixel
    DBuf img_, finalT_, finalIdx_, ncontrib_, vrend_, dbuf_, partials_, loss_;
    DBuf chunkT_, chunkC_, chunkIdx_, prefixT_, afterC_;
    DBuf nanc_;
    bool debug_valid_ = false;
    msk::ListBufs lists_{};
    // timing
    hipEvent_t mark_[2] = {nullptr, nullptr};
    int mark_n_ = 0;
    bool stage_timing_ = false;
    std::vector<hipEvent_t> ev_pool_;
    std::vector<int> ev_stage_;  // stage that ends at this event (-1 = step start)
    size_t ev_used_ = 0;
    double stage_ms_[ST_COUNT] = {0};
    int64_t stage_n_[ST_COUNT] = {0};
};

// Resident workgroups of a persistent raster launch by default: 8 per CU.
int default_persist_grid();
// Depth chunks per tile of a view with `num_tiles` tiles and a list capacity of `capacity`
// slots (1: the monolithic raster path).
uint32_t k_max_rule(int64_t capacity, int num_tiles);
msk::Params params_of(const MsplatDeviceState& st);

// Every live engine (msplat_sync / msplat_cleanup act on all of them).
void trim_all_engines();
void sync_all_engines();

}  // namespace msplat_hip

// Handle type of the kernel-level C ABI; `owned` is false for an engine borrowed from a
// trainer (msplat_engine_destroy then frees only the handle).
struct MsplatEngine {
    msplat_hip::Engine* e;
    bool owned;
};

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
    ST_ZERO = 0, ST_PROJECT, ST_RANK, ST_COUNTS, ST_SCATTER, ST_SORT, ST_RASTER_FWD, ST_SSIM, ST_RASTER_BWD,
    ST_SH_ADAM, ST_UPDATE, ST_COUNT
};

class Engine {
public:
    Engine();
    ~Engine();
    int train_step(const MsplatDeviceState& st, const MsplatCamera& cam, const uint8_t* gt, const MsplatStepParams& hp,
                   bool debug);
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

private:
    void prepare(int N, int W, int H, int nb3, bool train);
    void forward(const msk::Params& P, const MsplatCamera& cam, int degrees_to_use, const float bg[3]);
    msk::ListBufs list_bufs();
    void stage_event(int stage_end);
    void flush_events();

    hipStream_t stream_ = nullptr;
    int N_ = 0, W_ = 0, H_ = 0, tbx_ = 0, tby_ = 0;
    int64_t capacity_ = 0;
    uint32_t K_ = 1;
    // per-Gaussian
    DBuf radii_, tiles_, depths_, rect_, aabb_, recA_, recB_, recC_;
    DBuf v_xy_, v_conic_, v_rgb_, v_opac_;
    DBuf dbg_vm_, dbg_vs_, dbg_vq_, dbg_vo_;
    // per-tile
    DBuf counts_, exact_, bins_, totals_;
    DBuf dkeys_, dkeys_out_, dvals_, sorted_ids_, temp_;
    // per-intersection
    DBuf keys_, ids_, hist_;
    // per-pixel
    DBuf img_, finalT_, finalIdx_, vrend_, hbuf_, dbuf_, partials_, loss_;
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

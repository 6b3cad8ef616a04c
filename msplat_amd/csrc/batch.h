// batch.h — the multi-view batched training step: B views per step, each with its own
// engine slot (lists, images, gradients; slots take the views by decreasing size), one
// persistent raster launch over all views' tiles in each direction, gradient sums in slot
// order and one update. The whole
// iteration can be recorded once as a HIP graph and replayed: every per-step input the
// kernels read (cameras, targets, step scalars, the raster view table) sits in one device
// block that is rewritten before each replay.
#pragma once
#include <memory>
#include <vector>

#include "engine.h"

namespace msplat_hip {

struct BatchView {
    MsplatCamera cam;
    const uint8_t* gt;   // device [H W 3]
    float inv_n;         // loss normalisation of this view
    const uint8_t* mask = nullptr;  // device [H W] training mask or null
};

class BatchRunner {
public:
    explicit BatchRunner(hipStream_t s);
    ~BatchRunner();
    BatchRunner(const BatchRunner&) = delete;
    BatchRunner& operator=(const BatchRunner&) = delete;

    // One step over views[0..B). `A` holds the cap-sized gradient sums (zeroed here).
    void run(const MsplatDeviceState& st, const BatchView* views, int B, const MsplatStepParams& hp,
             msk::BatchGrads A, uint32_t* nan_counts);
    // Mean loss of view v of the last step (waits for the stream).
    double loss(int v);
    void set_graph(bool on) { graph_on_ = on; }
    // Views one after another through the single-view path (lists, raster forward, loss,
    // raster backward per view) instead of one persistent launch over all of them; never
    // recorded. The baseline the multi-view path is measured against.
    void set_sequential(bool on) { sequential_ = on; }
    bool graph() const { return graph_on_; }
    // Work-queue residency of the persistent raster kernels (0 = default).
    void set_persist_grid(int g);
    int persist_grid() const { return persist_grid_; }
    // Autotuner of the launch choices the results do not depend on. Stage one: pixels per
    // lane (the raster sub-tile a lane covers: 1 pixel of a 256-lane workgroup per tile, or
    // 2x2 pixels of one wave per tile) x residency of the persistent raster kernels. Stage
    // two, with the best of stage one: rank chunks of the tile histogram x waves per
    // block-list bucket workgroup. Stage three (when graphs are on): the best so far, run
    // eagerly instead of replayed from a recorded graph. The first steps cycle through the
    // candidates, each timed per megapixel over three steps after the one that records its
    // graph, and the fastest is kept for the remaining steps.
    void set_autotune(bool on);
    bool tuned() const { return tune_done_; }
    struct Choice {
        int ppl = 1, grid = 0, chunks = 0, waves = 0;
        int graph = 1;  // replay a recorded graph (when graphs are on), else launch eagerly
        bool operator==(const Choice& o) const {
            return ppl == o.ppl && grid == o.grid && chunks == o.chunks && waves == o.waves && graph == o.graph;
        }
    };
    struct TuneResult {
        Choice c;
        float ms;
    };
    const std::vector<TuneResult>& tune_results() const { return tune_ms_; }
    Choice choice() const { return {ppl_, persist_grid_, chunks_, bucket_waves_, graph_choice_ ? 1 : 0}; }
    void set_raster_ppl(int p);
    int raster_ppl() const { return ppl_; }
    // Rank chunks of the tile-list kernels and waves per bucket workgroup (0: defaults).
    void set_list_choice(int chunks, int waves);
    // Counters: steps replayed from a recorded graph, graphs recorded.
    int64_t replays() const { return replays_; }
    int64_t captures() const { return captures_; }
    // Tile-count geometry the recorded graph is launched with (grows monotonically).
    msk::Geom launch_geom() const { return hw_; }
    // Engine of slot v of the last step (slots hold the views by decreasing size).
    Engine* slot(int v) { return v >= 0 && v < (int)slots_.size() ? slots_[v].get() : nullptr; }

private:
    struct Key {
        int N = -1, B = 0, du = -1, degree = -1, bf16 = 0, grid = 0, ppl = 1, chunks = 0, waves = 0;
        msk::Geom g{};
        std::vector<int> slot_geom;   // recorded (W, H) per slot
        std::vector<int64_t> gens;    // (buffer generation, depth-chunk count) per slot
        const void* state[8] = {};
        const void* mom[12] = {};
        const void* own[3] = {};
        const void* acc[7] = {};
        const void* nanc = nullptr;
        bool operator==(const Key& o) const;
    };
    void enqueue(const msk::Params& P, int B, int du, int Tmax, msk::BatchGrads A, uint32_t* nan_counts);
    void release_graph();

    hipStream_t s_;
    std::vector<std::unique_ptr<Engine>> slots_;
    std::vector<std::pair<int, int>> sizes_;   // per slot
    std::vector<int> slot_of_view_;
    std::vector<msk::Geom> hwv_;               // high-water launch geometry per slot (graph mode)
    DBuf block_;                                     // [MsplatStepParams | ViewConsts x B | RasterView x B]
    DBuf wq_, wmisc_;     // batch work queue
    ConstRing ring_;
    bool graph_on_ = false;
    bool graph_choice_ = true;  // the autotuner's choice (replay or eager) when graphs are on
    bool sequential_ = false;
    int persist_grid_ = 0;
    msk::Geom hw_{};                                 // high-water launch geometry (graph mode)
    hipGraphExec_t exec_ = nullptr;
    bool autotune_ = false, tune_done_ = false;
    int tune_step_ = 0;
    std::vector<Choice> tune_cands_;
    std::vector<TuneResult> tune_ms_;
    struct TuneObs {
        int c;
        double ms, px;
    };
    std::vector<TuneObs> tune_obs_;  // every timed step: candidate, ms, pixels of its views
    void tune_score();
    int tune_stage1_ = 0;  // candidates of stage one
    int tune_stage2_ = 0;  // candidates of stages one and two
    int tune_stage_lo_ = 0, tune_stage_step_ = 0;  // the running stage's first candidate and step
    int ppl_ = 1, chunks_ = 0, bucket_waves_ = 0;
    void apply(const Choice& c);
    hipEvent_t tev_[2] = {nullptr, nullptr};
    void run_inner(const MsplatDeviceState& st, const BatchView* views, int B, const MsplatStepParams& hp,
                   msk::BatchGrads A, uint32_t* nan_counts);
    Key key_;
    int64_t replays_ = 0, captures_ = 0;
};

}  // namespace msplat_hip

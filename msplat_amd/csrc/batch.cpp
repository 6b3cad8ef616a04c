// batch.cpp — multi-view batched step (no counterpart in the reference, which trains one
// view per step: Model::fullIteration, core/src/model.cpp:901-980). Semantics: every view
// runs the single-view forward/backward on the same parameters; the projection VJPs and SH
// colour gradients are summed over the views in order of decreasing size (ties in view
// order); one Adam/SGLD update follows, SH Adam for the Gaussians that had a finite colour
// gradient in at least one view.
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace msplat_hip {

bool BatchRunner::Key::operator==(const Key& o) const {
    if (N != o.N || B != o.B || du != o.du || degree != o.degree || bf16 != o.bf16 || grid != o.grid || ppl != o.ppl ||
        chunks != o.chunks || waves != o.waves)
        return false;
    if (g.W != o.g.W || g.H != o.g.H || g.T != o.g.T || gens != o.gens || nanc != o.nanc) return false;
    if (slot_geom != o.slot_geom) return false;
    for (int k = 0; k < 8; k++)
        if (state[k] != o.state[k]) return false;
    for (int k = 0; k < 7; k++)
        if (acc[k] != o.acc[k]) return false;
    for (int k = 0; k < 12; k++)
        if (mom[k] != o.mom[k]) return false;
    for (int k = 0; k < 3; k++)
        if (own[k] != o.own[k]) return false;
    return true;
}

BatchRunner::BatchRunner(hipStream_t s) : s_(s) { persist_grid_ = default_persist_grid(); }

BatchRunner::~BatchRunner() {
    (void)hipStreamSynchronize(s_);
    release_graph();
    for (auto& e : tev_)
        if (e) (void)hipEventDestroy(e);
    DBuf* all[] = {&block_, &wq_, &wmisc_};
    for (DBuf* b : all) b->release();
}

void BatchRunner::release_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
    key_ = Key();
}

void BatchRunner::set_persist_grid(int g) {
    persist_grid_ = g > 0 ? g : default_persist_grid();
    for (auto& e : slots_) e->set_persist_grid(persist_grid_);
}

void BatchRunner::set_raster_ppl(int p) {
    ppl_ = p == 4 ? 4 : 1;
    for (auto& e : slots_) e->set_raster_ppl(ppl_);
}

void BatchRunner::set_list_choice(int chunks, int waves) {
    chunks_ = chunks > 0 ? chunks : 0;
    bucket_waves_ = (waves == 1 || waves == 2 || waves == 4) ? waves : 0;
    for (auto& e : slots_) {
        e->set_list_chunks(chunks_);
        e->set_list_bucket_waves(bucket_waves_);
    }
}

void BatchRunner::apply(const Choice& c) {
    set_raster_ppl(c.ppl);
    set_persist_grid(c.grid);
    set_list_choice(c.chunks, c.waves);
    graph_choice_ = c.graph != 0;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The launch sequence of one batched step (recorded once per key in graph mode).
void BatchRunner::enqueue(const msk::Params& P, int B, int du, int Tmax, msk::BatchGrads A, uint32_t* nan_counts) {
    const size_t n = (size_t)(P.N > 0 ? P.N : 1);
    const int nb3 = P.degree <= 0 ? 0 : (P.degree == 1 ? 9 : P.degree == 2 ? 24 : 45);
    msk::fill_zero(s_, A.v_mean, n * 12);
    msk::fill_zero(s_, A.v_scale, n * 12);
    msk::fill_zero(s_, A.v_quat, n * 16);
    msk::fill_zero(s_, A.v_opac, n * 4);
    msk::fill_zero(s_, A.v_dc, n * 12);
    msk::fill_zero(s_, A.v_rest, n * (size_t)std::max(nb3, 1) * 4);
    msk::fill_zero(s_, A.vis, n * 4);
    for (int v = 0; v < B; v++) slots_[v]->front(P, du, true, false);  // the batch builds one queue for all
    msk::RasterWork w{};
    w.views = reinterpret_cast<const msk::RasterView*>(block_.as<uint8_t>() + align256(sizeof(MsplatStepParams)) +
                                                       align256(sizeof(msk::ViewConsts) * B));
    w.B = B;
    w.Tmax = Tmax;
    w.queue = wq_.as<uint32_t>();
    w.n_work = wmisc_.as<uint32_t>();
    w.ctr = wmisc_.as<uint32_t>() + 1;
    w.grid = persist_grid_;
    w.ppl = ppl_;
    for (int v = 0; v < B; v++)
        if (slots_[v]->chunked()) slots_[v]->raster_fwd_own();
    msk::raster_work_build(s_, w);
    msk::raster_forward_views(s_, w);
    for (int v = 0; v < B; v++) slots_[v]->loss_to(0);
    for (int v = 0; v < B; v++)
        if (slots_[v]->chunked()) slots_[v]->raster_bwd_own();
    msk::raster_backward_views(s_, w);
    for (int v = 0; v < B; v++) {
        MsplatDeviceState st{};
        st.means = P.means; st.scales = P.scales; st.quats = P.quats; st.f_dc = P.f_dc; st.f_rest = P.f_rest;
        st.opac = P.opac;
        st.N = P.N;
        st.degree = P.degree;
        st.sh_bf16 = P.sh_bf16;
        st.f_dc_bf16 = P.f_dc_h;
        st.f_rest_bf16 = P.f_rest_h;
        slots_[v]->accumulate(st, du, A);
    }
    msk::batch_update(s_, P, A, block_.as<MsplatStepParams>(), du, nan_counts);
}

void BatchRunner::set_autotune(bool on) {
    autotune_ = on;
    tune_done_ = !on;
    tune_step_ = 0;
    tune_ms_.clear();
    tune_obs_.clear();
    tune_cands_.clear();
    if (on) {
        const int cus = default_persist_grid() / 8;  // the default is 8 workgroups per CU
        // the default first: a later candidate replaces the best only by a clear margin
        for (int k : {8, 2, 4, 16}) tune_cands_.push_back({1, cus * k, 0, 0, 1});
        for (int k : {8, 16, 32}) tune_cands_.push_back({4, cus * k, 0, 0, 1});
    }
    tune_stage1_ = (int)tune_cands_.size();
    tune_stage2_ = 0;
    tune_stage_lo_ = 0;
    tune_stage_step_ = 0;
}

// Stage two's candidates around the best of stage one (whose default list choice is
// already timed): tile-histogram rank chunks 384 / 1536 (default 768) and 2 waves per
// block-list bucket workgroup (default 4).
static std::vector<BatchRunner::Choice> stage_two(const BatchRunner::Choice& b) {
    std::vector<BatchRunner::Choice> out;
    for (const auto& cw : {std::make_pair(384, 0), std::make_pair(1536, 0), std::make_pair(0, 2), std::make_pair(384, 2)})
        out.push_back({b.ppl, b.grid, cw.first, cw.second, 1});
    return out;
}

// Each candidate runs kTuneSteps steps: the first records its graph (untimed), the others
// are timed (a stage's candidates take turns). The views (and their sizes) change from step to step, so a step's time is
// judged against what a step of its pixel count takes: a least-squares line
// ms = a + b * pixels through every timed step so far (all candidates: a is the per-step
// cost that does not scale with pixels, the Gaussians' projection, lists and update), and a
// candidate's score is the mean of its steps' measured / fitted ms (1 = an average step).
constexpr int kTuneSteps = 4;

void BatchRunner::tune_score() {
    double n = 0.0, sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0;
    for (const TuneObs& o : tune_obs_) {
        n += 1.0;
        sx += o.px;
        sy += o.ms;
        sxx += o.px * o.px;
        sxy += o.px * o.ms;
    }
    const double den = n * sxx - sx * sx;
    double b = den > 0.0 ? (n * sxy - sx * sy) / den : 0.0;
    double a = n > 0.0 ? (sy - b * sx) / n : 0.0;
    if (!(a > 0.0) || !(b >= 0.0)) {  // a degenerate fit: time per pixel
        a = 0.0;
        b = sx > 0.0 ? sy / sx : 1.0;
    }
    std::vector<double> sum(tune_ms_.size(), 0.0), cnt(tune_ms_.size(), 0.0);
    for (const TuneObs& o : tune_obs_) {
        sum[o.c] += o.ms / (a + b * o.px);
        cnt[o.c] += 1.0;
    }
    for (size_t k = 0; k < tune_ms_.size(); k++) tune_ms_[k].ms = cnt[k] > 0.0 ? (float)(sum[k] / cnt[k]) : 0.f;
}

void BatchRunner::run(const MsplatDeviceState& st, const BatchView* views, int B, const MsplatStepParams& hp,
                      msk::BatchGrads A, uint32_t* nan_counts) {
    if (!autotune_ || tune_done_) {
        run_inner(st, views, B, hp, A, nan_counts);
        return;
    }
    // Within a stage the candidates take turns, step by step (the scene and so the step time
    // drift while tuning runs; taking turns spreads that drift over all of them): the first
    // round records each candidate's graph, the others are timed.
    const int lo = tune_stage_lo_, ns = (int)tune_cands_.size() - lo;
    const int t = tune_step_ - tune_stage_step_;
    const int c = lo + t % ns;
    apply(tune_cands_[c]);
    const bool timed = t >= ns;
    while ((int)tune_ms_.size() < (int)tune_cands_.size()) tune_ms_.push_back({tune_cands_[tune_ms_.size()], 0.f});
    if (timed) {
        if (!tev_[0]) {
            hip_check(hipEventCreate(&tev_[0]), "hipEventCreate");
            hip_check(hipEventCreate(&tev_[1]), "hipEventCreate");
        }
        hip_check(hipEventRecord(tev_[0], s_), "hipEventRecord");
    }
    run_inner(st, views, B, hp, A, nan_counts);
    if (timed) {
        hip_check(hipEventRecord(tev_[1], s_), "hipEventRecord");
        hip_check(hipEventSynchronize(tev_[1]), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, tev_[0], tev_[1]), "hipEventElapsedTime");
        double px = 0.0;
        for (auto& wh : sizes_) px += (double)wh.first * wh.second;
        tune_obs_.push_back({c, (double)ms, px});
    }
    // Scores within 1% of each other are within the noise of a few steps of changing views:
    // the earlier candidate (the defaults come first) keeps its place unless beaten by more.
    auto best_of = [&](int upto) {
        TuneResult best = tune_ms_.front();
        for (int k = 0; k < upto; k++)
            if (tune_ms_[k].ms < 0.99f * best.ms) best = tune_ms_[k];
        return best;
    };
    if (++tune_step_ - tune_stage_step_ >= kTuneSteps * ns) {
        tune_score();
        tune_stage_lo_ = (int)tune_cands_.size();
        tune_stage_step_ = tune_step_;
        if ((int)tune_cands_.size() == tune_stage1_) {
            // stage one done: stage two around its best
            for (const Choice& ch : stage_two(best_of(tune_stage1_).c)) tune_cands_.push_back(ch);
            tune_stage2_ = (int)tune_cands_.size();
        } else if ((int)tune_cands_.size() == tune_stage2_ && graph_on_) {
            // stage two done: the best so far, launched eagerly
            Choice eager = best_of((int)tune_ms_.size()).c;
            eager.graph = 0;
            tune_cands_.push_back(eager);
        } else {
            apply(best_of((int)tune_ms_.size()).c);
            tune_done_ = true;
        }
    }
}

void BatchRunner::run_inner(const MsplatDeviceState& st, const BatchView* views, int B, const MsplatStepParams& hp,
                            msk::BatchGrads A, uint32_t* nan_counts) {
    if (B <= 0 || B > Engine::kLossSlots) throw std::invalid_argument("batch size must be in [1, 64]");
    const msk::Params P = params_of(st);
    const int du = hp.degrees_to_use < st.degree ? hp.degrees_to_use : st.degree;
    while ((int)slots_.size() < B) {
        slots_.emplace_back(new Engine(s_));
        slots_.back()->set_persist_grid(persist_grid_);
        slots_.back()->set_raster_ppl(ppl_);
        slots_.back()->set_list_chunks(chunks_);
        slots_.back()->set_list_bucket_waves(bucket_waves_);
    }
    for (int v = 0; v < B; v++) {
        const MsplatCamera& c = views[v].cam;
        if (c.width <= 0 || c.height <= 0 || !views[v].gt) throw std::invalid_argument("batched view without image");
        if (!msk::lists_supported(st.N, msk::geom_of(c.width, c.height).T))
            throw std::invalid_argument("view too large for the tile lists");
    }
    // slot k runs view order[k]: views by decreasing size (ties in view order), so that each
    // slot's recorded geometry tracks the sizes that rank of the batch actually takes
    std::vector<int> order((size_t)B);
    for (int v = 0; v < B; v++) order[v] = v;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (int64_t)views[a].cam.width * views[a].cam.height > (int64_t)views[b].cam.width * views[b].cam.height;
    });
    sizes_.assign(B, {0, 0});
    slot_of_view_.assign(B, 0);
    for (int k = 0; k < B; k++) {
        sizes_[k] = {views[order[k]].cam.width, views[order[k]].cam.height};
        slot_of_view_[order[k]] = k;
    }
    bool graphable = graph_on_ && graph_choice_ && !sequential_;
    if ((int)hwv_.size() < B) hwv_.resize(B, msk::Geom{});
    if (graphable) {
        for (int k = 0; k < B; k++) {
            hwv_[k] = msk::geom_of(std::max(hwv_[k].W, sizes_[k].first), std::max(hwv_[k].H, sizes_[k].second));
            if (!msk::lists_supported(st.N, hwv_[k].T)) graphable = false;
        }
        hw_ = hwv_[0];
    }
    // per-slot buffers: the slot's recorded geometry (graph) or its view's own
    for (int k = 0; k < B; k++) {
        const msk::Geom g = msk::geom_of(sizes_[k].first, sizes_[k].second);
        slots_[k]->prepare(st.N, graphable ? hwv_[k] : g, g, true);
    }
    int Tmax = 0;
    for (int k = 0; k < B; k++)
        Tmax = std::max(Tmax, graphable ? hwv_[k].T : msk::geom_of(sizes_[k].first, sizes_[k].second).T);
    const size_t items = (size_t)B * Tmax;
    wq_.ensure(items * 4);
    wmisc_.ensure(64);
    const size_t off_vc = align256(sizeof(MsplatStepParams));
    const size_t off_rv = off_vc + align256(sizeof(msk::ViewConsts) * B);
    const size_t bytes = off_rv + sizeof(msk::RasterView) * B;
    block_.ensure(bytes);
    // host image of the step block
    std::vector<uint8_t> h(bytes, 0);
    std::memcpy(h.data(), &hp, sizeof(hp));
    for (int k = 0; k < B; k++) {
        const BatchView& bv = views[order[k]];
        const msk::ViewConsts vc = Engine::make_view(bv.cam, bv.gt, hp.background, hp.ssim_weight, bv.inv_n, bv.mask);
        const msk::ViewConsts* d = reinterpret_cast<const msk::ViewConsts*>(block_.as<uint8_t>() + off_vc) + k;
        slots_[k]->bind_view(vc, d);
        std::memcpy(h.data() + off_vc + sizeof(vc) * k, &vc, sizeof(vc));
        const msk::RasterView rv = slots_[k]->raster_view(true);
        std::memcpy(h.data() + off_rv + sizeof(rv) * k, &rv, sizeof(rv));
    }
    ring_.upload(s_, block_.p, h.data(), bytes);
    if (sequential_) {
        const size_t n = (size_t)(P.N > 0 ? P.N : 1);
        const int nb3 = P.degree <= 0 ? 0 : (P.degree == 1 ? 9 : P.degree == 2 ? 24 : 45);
        void* zs[7] = {A.v_mean, A.v_scale, A.v_quat, A.v_opac, A.v_dc, A.v_rest, A.vis};
        const size_t zb[7] = {n * 12, n * 12, n * 16, n * 4, n * 12, n * (size_t)std::max(nb3, 1) * 4, n * 4};
        for (int k = 0; k < 7; k++) msk::fill_zero(s_, zs[k], zb[k]);
        for (int k = 0; k < B; k++) {
            Engine& e = *slots_[k];
            const BatchView& bv = views[order[k]];
            e.own_view(Engine::make_view(bv.cam, bv.gt, hp.background, hp.ssim_weight, bv.inv_n, bv.mask));
            e.front(P, du, true);
            e.raster_fwd_own();
            e.loss_to(0);
            e.raster_bwd_own();
            e.accumulate(st, du, A);
        }
        msk::batch_update(s_, P, A, block_.as<MsplatStepParams>(), du, nan_counts);
        hip_check(hipGetLastError(), "batched step");
        return;
    }
    if (!graphable) {
        enqueue(P, B, du, Tmax, A, nan_counts);
        hip_check(hipGetLastError(), "batched step");
        return;
    }
    Key k;
    k.N = st.N;
    k.B = B;
    k.du = du;
    k.degree = st.degree;
    k.bf16 = st.sh_bf16;
    k.grid = persist_grid_;
    k.ppl = ppl_;
    k.chunks = chunks_;
    k.waves = bucket_waves_;
    k.g = hw_;
    k.slot_geom.clear();
    for (int j = 0; j < B; j++) {
        k.slot_geom.push_back(hwv_[j].W);
        k.slot_geom.push_back(hwv_[j].H);
    }
    // per slot: buffer generation and depth-chunk count (chunked views record their own path),
    // compared entry by entry: a recorded graph is replayed only against the very buffers
    // it was recorded with
    k.gens.clear();
    for (int v = 0; v < B; v++) {
        k.gens.push_back(slots_[v]->generation());
        k.gens.push_back(slots_[v]->chunks());
    }
    const DBuf* own[3] = {&block_, &wq_, &wmisc_};
    for (int j = 0; j < 3; j++) k.own[j] = own[j]->p;
    const void* ptrs[8] = {st.means, st.scales, st.quats, st.f_dc, st.f_rest, st.opac, st.f_dc_bf16, st.f_rest_bf16};
    for (int j = 0; j < 8; j++) k.state[j] = ptrs[j];
    for (int j = 0; j < 6; j++) {
        k.mom[j] = st.m[j];
        k.mom[6 + j] = st.v[j];
    }
    const void* accs[7] = {A.v_mean, A.v_scale, A.v_quat, A.v_opac, A.v_dc, A.v_rest, A.vis};
    for (int j = 0; j < 7; j++) k.acc[j] = accs[j];
    k.nanc = nan_counts;
    if (!(exec_ && key_ == k)) {
        release_graph();
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(s_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
            enqueue(P, B, du, Tmax, A, nan_counts);
        } catch (...) {
            (void)hipStreamEndCapture(s_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hip_check(hipStreamEndCapture(s_, &graph), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&exec_, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hip_check(e, "hipGraphInstantiate");
        key_ = k;
        captures_++;
    } else {
        replays_++;
    }
    hip_check(hipGraphLaunch(exec_, s_), "hipGraphLaunch");
}

double BatchRunner::loss(int v) {
    if (v < 0 || v >= (int)slot_of_view_.size()) return 0.0;
    const int k = slot_of_view_[v];
    return slots_[k]->loss_slot(0, sizes_[k].first, sizes_[k].second);
}

}  // namespace msplat_hip

// engine.cpp — the per-step dispatcher. Replaces the reference's Metal dispatcher
// msplat_train_step / forward_pipeline (core/metal/msplat_metal.mm:803-1097, 1118-1740):
// buffers are cached and grown on demand (the FusedTensorCache role, mm:492-625), the whole
// step is enqueued on one HIP stream with no host synchronisation, and accumulators are
// cleared by fill kernels (safe under graph capture) instead of blits.
#include "engine.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <stdexcept>

namespace msplat_hip {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::string msg = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
        throw std::runtime_error(msg);
    }
}

void DBuf::ensure(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (bytes <= cap) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    size_t rounded = (bytes + 4095) & ~(size_t)4095;
    hip_check(hipMalloc(&p, rounded), "hipMalloc");
    cap = rounded;
}

void DBuf::release() {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
}

static const char* kStageNames[ST_COUNT] = {"zero_grads",  "project_sh_fwd", "depth_rank", "tile_count_scan",
                                            "tile_bucket", "raster_fwd", "ssim_loss",
                                            "raster_bwd",  "sh_rest_adam",   "gaussian_update"};

static std::mutex g_engines_mu;
static std::set<Engine*>& engines() {
    static std::set<Engine*> s;
    return s;
}

void trim_all_engines() {
    std::lock_guard<std::mutex> lk(g_engines_mu);
    for (Engine* e : engines()) e->trim();
}

void sync_all_engines() {
    std::lock_guard<std::mutex> lk(g_engines_mu);
    for (Engine* e : engines()) (void)e->sync();
}

ConstRing::ConstRing() {}

ConstRing::~ConstRing() {
    for (int k = 0; k < kSlots; k++) {
        if (armed_[k]) (void)hipEventSynchronize(ev_[k]);
        if (ev_[k]) (void)hipEventDestroy(ev_[k]);
    }
    if (host_) (void)hipHostFree(host_);
}

void ConstRing::upload(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (bytes > kSlotBytes) throw std::invalid_argument("constant block larger than a staging slot");
    if (!host_) {  // on first use: engines that only run views bound by a batch never upload
        hip_check(hipHostMalloc((void**)&host_, kSlots * kSlotBytes, hipHostMallocDefault), "hipHostMalloc");
        for (int j = 0; j < kSlots; j++)
            hip_check(hipEventCreateWithFlags(&ev_[j], hipEventDisableTiming), "hipEventCreate");
    }
    const int k = next_;
    next_ = (next_ + 1) % kSlots;
    if (armed_[k]) hip_check(hipEventSynchronize(ev_[k]), "hipEventSynchronize");
    uint8_t* h = host_ + (size_t)k * kSlotBytes;
    std::memcpy(h, src, bytes);
    hip_check(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipEventRecord(ev_[k], s), "hipEventRecord");
    armed_[k] = true;
}

int default_persist_grid() {
    static int g = 0;
    if (g == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        g = cus * 8;
    }
    return g;
}

Engine::Engine(hipStream_t shared) {
    if (shared) {
        stream_ = shared;
        owns_stream_ = false;
    } else {
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
    persist_grid_ = default_persist_grid();
    hip_check(hipEventCreate(&mark_[0]), "hipEventCreate");
    hip_check(hipEventCreate(&mark_[1]), "hipEventCreate");
    nanc_.ensure(64);
    hip_check(hipMemset(nanc_.p, 0, 64), "hipMemset");
    totals_.ensure(64);
    loss_.ensure(kLossSlots * 8);
    hip_check(hipMemset(loss_.p, 0, kLossSlots * 8), "hipMemset");
    std::lock_guard<std::mutex> lk(g_engines_mu);
    engines().insert(this);
}

void Engine::trim() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    DBuf* all[] = {&radii_, &tiles_, &depths_, &rect_, &aabb_, &recA_, &recB_, &recC_, &g2d_,
                   &dbg_vm_, &dbg_vs_, &dbg_vq_, &dbg_vo_, &counts_, &exact_,
                   &bins_, &dkeys_, &dkeys_out_, &dvals_, &sorted_ids_, &temp_, &ids_, &l1_, &l1m_, &blocks_, &hist_, &img_,
                   &finalT_, &finalIdx_, &ncontrib_, &vrend_, &dbuf_, &partials_, &chunkT_, &chunkC_, &chunkIdx_,
                   &prefixT_, &afterC_, &wq_, &wmisc_, &consts_, &srect_};
    for (DBuf* b : all) b->release();
    debug_valid_ = false;
    gen_++;
}

Engine::~Engine() {
    {
        std::lock_guard<std::mutex> lk(g_engines_mu);
        engines().erase(this);
    }
    if (stream_) hipStreamSynchronize(stream_);
    for (auto e : ev_pool_) hipEventDestroy(e);
    hipEventDestroy(mark_[0]);
    hipEventDestroy(mark_[1]);
    DBuf* all[] = {&radii_, &tiles_, &depths_, &rect_, &aabb_, &recA_, &recB_, &recC_, &g2d_,
                   &dbg_vm_, &dbg_vs_, &dbg_vq_, &dbg_vo_, &counts_, &exact_,
                   &bins_, &totals_, &dkeys_, &dkeys_out_, &dvals_, &sorted_ids_, &temp_, &ids_, &l1_, &l1m_, &blocks_, &hist_, &img_, &finalT_, &finalIdx_, &ncontrib_,
                   &vrend_, &dbuf_, &partials_, &loss_, &chunkT_, &chunkC_, &chunkIdx_, &prefixT_,
                   &afterC_, &nanc_, &wq_, &wmisc_, &consts_, &srect_};
    for (DBuf* b : all) b->release();
    if (stream_ && owns_stream_) hipStreamDestroy(stream_);
}

// K_max rule of the depth-chunked raster path (reference msplat_metal.mm:1204-1224).
uint32_t k_max_rule(int64_t capacity, int num_tiles) {
    if (num_tiles >= 400) return 1;
    uint32_t avg = (uint32_t)(capacity / (num_tiles > 1 ? num_tiles : 1));
    uint32_t cons = avg * 6u;
    uint32_t K = (cons + msk::CHUNK - 1) / msk::CHUNK;
    if (K < 2) K = 2;
    uint32_t abs_max = (uint32_t)((capacity + msk::CHUNK - 1) / msk::CHUNK);
    if (K > abs_max) K = abs_max;
    return K;
}

void Engine::ensure(DBuf& b, size_t bytes) {
    const size_t cap = b.cap;
    b.ensure(bytes);
    if (b.cap != cap) gen_++;
}

void Engine::prepare(int N, const msk::Geom& launch, const msk::Geom& view, bool train) {
    N_ = N;
    launch_ = launch;
    view_ = view;
    W_ = view.W;
    H_ = view.H;
    tbx_ = view.tbx;
    tby_ = view.tby;
    const int T = launch.T;
    const size_t n = (size_t)(N > 0 ? N : 1), P = (size_t)launch.W * launch.H;
    capacity_ = (int64_t)N * 48;  // 48 N slots (reference msplat_metal.mm:518, 1142)
    if (capacity_ < 64) capacity_ = 64;
    K_ = k_max_rule(capacity_, view.T);
    ensure(radii_, n * 4); ensure(tiles_, n * 4); ensure(depths_, n * 4); ensure(rect_, n * 8); ensure(aabb_, n * 8);
    ensure(recA_, (n + 1) * 16); ensure(recB_, (n + 1) * 16); ensure(recC_, (n + 1) * 8);  // row N: zero record
    ensure(counts_, (size_t)T * 4); ensure(exact_, (size_t)T * 4); ensure(bins_, (size_t)T * 8);
    ensure(srect_, n * 8);
    ensure(dkeys_, n * 4); ensure(dkeys_out_, n * 4); ensure(dvals_, n * 4); ensure(sorted_ids_, n * 4);
    ensure(temp_, msk::lists_temp_bytes(N));
    ensure(ids_, (size_t)capacity_ * 4);
    ensure(l1_, (size_t)capacity_ * 4);
    ensure(l1m_, (size_t)capacity_ * 2);
    ensure(blocks_, msk::lists_block_bytes(launch));
    ensure(hist_, msk::lists_hist_bytes(N, T, chunk_target_));
    ensure(img_, P * 12); ensure(finalT_, P * 4); ensure(finalIdx_, P * 4); ensure(ncontrib_, P * 4);
    ensure(wq_, (size_t)T * 4);
    ensure(wmisc_, 64);
    ensure(consts_, sizeof(msk::ViewConsts) + sizeof(msk::RasterView) + 64);
    if (K_ > 1) {
        const size_t Pv = (size_t)view.W * view.H;
        ensure(chunkT_, (size_t)K_ * Pv * 4); ensure(chunkC_, (size_t)K_ * Pv * 12); ensure(chunkIdx_, (size_t)K_ * Pv * 4);
        if (train) { ensure(prefixT_, (size_t)K_ * Pv * 4); ensure(afterC_, (size_t)K_ * Pv * 12); }
    }
    if (train) {
        ensure(g2d_, n * msk::G2D * 4);
        ensure(vrend_, P * 12); ensure(dbuf_, P * 36);
        const size_t nblk = (size_t)((launch.W + 15) / 16) * ((launch.H + 15) / 16);
        ensure(partials_, nblk * 4);
    }
}

msk::ViewConsts Engine::make_view(const MsplatCamera& cam, const uint8_t* gt, const float bg[3], float ssim_w,
                                  float inv_n, const uint8_t* mask) {
    msk::ViewConsts v{};
    v.cam = cam;
    const msk::Geom g = msk::geom_of(cam.width, cam.height);
    v.W = g.W;
    v.H = g.H;
    v.tbx = g.tbx;
    v.tby = g.tby;
    v.T = g.T;
    for (int c = 0; c < 3; c++) v.bg[c] = bg[c];
    v.ssim_w = ssim_w;
    v.inv_n = inv_n;
    v.gt = gt;
    v.mask = mask;
    return v;
}

void Engine::bind_view(const msk::ViewConsts& h, const msk::ViewConsts* d) {
    vch_ = h;
    vcd_ = d;
}

msk::RasterView Engine::raster_view(bool train) const {
    msk::RasterView r{};
    r.vc = vcd_;
    r.bins = bins_.as<int2>();
    r.pk = {ids_.as<uint32_t>(), {recA_.as<float4>(), recB_.as<float4>(), recC_.as<float2>()}};
    r.img = img_.as<float>();
    r.final_T = finalT_.as<float>();
    r.final_idx = finalIdx_.as<int32_t>();
    r.v_rend = train ? vrend_.as<float>() : nullptr;
    r.G = train ? msk::GradBufs{g2d_.as<float>()}
                : msk::GradBufs{nullptr};
    r.chunked = K_ > 1 ? 1 : 0;
    r.n_contrib = ncontrib_.as<int32_t>();
    return r;
}

// The engine's own constant block: [ViewConsts | RasterView of this view], one copy.
void Engine::upload_view(const msk::ViewConsts& h, bool train) {
    vch_ = h;
    vcd_ = consts_.as<msk::ViewConsts>();
    struct {
        msk::ViewConsts v;
        msk::RasterView r;
    } blk;
    blk.v = h;
    blk.r = raster_view(train);
    ring_.upload(stream_, consts_.p, &blk, sizeof(blk));
}

msk::RasterWork Engine::own_work() {
    msk::RasterWork w{};
    w.views = reinterpret_cast<const msk::RasterView*>(consts_.as<uint8_t>() + sizeof(msk::ViewConsts));
    w.B = 1;
    w.Tmax = launch_.T;
    w.queue = wq_.as<uint32_t>();
    w.n_work = wmisc_.as<uint32_t>();
    w.ctr = wmisc_.as<uint32_t>() + 1;
    w.grid = persist_grid_;
    w.ppl = ppl_;
    return w;
}

void Engine::stage_event(int stage_end) {
    if (!stage_timing_) return;
    if (ev_used_ >= ev_pool_.size()) {
        if (ev_pool_.size() >= 8192) {
            flush_events();
        } else {
            hipEvent_t ev;
            hip_check(hipEventCreate(&ev), "hipEventCreate");
            ev_pool_.push_back(ev);
            ev_stage_.push_back(-1);
        }
    }
    hip_check(hipEventRecord(ev_pool_[ev_used_], stream_), "hipEventRecord");
    ev_stage_[ev_used_] = stage_end;
    ev_used_++;
}

void Engine::flush_events() {
    if (ev_used_ == 0) return;
    hip_check(hipEventSynchronize(ev_pool_[ev_used_ - 1]), "hipEventSynchronize");
    for (size_t k = 1; k < ev_used_; k++) {
        int st = ev_stage_[k];
        if (st < 0) continue;
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ev_pool_[k - 1], ev_pool_[k]), "hipEventElapsedTime");
        stage_ms_[st] += ms;
        stage_n_[st] += 1;
    }
    ev_used_ = 0;
}

void Engine::front(const msk::Params& P, int degrees_to_use, bool train, bool own_queue) {
    // accumulators of the raster backward (the reference blit-zeroes them, mm:1496-1512):
    // zeroed by the projection, row by row
    const msk::GradBufs zero{g2d_.as<float>()};
    if (train) stage_event(ST_ZERO);
    msk::ProjOut pj{radii_.as<int32_t>(), tiles_.as<int32_t>(), depths_.as<float>(), rect_.as<ushort4>(),
                    aabb_.as<float2>(), {recA_.as<float4>(), recB_.as<float4>(), recC_.as<float2>()}};
    msk::project_forward(stream_, P, vcd_, degrees_to_use, pj, nanc_.as<uint32_t>(), train ? &zero : nullptr);
    // (the projection also writes row N, the transparent record)
    stage_event(ST_PROJECT);
    lists_ = list_bufs();
    msk::ListBufs& L = lists_;
    msk::lists_rank(stream_, P.N, pj, L);
    stage_event(ST_RANK);
    // an unchunked view's own raster work queue is written by the tile-offset scan
    queue_built_ = own_queue && K_ <= 1;
    const msk::RasterWork w = own_work();
    msk::lists_count(stream_, P.N, pj, vcd_, launch_, L, queue_built_ ? &w : nullptr);
    stage_event(ST_COUNTS);
    msk::lists_bucket(stream_, P.N, vcd_, launch_, L);
    stage_event(ST_BUCKET);
}

void Engine::raster_fwd_own() {
    msk::RasterBufs R{img_.as<float>(), finalT_.as<float>(), finalIdx_.as<int32_t>(), chunkT_.as<float>(),
                      chunkC_.as<float>(), chunkIdx_.as<int32_t>(), prefixT_.as<float>(), afterC_.as<float>(), K_};
    const msk::RasterWork w = own_work();
    if (K_ <= 1 && !queue_built_) msk::raster_work_build(stream_, w);
    queue_built_ = false;  // the forward consumes the queue's forward head
    msk::raster_forward(stream_, lists_, vcd_, launch_, R, &w);
    stage_event(ST_RASTER_FWD);
}

void Engine::loss_to(int slot) {
    msk::ssim_loss(stream_, vcd_, launch_, img_.as<float>(), dbuf_.as<float>(), vrend_.as<float>(),
                   partials_.as<float>(), loss_.as<double>() + slot, nanc_.as<uint32_t>());
    stage_event(ST_SSIM);
}

void Engine::raster_bwd_own() {
    msk::RasterBufs R{img_.as<float>(), finalT_.as<float>(), finalIdx_.as<int32_t>(), chunkT_.as<float>(),
                      chunkC_.as<float>(), chunkIdx_.as<int32_t>(), prefixT_.as<float>(), afterC_.as<float>(), K_};
    msk::GradBufs G{g2d_.as<float>()};
    const msk::RasterWork w = own_work();
    msk::raster_backward(stream_, lists_, R, vcd_, launch_, vrend_.as<float>(), G, &w);
    stage_event(ST_RASTER_BWD);
}

msk::ListBufs Engine::list_bufs() {
    msk::ListBufs L;
    L.tile_counts = counts_.as<uint32_t>();
    L.exact_off = exact_.as<uint32_t>();
    L.bins = bins_.as<int2>();
    L.srect = srect_.as<ushort4>();
    L.ids = ids_.as<uint32_t>();
    L.view = {ids_.as<uint32_t>(), {recA_.as<float4>(), recB_.as<float4>(), recC_.as<float2>()}};
    L.hist = hist_.as<uint32_t>();
    L.l1_ids = l1_.as<uint32_t>();
    L.l1_mask = l1m_.as<uint16_t>();
    L.blocks = blocks_.as<uint32_t>();
    L.totals = totals_.as<uint32_t>();
    L.capacity = capacity_;
    L.num_tiles = tbx_ * tby_;
    L.dkeys = dkeys_.as<uint32_t>();
    L.dkeys_out = dkeys_out_.as<uint32_t>();
    L.dvals = dvals_.as<uint32_t>();
    L.sorted_ids = sorted_ids_.as<uint32_t>();
    L.temp = temp_.p;
    L.temp_bytes = temp_.cap;
    L.chunk_target = chunk_target_;
    L.bucket_waves = bucket_waves_;
    return L;
}

msk::Params params_of(const MsplatDeviceState& st) {
    msk::Params P;
    P.means = st.means; P.scales = st.scales; P.quats = st.quats; P.f_dc = st.f_dc; P.f_rest = st.f_rest;
    P.opac = st.opac;
    for (int g = 0; g < 6; g++) { P.m[g] = st.m[g]; P.v[g] = st.v[g]; }
    P.N = st.N;
    P.degree = st.degree;
    P.sh_bf16 = st.sh_bf16;
    P.f_dc_h = st.f_dc_bf16;
    P.f_rest_h = st.f_rest_bf16;
    return P;
}


int Engine::forward_backward(const MsplatDeviceState& st, const MsplatCamera& cam, const uint8_t* gt,
                             const MsplatStepParams& hp, int slot, const uint8_t* mask) {
    if (st.N < 0 || st.degree < 0 || st.degree > 3 || cam.width <= 0 || cam.height <= 0 || !gt) return -1;
    if (slot < 0 || slot >= kLossSlots) return -1;
    if (st.sh_bf16 && (!st.f_dc_bf16 || (st.degree > 0 && !st.f_rest_bf16))) return -1;
    if (cam.width > 65535 * 16 || cam.height > 65535 * 16) return -1;
    const msk::Geom g = msk::geom_of(cam.width, cam.height);
    if (!msk::lists_supported(st.N, g.T)) return -1;
    const msk::Params P = params_of(st);
    prepare(st.N, g, g, true);
    last_W = cam.width;
    last_H = cam.height;
    stage_event(-1);
    upload_view(make_view(cam, gt, hp.background, hp.ssim_weight, hp.loss_inv_n, mask), true);
    front(P, hp.degrees_to_use, true);
    raster_fwd_own();
    loss_to(slot);
    raster_bwd_own();
    hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : -2;
}

void Engine::accumulate(const MsplatDeviceState& st, int degrees_to_use, msk::BatchGrads& acc) {
    const msk::Params P = params_of(st);
    msk::ProjOut pj{radii_.as<int32_t>(), tiles_.as<int32_t>(), depths_.as<float>(), rect_.as<ushort4>(),
                    aabb_.as<float2>(), {recA_.as<float4>(), recB_.as<float4>(), recC_.as<float2>()}};
    msk::GradBufs G{g2d_.as<float>()};
    msk::view_accumulate(stream_, P, vcd_, degrees_to_use, pj, G, acc, nanc_.as<uint32_t>());
    stage_event(ST_UPDATE);
}

double Engine::loss_slot(int k, int W, int H) {
    if (k < 0 || k >= kLossSlots) return 0.0;
    double l = 0.0;
    hipMemcpyAsync(&l, loss_.as<double>() + k, 8, hipMemcpyDeviceToHost, stream_);
    hipStreamSynchronize(stream_);
    return (W > 0 && H > 0) ? l / ((double)W * H) : 0.0;
}

int Engine::train_step(const MsplatDeviceState& st, const MsplatCamera& cam, const uint8_t* gt,
                       const MsplatStepParams& hp, bool debug, const uint8_t* mask) {
    const int rc = forward_backward(st, cam, gt, hp, 0, mask);
    if (rc != 0) return rc;
    const msk::Params P = params_of(st);
    const size_t n = (size_t)(st.N > 0 ? st.N : 1);
    msk::GradBufs G{g2d_.as<float>()};
    msk::sh_rest_adam(stream_, P, cam, radii_.as<int32_t>(), g2d_.as<float>(), hp);
    stage_event(ST_SH_ADAM);
    msk::ProjOut pj{radii_.as<int32_t>(), tiles_.as<int32_t>(), depths_.as<float>(), rect_.as<ushort4>(),
                    aabb_.as<float2>(), {recA_.as<float4>(), recB_.as<float4>(), recC_.as<float2>()}};
    float *dvm = nullptr, *dvs = nullptr, *dvq = nullptr, *dvo = nullptr;
    if (debug) {
        dbg_vm_.ensure(n * 12); dbg_vs_.ensure(n * 12); dbg_vq_.ensure(n * 16); dbg_vo_.ensure(n * 4);
        dvm = dbg_vm_.as<float>(); dvs = dbg_vs_.as<float>(); dvq = dbg_vq_.as<float>(); dvo = dbg_vo_.as<float>();
    }
    debug_valid_ = debug;
    msk::gaussian_update(stream_, P, cam, pj, G, hp, nanc_.as<uint32_t>(), dvm, dvs, dvq, dvo);
    stage_event(ST_UPDATE);
    hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : -2;
}

int Engine::render(const MsplatDeviceState& st, const MsplatCamera& cam, int degrees_to_use, const float bg[3]) {
    if (st.N < 0 || st.degree < 0 || st.degree > 3 || cam.width <= 0 || cam.height <= 0) return -1;
    if (st.sh_bf16 && (!st.f_dc_bf16 || (st.degree > 0 && !st.f_rest_bf16))) return -1;
    if (cam.width > 65535 * 16 || cam.height > 65535 * 16) return -1;
    const msk::Geom g = msk::geom_of(cam.width, cam.height);
    if (!msk::lists_supported(st.N, g.T)) return -1;
    const msk::Params P = params_of(st);
    prepare(st.N, g, g, false);
    last_W = cam.width;
    last_H = cam.height;
    stage_event(-1);
    upload_view(make_view(cam, nullptr, bg, 0.f, 0.f), false);
    front(P, degrees_to_use, false);
    raster_fwd_own();
    hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : -2;
}

int Engine::buffer(const std::string& name, void** p, size_t* bytes) {
    const size_t n = (size_t)N_, P = (size_t)W_ * H_, T = (size_t)tbx_ * tby_;
    struct Ent { const char* nm; DBuf* b; size_t sz; };
    const Ent ents[] = {
        {"radii", &radii_, n * 4}, {"tiles_hit", &tiles_, n * 4}, {"depths", &depths_, n * 4}, {"rect", &rect_, n * 8},
        {"aabb", &aabb_, n * 8}, {"recA", &recA_, n * 16}, {"recB", &recB_, n * 16}, {"recC", &recC_, n * 8},
        {"tile_counts", &counts_, T * 4}, {"tile_bins", &bins_, T * 8},
        {"ids", &ids_, (size_t)capacity_ * 4}, {"img", &img_, P * 12}, {"final_T", &finalT_, P * 4},
        {"final_idx", &finalIdx_, P * 4}, {"n_contrib", &ncontrib_, P * 4}, {"v_rendered", &vrend_, P * 12}, {"grad2d", &g2d_, n * msk::G2D * 4},
        {"dbg_v_mean", &dbg_vm_, n * 12}, {"dbg_v_scale", &dbg_vs_, n * 12}, {"dbg_v_quat", &dbg_vq_, n * 16},
        {"dbg_v_opac", &dbg_vo_, n * 4}, {"nan_counts", &nanc_, 64}, {"totals", &totals_, 32}, {"sorted_ids", &sorted_ids_, n * 4},
    };
    for (const Ent& e : ents) {
        if (name == e.nm) {
            if (!e.b->p) return -1;
            *p = e.b->p;
            *bytes = e.sz;
            return 0;
        }
    }
    return -1;
}

int Engine::sync() { return hipStreamSynchronize(stream_) == hipSuccess ? 0 : -2; }

void Engine::stats(int64_t out[6]) {
    uint32_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (totals_.p) (void)hipMemcpyAsync(t, totals_.p, 32, hipMemcpyDeviceToHost, stream_);
    hipStreamSynchronize(stream_);
    out[0] = t[0];
    out[1] = t[1];
    out[2] = t[2];
    out[3] = t[3];
    out[4] = K_;
    out[5] = (int64_t)tbx_ * tby_;
}

double Engine::loss() {
    double l = 0.0;
    hipMemcpyAsync(&l, loss_.p, 8, hipMemcpyDeviceToHost, stream_);
    hipStreamSynchronize(stream_);
    return (W_ > 0 && H_ > 0) ? l / ((double)W_ * H_) : 0.0;
}

void Engine::mark() {
    hipEventRecord(mark_[mark_n_ & 1], stream_);
    mark_n_++;
}

float Engine::elapsed_ms() {
    if (mark_n_ < 2) return 0.f;
    hipEvent_t a = mark_[(mark_n_ - 2) & 1], b = mark_[(mark_n_ - 1) & 1];
    hipEventSynchronize(b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

void Engine::set_stage_timing(bool on) {
    if (!on) flush_events();
    stage_timing_ = on;
}

int Engine::stage_times(double* ms, int max_stages, int64_t* launches) {
    flush_events();
    int n = max_stages < ST_COUNT ? max_stages : ST_COUNT;
    for (int k = 0; k < n; k++) {
        ms[k] = stage_ms_[k];
        if (launches) launches[k] = stage_n_[k];
    }
    return ST_COUNT;
}

void Engine::reset_stage_times() {
    flush_events();
    for (int k = 0; k < ST_COUNT; k++) { stage_ms_[k] = 0; stage_n_[k] = 0; }
}

}  // namespace msplat_hip

// ---------------------------------------------------------------------------------------
// extern "C" kernel-level ABI (include/msplat_kernels.h)
using msplat_hip::Engine;

extern "C" {

int msplat_dev_count(void) {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}
void* msplat_dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 256) != hipSuccess) return nullptr;
    return p;
}
void msplat_dev_free(void* p) {
    if (p) (void)hipFree(p);
}
int msplat_dev_upload(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -2;
}
int msplat_dev_download(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -2;
}
int msplat_dev_memset(void* dst, int value, size_t bytes) {
    return hipMemset(dst, value, bytes) == hipSuccess ? 0 : -2;
}

int msplat_set_list_bucket_waves(int waves) {
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4) return -1;
    msk::set_bucket_waves(waves);
    return 0;
}

int msplat_list_bucket_waves_for(int tiles, int64_t lds_limit, int request) {
    if (tiles <= 0 || lds_limit <= 0) return 0;
    return msk::bucket_waves(tiles, (size_t)lds_limit, request);
}

int msplat_dev_copy_kernel(void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return -1;
    msk::copy16(nullptr, src, dst, bytes);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

MsplatEngine* msplat_engine_create(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nullptr;
    try {
        MsplatEngine* h = new MsplatEngine;
        h->e = new Engine();
        h->owned = true;
        return h;
    } catch (const std::exception& ex) {
        fprintf(stderr, "msplat: engine creation failed: %s\n", ex.what());
        return nullptr;
    }
}

void msplat_engine_destroy(MsplatEngine* e) {
    if (!e) return;
    if (e->owned) delete e->e;
    delete e;
}

int msplat_engine_train_step(MsplatEngine* e, const MsplatDeviceState* st, const MsplatCamera* cam, const uint8_t* gt,
                             const MsplatStepParams* hp, int debug) {
    return msplat_engine_train_step_masked(e, st, cam, gt, nullptr, hp, debug);
}

int msplat_engine_train_step_masked(MsplatEngine* e, const MsplatDeviceState* st, const MsplatCamera* cam,
                                    const uint8_t* gt, const uint8_t* mask, const MsplatStepParams* hp, int debug) {
    if (!e || !st || !cam || !hp) return -1;
    try {
        return e->e->train_step(*st, *cam, gt, *hp, debug != 0, mask);
    } catch (const std::exception& ex) {
        fprintf(stderr, "msplat: %s\n", ex.what());
        return -2;
    }
}

int msplat_engine_render(MsplatEngine* e, const MsplatDeviceState* st, const MsplatCamera* cam, int degrees_to_use,
                         const float background[3]) {
    if (!e || !st || !cam || !background) return -1;
    try {
        return e->e->render(*st, *cam, degrees_to_use, background);
    } catch (const std::exception& ex) {
        fprintf(stderr, "msplat: %s\n", ex.what());
        return -2;
    }
}

int msplat_engine_buffer(MsplatEngine* e, const char* name, void** dptr, size_t* bytes) {
    if (!e || !name || !dptr || !bytes) return -1;
    return e->e->buffer(name, dptr, bytes);
}

int64_t msplat_engine_read(MsplatEngine* e, const char* name, void* host, size_t bytes) {
    if (!e || !name || !host) return -1;
    void* d = nullptr;
    size_t sz = 0;
    if (e->e->buffer(name, &d, &sz) != 0) return -1;
    size_t n = bytes < sz ? bytes : sz;
    if (hipMemcpyAsync(host, d, n, hipMemcpyDeviceToHost, e->e->stream()) != hipSuccess) return -1;
    if (hipStreamSynchronize(e->e->stream()) != hipSuccess) return -1;
    return (int64_t)n;
}

int msplat_engine_sync(MsplatEngine* e) { return e ? e->e->sync() : -1; }

int msplat_engine_stats(MsplatEngine* e, int64_t out[6]) {
    if (!e || !out) return -1;
    e->e->stats(out);
    return 0;
}

double msplat_engine_loss(MsplatEngine* e) { return e ? e->e->loss() : 0.0; }

void* msplat_engine_stream(MsplatEngine* e) { return e ? (void*)e->e->stream() : nullptr; }

void msplat_engine_mark(MsplatEngine* e) {
    if (e) e->e->mark();
}
float msplat_engine_elapsed_ms(MsplatEngine* e) { return e ? e->e->elapsed_ms() : 0.f; }

void msplat_engine_set_stage_timing(MsplatEngine* e, int enable) {
    if (e) e->e->set_stage_timing(enable != 0);
}
int msplat_engine_stage_times(MsplatEngine* e, double* ms, int max_stages, int64_t* launches) {
    return e ? e->e->stage_times(ms, max_stages, launches) : 0;
}
const char* msplat_engine_stage_name(int k) {
    return (k >= 0 && k < msplat_hip::ST_COUNT) ? msplat_hip::kStageNames[k] : "";
}
void msplat_engine_reset_stage_times(MsplatEngine* e) {
    if (e) e->e->reset_stage_times();
}

}  // extern "C"

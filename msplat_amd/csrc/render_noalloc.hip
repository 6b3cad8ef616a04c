// render_noalloc.hip — msplat_render_noalloc: the render of msplat_engine_render (projection,
// depth rank, tile lists, raster forward) for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops). Every intermediate buffer lives in
// one file-scope device array, `scratch`, of kScratchBytes; inputs and outputs are the
// caller's. The per-view constants are written into the array by a kernel that takes them
// by value, so a captured launch sequence carries them with it.
//
// Calls on one GPU share that GPU's array and the last call's contents stay in it: calls
// that overlap on one GPU (two streams, two threads) race. Everything a call reads is
// written earlier in the same call (counters are cleared by fill kernels, never by
// hipMemsetAsync).
#include <algorithm>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch` (looked up once per device, with
// that device current; the symbol's name is a host-side stand-in, not a device address).
void* scratch_ptr() {
    int dev = 0;
    msplat_hip::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        msplat_hip::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("no device address for the render scratch");
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

struct ConstBlock {
    msk::ViewConsts v;
    msk::RasterView r;
};

__global__ void k_store_consts(ConstBlock b, ConstBlock* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = b;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Offsets of every buffer of one render inside the scratch array.
struct Layout {
    size_t consts, nanc, totals, wmisc;
    size_t radii, tiles, depths, rect, aabb, recA, recB, recC;
    size_t srect, dkeys, dkeys_out, dvals, sorted_ids, temp, hist, ids, l1, l1m, blocks;
    size_t counts, exact, bins, wq;
    size_t chunkT, chunkC, chunkIdx;
    size_t total;
    int64_t capacity;
    uint32_t K;
};

// The same sizes Engine::prepare gives its buffers for a render (no gradient buffers).
Layout layout(int N, const msk::Geom& g) {
    Layout L{};
    const size_t n = (size_t)(N > 0 ? N : 1), P = (size_t)g.W * g.H, T = (size_t)g.T;
    L.capacity = std::max<int64_t>((int64_t)N * 48, 64);  // 48 N list slots (reference msplat_metal.mm:518)
    L.K = msplat_hip::k_max_rule(L.capacity, g.T);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align256(bytes ? bytes : 4);
        return o;
    };
    L.consts = take(sizeof(ConstBlock));
    L.nanc = take(64);
    L.totals = take(64);
    L.wmisc = take(64);
    L.radii = take(n * 4); L.tiles = take(n * 4); L.depths = take(n * 4); L.rect = take(n * 8); L.aabb = take(n * 8);
    L.recA = take((n + 1) * 16); L.recB = take((n + 1) * 16); L.recC = take((n + 1) * 8);  // row N: empty record
    L.srect = take(n * 8);
    L.dkeys = take(n * 4); L.dkeys_out = take(n * 4); L.dvals = take(n * 4); L.sorted_ids = take(n * 4);
    L.temp = take(msk::lists_temp_bytes(N));
    L.hist = take(msk::lists_hist_bytes(N, g.T));
    L.ids = take((size_t)L.capacity * 4);
    L.l1 = take((size_t)L.capacity * 4);
    L.l1m = take((size_t)L.capacity * 2);
    L.blocks = take(msk::lists_block_bytes(g));
    L.counts = take(T * 4); L.exact = take(T * 4); L.bins = take(T * 8); L.wq = take(T * 4);
    if (L.K > 1) {
        L.chunkT = take((size_t)L.K * P * 4);
        L.chunkC = take((size_t)L.K * P * 12);
        L.chunkIdx = take((size_t)L.K * P * 4);
    }
    L.total = off;
    return L;
}

int render(const MsplatDeviceState& st, const MsplatCamera& cam, int degrees_to_use, const float bg[3], float* img,
           float* final_T, int32_t* final_idx, hipStream_t s) {
    if (st.N < 0 || st.degree < 0 || st.degree > 3 || cam.width <= 0 || cam.height <= 0) return -1;
    if (!img || !final_T || !final_idx) return -1;
    if (st.sh_bf16 && (!st.f_dc_bf16 || (st.degree > 0 && !st.f_rest_bf16))) return -1;
    if (cam.width > 65535 * 16 || cam.height > 65535 * 16) return -1;
    const msk::Geom g = msk::geom_of(cam.width, cam.height);
    if (!msk::lists_supported(st.N, g.T)) return -1;
    const Layout L = layout(st.N, g);
    if (L.total > kScratchBytes) {
        std::fprintf(stderr,
                     "msplat_render_noalloc: %d Gaussians at %dx%d need %zu bytes of scratch, more than its %zu "
                     "(use msplat_engine_render)\n",
                     st.N, cam.width, cam.height, L.total, kScratchBytes);
        return -3;
    }
    uint8_t* base = static_cast<uint8_t*>(scratch_ptr());
    auto at = [&](size_t o) { return static_cast<void*>(base + o); };
    const msk::Params P = msplat_hip::params_of(st);
    // per-view constants and the raster view table, stored by a kernel (by-value arguments)
    ConstBlock blk{};
    blk.v = msplat_hip::Engine::make_view(cam, nullptr, bg, 0.f, 0.f);
    ConstBlock* dblk = static_cast<ConstBlock*>(at(L.consts));
    blk.r.vc = &dblk->v;
    blk.r.bins = static_cast<const int2*>(at(L.bins));
    blk.r.pk = {static_cast<const uint32_t*>(at(L.ids)),
                {static_cast<float4*>(at(L.recA)), static_cast<float4*>(at(L.recB)), static_cast<float2*>(at(L.recC))}};
    blk.r.img = img;
    blk.r.final_T = final_T;
    blk.r.final_idx = final_idx;
    blk.r.v_rend = nullptr;
    blk.r.G = msk::GradBufs{nullptr};
    blk.r.chunked = L.K > 1 ? 1 : 0;
    hipLaunchKernelGGL(k_store_consts, dim3(1), dim3(64), 0, s, blk, dblk);
    const msk::ViewConsts* vcd = &dblk->v;
    msk::fill_zero(s, at(L.nanc), 64);
    msk::ProjOut pj{static_cast<int32_t*>(at(L.radii)),  static_cast<int32_t*>(at(L.tiles)),
                    static_cast<float*>(at(L.depths)),   static_cast<ushort4*>(at(L.rect)),
                    static_cast<float2*>(at(L.aabb)),
                    {static_cast<float4*>(at(L.recA)), static_cast<float4*>(at(L.recB)), static_cast<float2*>(at(L.recC))}};
    msk::project_forward(s, P, vcd, degrees_to_use, pj, static_cast<uint32_t*>(at(L.nanc)));
    msk::ListBufs lists;
    lists.tile_counts = static_cast<uint32_t*>(at(L.counts));
    lists.exact_off = static_cast<uint32_t*>(at(L.exact));
    lists.bins = static_cast<int2*>(at(L.bins));
    lists.srect = static_cast<ushort4*>(at(L.srect));
    lists.ids = static_cast<uint32_t*>(at(L.ids));
    lists.view = {lists.ids, pj.rec};
    lists.hist = static_cast<uint32_t*>(at(L.hist));
    lists.l1_ids = static_cast<uint32_t*>(at(L.l1));
    lists.l1_mask = static_cast<uint16_t*>(at(L.l1m));
    lists.blocks = static_cast<uint32_t*>(at(L.blocks));
    lists.totals = static_cast<uint32_t*>(at(L.totals));
    lists.capacity = L.capacity;
    lists.num_tiles = g.T;
    lists.dkeys = static_cast<uint32_t*>(at(L.dkeys));
    lists.dkeys_out = static_cast<uint32_t*>(at(L.dkeys_out));
    lists.dvals = static_cast<uint32_t*>(at(L.dvals));
    lists.sorted_ids = static_cast<uint32_t*>(at(L.sorted_ids));
    lists.temp = at(L.temp);
    lists.temp_bytes = msk::lists_temp_bytes(st.N);
    msk::lists_rank(s, st.N, pj, lists);
    msk::lists_count(s, st.N, pj, vcd, g, lists);
    msk::lists_bucket(s, st.N, vcd, g, lists);
    msk::RasterBufs R{img, final_T, final_idx,
                      L.K > 1 ? static_cast<float*>(at(L.chunkT)) : nullptr,
                      L.K > 1 ? static_cast<float*>(at(L.chunkC)) : nullptr,
                      L.K > 1 ? static_cast<int32_t*>(at(L.chunkIdx)) : nullptr,
                      nullptr, nullptr, L.K};
    msk::RasterWork w{};
    w.views = &dblk->r;
    w.B = 1;
    w.Tmax = g.T;
    w.queue = static_cast<uint32_t*>(at(L.wq));
    w.n_work = static_cast<uint32_t*>(at(L.wmisc));
    w.ctr = static_cast<uint32_t*>(at(L.wmisc)) + 1;
    w.grid = msplat_hip::default_persist_grid();
    w.ppl = 1;
    if (L.K <= 1) msk::raster_work_build(s, w);
    msk::raster_forward(s, lists, vcd, g, R, &w);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace

extern "C" int msplat_render_noalloc(const MsplatDeviceState* st, const MsplatCamera* cam, int degrees_to_use,
                                     const float background[3], float* img, float* final_T, int32_t* final_idx,
                                     void* stream) {
    if (!st || !cam || !background) return -1;
    try {
        return render(*st, *cam, degrees_to_use, background, img, final_T, final_idx, static_cast<hipStream_t>(stream));
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "msplat_render_noalloc: %s\n", ex.what());
        return -2;
    }
}

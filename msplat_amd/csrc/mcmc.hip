// mcmc.hip — GPU-resident MCMC relocation and growth (the reference does this on the CPU
// over unified memory: Model::mcmcAfterTrain core/src/model.cpp:497-565, relocate
// model.cpp:336-416, addNewGaussians model.cpp:419-495, computeRelocation
// model.cpp:311-332).
//
// Sampling follows this build's own definition (libc++'s discrete_distribution stream
// cannot be reproduced): weight_k = floor(sigmoid(o_k) 2^32) as u64, cdf = inclusive
// prefix sum, draw j = splitmix64((seed << 32) | j), target = floor(draw total / 2^64),
// source = first k with cdf_k > target. All integer, so any evaluation order gives the same
// sources. Seeds: 0x9E3779B9 ^ count (relocation), 0xA24BAED4 ^ count (growth), the counter
// advancing only when a draw happens.
#include <algorithm>

#include "mcmc.h"
#include "ms_math.h"

namespace msk {

static inline unsigned cdiv2(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t reloc_weight(float logit) {
    const float p = 1.0f / (1.0f + ms_expf(-logit));
    if (!ms_finite(p) || !(p > 0.0f)) return 0;
    return (uint64_t)((double)p * 4294967296.0);
}

// Learned proposal weight: sigmoid(o) exp(clamp(z, -8, 8)) in 2^-20 units (fits u64 sums
// over 2^31 rows).
__device__ __forceinline__ uint64_t learned_weight(float logit, float z) {
    const float p = 1.0f / (1.0f + ms_expf(-logit));
    if (!ms_finite(p) || !(p > 0.0f) || !ms_finite(z)) return 0;
    const float e = ms_expf(fminf(fmaxf(z, -8.0f), 8.0f));
    return (uint64_t)((double)p * (double)e * 1048576.0);
}

// dead = non-finite parameter, sigmoid(opacity) <= 0.005, or |mean|^2 > r^2 (model.cpp:521-558)
__global__ void k_classify(int N, int nr, const float* __restrict__ means, const float* __restrict__ scales,
                           const float* __restrict__ quats, const float* __restrict__ f_dc,
                           const float* __restrict__ f_rest, const float* __restrict__ opac, float r2,
                           const float* __restrict__ mlp_out, uint32_t* __restrict__ dead,
                           uint32_t* __restrict__ alive, uint64_t* __restrict__ w, uint32_t* __restrict__ counters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    bool fin = ms_finite(opac[i]);
    for (int k = 0; k < 3; k++) fin = fin && ms_finite(means[3 * (size_t)i + k]) && ms_finite(scales[3 * (size_t)i + k]);
    for (int k = 0; k < 4; k++) fin = fin && ms_finite(quats[4 * (size_t)i + k]);
    for (int k = 0; k < 3; k++) fin = fin && ms_finite(f_dc[3 * (size_t)i + k]);
    for (int k = 0; k < nr && fin; k++) fin = ms_finite(f_rest[(size_t)nr * i + k]);
    if (!fin) atomicAdd(&counters[0], 1u);
    const float sig = 1.0f / (1.0f + ms_expf(-opac[i]));
    bool is_dead = !fin || (sig <= 0.005f);
    if (!is_dead && r2 > 0.0f) {
        const float x = means[3 * (size_t)i], y = means[3 * (size_t)i + 1], z = means[3 * (size_t)i + 2];
        if (x * x + y * y + z * z > r2) {
            is_dead = true;
            atomicAdd(&counters[1], 1u);
        }
    }
    dead[i] = is_dead ? 1u : 0u;
    alive[i] = is_dead ? 0u : 1u;
    w[i] = is_dead ? 0ull : (mlp_out ? learned_weight(opac[i], mlp_out[4 * (size_t)i]) : reloc_weight(opac[i]));
}

__global__ void k_compact(int N, const uint32_t* __restrict__ dead, const uint32_t* __restrict__ dpos,
                          const uint32_t* __restrict__ apos, const uint64_t* __restrict__ w,
                          uint32_t* __restrict__ dead_list, uint32_t* __restrict__ alive_list,
                          uint64_t* __restrict__ alive_w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (dead[i]) {
        dead_list[dpos[i]] = (uint32_t)i;
    } else {
        alive_list[apos[i]] = (uint32_t)i;
        alive_w[apos[i]] = w[i];
    }
}

__global__ void k_weights_all(int N, const float* __restrict__ opac, uint64_t* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) w[i] = reloc_weight(opac[i]);
}

__global__ void k_sample(int num, uint32_t seed, int ncand, const uint64_t* __restrict__ cdf,
                         uint32_t* __restrict__ chosen, uint32_t* __restrict__ counts) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= num) return;
    const uint64_t total = cdf[ncand - 1];
    const uint64_t u = splitmix64(((uint64_t)seed << 32) | (uint64_t)(uint32_t)j);
    const uint64_t r = __umul64hi(u, total);
    int lo = 0, hi = ncand;  // first k with cdf[k] > r (exists: cdf[ncand-1] = total > r)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > r) hi = mid; else lo = mid + 1;
    }
    chosen[j] = (uint32_t)lo;
    atomicAdd(&counts[lo], 1u);
}

// Eq. 9 (model.cpp:311-332, 389-401) on every candidate that was drawn at least once;
// the ensemble counts the source itself (N = draws + 1).
__device__ void eq9(float o_old, const float s_old[3], int n, float* o_new_out, float s_new[3]) {
    if (n < 1) n = 1;
    if (n > 51) n = 51;
    const float o_new = (float)(1.0 - pow(1.0 - (double)o_old, 1.0 / (double)n));
    const double on = (double)o_new;
    double denom = 0.0, binom = 1.0;
    for (int k = 1; k <= n; k++) {
        binom = binom * (double)(n - k + 1) / (double)k;
        const double term = binom * pow(on, (double)k) / sqrt((double)k);
        denom += (k % 2 == 1) ? term : -term;
    }
    double coeff = (denom > 1e-12) ? ((double)o_old / denom) : 1.0;
    if (!isfinite(coeff) || coeff <= 0.0) coeff = 1.0;
    for (int i = 0; i < 3; i++) s_new[i] = (float)((double)s_old[i] * coeff);
    *o_new_out = o_new;
}

// Row-wise view of the trainable state: the six parameter groups, the two bf16 SH shadows
// (absent in fp32 mode) and the Adam moments, each with its row size in bytes (all even).
struct Rows {
    void* p[8];
    void* m[6];
    void* v[6];
    int pb[8];
    int mb[6];
};

__device__ __forceinline__ void zero_row(void* base, int rb, int idx) {
    uint16_t* r = reinterpret_cast<uint16_t*>(base) + (size_t)idx * (rb >> 1);
    for (int k = 0; k < (rb >> 1); k++) r[k] = 0;
}
__device__ __forceinline__ void copy_row(void* base, int rb, int dst, int src) {
    uint16_t* b = reinterpret_cast<uint16_t*>(base);
    const size_t n = (size_t)(rb >> 1);
    for (size_t k = 0; k < n; k++) b[(size_t)dst * n + k] = b[(size_t)src * n + k];
}
__device__ __forceinline__ void zero_moments(const Rows& R, int idx) {
    for (int g = 0; g < 6; g++) {
        if (!R.mb[g]) continue;
        zero_row(R.m[g], R.mb[g], idx);
        zero_row(R.v[g], R.mb[g], idx);
    }
}

__global__ void k_eq9_sources(int ncand, const uint32_t* __restrict__ cand_list, const uint32_t* __restrict__ counts,
                              Rows R) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncand || counts[k] == 0) return;
    const int src = cand_list ? (int)cand_list[k] : k;
    float* op = reinterpret_cast<float*>(R.p[5]);
    float* sc = reinterpret_cast<float*>(R.p[1]);
    const float o_old = 1.0f / (1.0f + ms_expf(-op[src]));
    const float s_old[3] = {ms_expf(sc[3 * (size_t)src]), ms_expf(sc[3 * (size_t)src + 1]), ms_expf(sc[3 * (size_t)src + 2])};
    float o_new, s_new[3];
    eq9(o_old, s_old, (int)counts[k] + 1, &o_new, s_new);
    o_new = fminf(fmaxf(o_new, 1e-6f), 1.0f - 1e-6f);
    op[src] = (float)log((double)o_new / (1.0 - (double)o_new));
    for (int j = 0; j < 3; j++) sc[3 * (size_t)src + j] = (float)log((double)fmaxf(s_new[j], 1e-10f));
    zero_moments(R, src);
}

__global__ void k_copy_rows(int num, const uint32_t* __restrict__ dst_list, int dst_base,
                            const uint32_t* __restrict__ cand_list, const uint32_t* __restrict__ chosen, Rows R,
                            const float* __restrict__ mlp_out, uint32_t* __restrict__ pair_src,
                            uint32_t* __restrict__ pair_dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= num) return;
    const int dst = dst_list ? (int)dst_list[j] : dst_base + j;
    const int src = cand_list ? (int)cand_list[chosen[j]] : (int)chosen[j];
    for (int g = 0; g < 8; g++)
        if (R.pb[g]) copy_row(R.p[g], R.pb[g], dst, src);
    zero_moments(R, dst);
    if (mlp_out) {  // learned placement: offset in units of the copy's largest scale
        float* mu = reinterpret_cast<float*>(R.p[0]) + 3 * (size_t)dst;
        const float* ls = reinterpret_cast<const float*>(R.p[1]) + 3 * (size_t)dst;
        const float s = ms_expf(fmaxf(fmaxf(ls[0], ls[1]), ls[2]));
        for (int k = 0; k < 3; k++) {
            const float o = mlp_out[4 * (size_t)src + 1 + k];
            mu[k] = fmaf(s, ms_finite(o) ? fminf(fmaxf(o, -3.0f), 3.0f) : 0.0f, mu[k]);
        }
    }
    if (pair_src) {
        pair_src[j] = (uint32_t)src;
        pair_dst[j] = (uint32_t)dst;
    }
}

// Eq. 9 on explicit (opacity, count) pairs, for the known-answer tests.
__global__ void k_eq9_cases(int n, const float* __restrict__ o_old, const int* __restrict__ cnt, float* __restrict__ o_new,
                            float* __restrict__ s_new) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float one[3] = {1.f, 1.f, 1.f};
    float sn[3];
    eq9(o_old[i], one, cnt[i], &o_new[i], sn);
    s_new[i] = sn[0];
}

void eq9_cases(hipStream_t s, int n, const float* o_old, const int* cnt, float* o_new, float* s_new) {
    if (n > 0) hipLaunchKernelGGL(k_eq9_cases, dim3(cdiv2(n, 64)), dim3(64), 0, s, n, o_old, cnt, o_new, s_new);
}

// ---- device-wide prefix sums (u32 counts, u64 weights) --------------------------------
// Three launches: each workgroup scans a 2048-item tile (8 consecutive items per thread,
// wave prefix by shuffles, waves combined through LDS) and writes the tile's total; one
// workgroup scans the totals; the tiles add their offsets. Integer sums: the result does not
// depend on the order of evaluation.
constexpr int SC_THREADS = 256, SC_ITEMS = 8, SC_TILE = SC_THREADS * SC_ITEMS;

template <class T> __device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive prefix of v over the workgroup (blockDim.x <= 1024) and the workgroup's total.
template <class T> __device__ __forceinline__ T block_excl_scan(T v, T& total) {
    __shared__ T ws[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const T incl = wave_incl_scan(v);
    if (lane == 63) ws[wid] = incl;
    __syncthreads();
    T before = 0, all = 0;
    for (int w = 0; w < nw; w++) {
        const T x = ws[w];
        if (w < wid) before += x;
        all += x;
    }
    __syncthreads();
    total = all;
    return before + incl - v;
}

template <class T, bool INCLUSIVE>
__global__ __launch_bounds__(SC_THREADS) void k_scan_tiles(int n, const T* __restrict__ in, T* __restrict__ out,
                                                           T* __restrict__ sums) {
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    T x[SC_ITEMS], run = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        x[k] = base + k < n ? in[base + k] : (T)0;
        run += x[k];
    }
    T total;
    T pre = block_excl_scan<T>(run, total);
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        if (INCLUSIVE) pre += x[k];
        if (base + k < n) out[base + k] = pre;
        if (!INCLUSIVE) pre += x[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the tile totals in place, one workgroup.
template <class T> __global__ __launch_bounds__(1024) void k_scan_sums(int nb, T* __restrict__ sums) {
    T carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int b = b0 + (int)threadIdx.x;
        const T v = b < nb ? sums[b] : (T)0;
        T total;
        const T pre = block_excl_scan<T>(v, total);
        if (b < nb) sums[b] = carry + pre;
        carry += total;
    }
}

template <class T> __global__ __launch_bounds__(SC_THREADS) void k_scan_add(int n, T* __restrict__ out,
                                                                           const T* __restrict__ sums) {
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    const T off = sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++)
        if (base + k < n) out[base + k] += off;
}

// out = prefix sums of in[0..n) (exclusive or inclusive); `sums` holds ceil(n / 2048) items.
template <class T, bool INCLUSIVE>
static void device_scan(hipStream_t s, int n, const T* in, T* out, T* sums) {
    if (n <= 0) return;
    const int nb = (n + SC_TILE - 1) / SC_TILE;
    hipLaunchKernelGGL((k_scan_tiles<T, INCLUSIVE>), dim3(nb), dim3(SC_THREADS), 0, s, n, in, out, sums);
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(1024), 0, s, nb, sums);
    hipLaunchKernelGGL(k_scan_add<T>, dim3(nb), dim3(SC_THREADS), 0, s, n, out, (const T*)sums);
}

static void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

template <class T> static T* grow(McmcScratch& S, int slot, size_t count) {
    size_t bytes = count * sizeof(T) + 256;
    if (S.cap[slot] < bytes) {
        if (S.buf[slot]) check(hipFree(S.buf[slot]), "hipFree");
        S.buf[slot] = nullptr;
        check(hipMalloc(&S.buf[slot], bytes), "hipMalloc");
        S.cap[slot] = bytes;
    }
    return reinterpret_cast<T*>(S.buf[slot]);
}

void McmcScratch::release() {
    for (int k = 0; k < 16; k++) {
        if (buf[k]) (void)hipFree(buf[k]);
        buf[k] = nullptr;
        cap[k] = 0;
    }
}

static Rows make_rows(const Params& P) {
    Rows R;
    const int nr = (P.degree <= 0 ? 0 : (P.degree == 1 ? 3 : P.degree == 2 ? 8 : 15)) * 3;
    const int lens[6] = {3, 3, 4, 3, nr, 1};
    void* ps[6] = {P.means, P.scales, P.quats, P.f_dc, P.f_rest, P.opac};
    for (int g = 0; g < 6; g++) {
        R.p[g] = ps[g];
        R.pb[g] = lens[g] * 4;
        R.m[g] = P.m[g];
        R.v[g] = P.v[g];
        const bool half = P.sh_bf16 && (g == 3 || g == 4);
        R.mb[g] = lens[g] * (half ? 2 : 4);
    }
    R.p[6] = P.sh_bf16 ? (void*)P.f_dc_h : nullptr;
    R.pb[6] = P.sh_bf16 ? 6 : 0;
    R.p[7] = P.sh_bf16 ? (void*)P.f_rest_h : nullptr;
    R.pb[7] = (P.sh_bf16 && nr) ? nr * 2 : 0;
    return R;
}

// Samples `num` candidates from weights w[0..ncand) and applies Eq. 9 + copies.
static bool draw_and_apply(hipStream_t s, McmcScratch& S, const Rows& R, uint64_t* w, int ncand, int num,
                           const uint32_t* cand_list, const uint32_t* dst_list, int dst_base, uint32_t seed_base,
                           uint32_t& count, const float* mlp_out = nullptr, uint32_t* pair_src = nullptr,
                           uint32_t* pair_dst = nullptr) {
    uint64_t* cdf = grow<uint64_t>(S, 8, (size_t)ncand);
    uint64_t* sums = grow<uint64_t>(S, 9, (size_t)ncand / SC_TILE + 1);
    device_scan<uint64_t, true>(s, ncand, w, cdf, sums);
    uint64_t total = 0;
    check(hipMemcpyAsync(&total, cdf + (ncand - 1), 8, hipMemcpyDeviceToHost, s), "copy");
    check(hipStreamSynchronize(s), "sync");
    if (total == 0) return false;  // no usable sources (model.cpp:366, 448)
    const uint32_t seed = seed_base ^ (++count);
    uint32_t* chosen = grow<uint32_t>(S, 10, (size_t)num);
    uint32_t* cnt = grow<uint32_t>(S, 11, (size_t)ncand);
    fill_zero(s, cnt, (size_t)ncand * 4);
    hipLaunchKernelGGL(k_sample, dim3(cdiv2(num, 256)), dim3(256), 0, s, num, seed, ncand, cdf, chosen, cnt);
    hipLaunchKernelGGL(k_eq9_sources, dim3(cdiv2(ncand, 256)), dim3(256), 0, s, ncand, cand_list, cnt, R);
    hipLaunchKernelGGL(k_copy_rows, dim3(cdiv2(num, 256)), dim3(256), 0, s, num, dst_list, dst_base, cand_list, chosen, R,
                       mlp_out, pair_src, pair_dst);
    check(hipGetLastError(), "launch");
    return true;
}

McmcInfo mcmc_relocate_and_grow(hipStream_t s, const Params& P, int cap_max, float cull_radius, uint32_t& count,
                                McmcScratch& S, ReloHooks* hooks) {
    const float* mlp_out = hooks ? hooks->mlp_out : nullptr;
    if (hooks) hooks->pairs = 0;
    McmcInfo info;
    const int N = P.N;
    info.num_active = N;
    if (N <= 0) return info;
    const Rows R = make_rows(P);
    const int nr = R.pb[4] / 4;
    uint32_t* dead = grow<uint32_t>(S, 0, N);
    uint32_t* alive = grow<uint32_t>(S, 1, N);
    uint64_t* w = grow<uint64_t>(S, 2, N);
    uint32_t* ctr = grow<uint32_t>(S, 3, 4);
    uint32_t* dpos = grow<uint32_t>(S, 4, N);
    uint32_t* apos = grow<uint32_t>(S, 5, N);
    fill_zero(s, ctr, 16);
    const float r2 = cull_radius > 0.0f ? cull_radius * cull_radius : -1.0f;
    hipLaunchKernelGGL(k_classify, dim3(cdiv2(N, 256)), dim3(256), 0, s, N, nr, P.means, P.scales, P.quats, P.f_dc,
                       P.f_rest, P.opac, r2, mlp_out, dead, alive, w, ctr);
    uint32_t* sums = grow<uint32_t>(S, 6, (size_t)N / SC_TILE + 1);
    device_scan<uint32_t, false>(s, N, dead, dpos, sums);
    device_scan<uint32_t, false>(s, N, alive, apos, sums);
    uint32_t* dead_list = grow<uint32_t>(S, 12, N);
    uint32_t* alive_list = grow<uint32_t>(S, 13, N);
    uint64_t* alive_w = grow<uint64_t>(S, 14, N);
    hipLaunchKernelGGL(k_compact, dim3(cdiv2(N, 256)), dim3(256), 0, s, N, dead, dpos, apos, w, dead_list, alive_list,
                       alive_w);
    uint32_t h[4] = {0, 0, 0, 0}, last_pos = 0, last_dead = 0;
    check(hipMemcpyAsync(h, ctr, 16, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(&last_pos, dpos + (N - 1), 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(&last_dead, dead + (N - 1), 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipStreamSynchronize(s), "sync");
    const int D = (int)(last_pos + last_dead), A = N - D;
    info.dead = D;
    info.nonfinite = (int)h[0];
    info.culled = (int)h[1];
    if (D > 0 && A > 0 &&
        draw_and_apply(s, S, R, alive_w, A, D, alive_list, dead_list, 0, 0x9E3779B9u, count, mlp_out,
                       hooks ? hooks->pair_src : nullptr, hooks ? hooks->pair_dst : nullptr) &&
        hooks && hooks->pair_src)
        hooks->pairs = D;
    // growth toward cap_max by ~5% (model.cpp:420-421)
    const int target = std::min(cap_max, (int)(1.05f * (float)N));
    const int num_new = target - N;
    if (num_new > 0) {
        uint64_t* wa = grow<uint64_t>(S, 15, N);
        hipLaunchKernelGGL(k_weights_all, dim3(cdiv2(N, 256)), dim3(256), 0, s, N, P.opac, wa);
        if (draw_and_apply(s, S, R, wa, N, num_new, nullptr, nullptr, N, 0xA24BAED4u, count)) {
            info.grown = num_new;
            info.num_active = N + num_new;
        }
    }
    info.reloc_count_out = count;
    return info;
}

}  // namespace msk

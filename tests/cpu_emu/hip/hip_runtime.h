// CPU emulation of the HIP runtime and device intrinsics used by msplat_amd/csrc
// (TEST INFRASTRUCTURE). It lets the library's kernels and host code be compiled with g++
// and run under AddressSanitizer / UndefinedBehaviorSanitizer on the host, so that every
// global and LDS access of a kernel is bounds-checked on the CPU before it runs on a GPU.
//
// Model: a launch runs its workgroups one after another on the calling thread; every
// work-item of a workgroup is a fiber (ucontext, with AddressSanitizer's fiber-switch
// annotations). __syncthreads is a barrier over the workgroup's live work-items (a
// work-item that returns leaves it, as a wave that ends leaves s_barrier), wave operations
// (__shfl_*, __ballot, DPP row moves, wave barriers) synchronise the 64 work-items of a
// wave through a double-buffered exchange area; a barrier that can never open (divergent
// barrier) aborts with the kernel and workgroup named. Static __shared__ arrays become function statics,
// which is sound because workgroups run one at a time. Streams are synchronous. Device
// allocations are exact-size malloc blocks filled with 0xFF bytes, so reads of unwritten
// memory show up as NaN / 0xFFFFFFFF rather than as zeros.
#pragma once

#include <setjmp.h>
#include <ucontext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __constant__ static
#define __HIP_EMU__ 1

// ---- vector types ------------------------------------------------------------------
struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct uint2 { unsigned x, y; };
struct uint4 { unsigned x, y, z, w; };
struct ushort4 { unsigned short x, y, z, w; };
struct uchar4 { unsigned char x, y, z, w; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline ushort4 make_ushort4(unsigned short a, unsigned short b, unsigned short c, unsigned short d) {
    return {a, b, c, d};
}

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// ---- work-item context: one fiber per work-item, one OS thread per launch ------------
struct EmuIdx { unsigned x = 0, y = 0, z = 0; };

struct EmuFiber {
    ucontext_t ctx;       // first entry only
    jmp_buf jb;           // later switches (no signal-mask system call)
    bool started = false;
    char* stack = nullptr;
    void* fake = nullptr;
    int state = 0;        // 0 runnable, 1 at the workgroup barrier, 2 at its wave barrier, 3 returned
    unsigned wave_ticket = 0;
    EmuIdx tid;
};

struct EmuBlock {
    std::vector<EmuFiber> fibers;
    unsigned nthreads = 0, nwaves = 0, cur = 0;
    std::vector<uint64_t> xch[2];        // double-buffered wave exchange, 64 entries per wave
    std::vector<unsigned> wave_op;       // per lane: wave operations issued in this workgroup
    std::vector<int> and_votes;
    unsigned char* dyn_smem = nullptr;
    ucontext_t sched;
    jmp_buf sched_jb;
    void* sched_fake = nullptr;
    const void* sched_bottom = nullptr;
    size_t sched_size = 0;
    std::function<void()> body;
};

inline EmuIdx threadIdx, blockIdx;
inline dim3 blockDim, gridDim;
inline EmuBlock* emu_blk = nullptr;

extern "C" void __sanitizer_start_switch_fiber(void** fake_stack_save, const void* bottom, size_t size)
    __attribute__((weak));
extern "C" void __sanitizer_finish_switch_fiber(void* fake_stack_save, const void** bottom_old, size_t* size_old)
    __attribute__((weak));

inline void* emu_dyn_smem() { return emu_blk->dyn_smem; }
inline unsigned emu_tid() { return emu_blk->cur; }
inline unsigned emu_lane() { return emu_tid() & 63u; }
inline unsigned emu_wave() { return emu_tid() >> 6; }

constexpr size_t kEmuStack = 512 * 1024;

// fiber-switch annotations for AddressSanitizer (off with EMU_NO_FIBER_ANNOTATIONS=1: faster,
// at the price of possible false stack reports)
inline bool emu_annotate() {
    static const bool on = std::getenv("EMU_NO_FIBER_ANNOTATIONS") == nullptr;
    return on && __sanitizer_start_switch_fiber != nullptr;
}

// switch from the current fiber back to the scheduler
inline void emu_yield(int state) {
    EmuBlock* b = emu_blk;
    EmuFiber& f = b->fibers[b->cur];
    f.state = state;
    if (!_setjmp(f.jb)) {
        if (emu_annotate()) __sanitizer_start_switch_fiber(&f.fake, b->sched_bottom, b->sched_size);
        _longjmp(b->sched_jb, 1);
    }
    if (emu_annotate()) __sanitizer_finish_switch_fiber(f.fake, &b->sched_bottom, &b->sched_size);
    threadIdx = f.tid;
}

inline void __syncthreads() { emu_yield(1); }

inline int __syncthreads_and(int p) {
    emu_blk->and_votes[emu_tid()] = p ? 1 : 0;
    emu_yield(1);
    int r = 1;
    for (unsigned t = 0; t < emu_blk->nthreads; t++)
        if (emu_blk->fibers[t].state != 3 && !emu_blk->and_votes[t]) r = 0;
    emu_yield(1);
    return r;
}

inline void emu_wave_sync() { emu_yield(2); }

inline uint64_t emu_exchange(uint64_t v, unsigned src_lane) {
    const unsigned t = emu_tid();
    std::vector<uint64_t>& x = emu_blk->xch[emu_blk->wave_op[t]++ & 1u];
    const size_t base = (size_t)emu_wave() * 64;
    x[base + emu_lane()] = v;
    emu_wave_sync();
    return x[base + (src_lane & 63u)];
}

inline void __builtin_amdgcn_wave_barrier() { emu_wave_sync(); }
#define __builtin_amdgcn_fence(order, scope) ((void)0)

inline uint64_t __ballot(int p) {
    const unsigned t = emu_tid();
    std::vector<uint64_t>& x = emu_blk->xch[emu_blk->wave_op[t]++ & 1u];
    const size_t base = (size_t)emu_wave() * 64;
    x[base + emu_lane()] = p ? 1u : 0u;
    emu_wave_sync();
    uint64_t m = 0;
    const unsigned live = std::min<unsigned>(64u, emu_blk->nthreads - (unsigned)base);
    for (unsigned l = 0; l < live; l++)
        if (x[base + l] && emu_blk->fibers[base + l].state != 3) m |= (1ull << l);
    return m;
}

template <class T> inline T __shfl_xor(T v, int m, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of a wide type");
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    const unsigned l = emu_lane();
    const unsigned src = (l & ~(unsigned)(width - 1)) | ((l ^ (unsigned)m) & (unsigned)(width - 1));
    const uint64_t r = emu_exchange(b, src);
    T o;
    std::memcpy(&o, &r, sizeof(T));
    return o;
}

template <class T> inline T __shfl_up(T v, unsigned d, int width = 64) {
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    const unsigned l = emu_lane();
    const unsigned lw = l & (unsigned)(width - 1);
    const unsigned src = lw >= d ? l - d : l;
    const uint64_t r = emu_exchange(b, src);
    T o;
    std::memcpy(&o, &r, sizeof(T));
    return o;
}

// DPP moves used by the kernels: row_shr:1..15 (0x111-0x11F), row_bcast:15 (0x142),
// row_bcast:31 (0x143). Lanes whose row is masked off, or whose source lane is invalid
// (bound_ctrl off), keep `old`.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    (void)bank_mask;
    const unsigned l = emu_lane(), row = l >> 4;
    int s_lane = -1;
    if (ctrl >= 0x111 && ctrl <= 0x11F) {
        const unsigned n = (unsigned)(ctrl - 0x110);
        if ((l & 15u) >= n) s_lane = (int)(l - n);
    } else if (ctrl == 0x142) {
        if (row >= 1) s_lane = (int)(row * 16 - 1);
    } else if (ctrl == 0x143) {
        if (row >= 2) s_lane = 31;
    } else {
        std::fprintf(stderr, "emu: DPP control 0x%x not emulated\n", ctrl);
        std::abort();
    }
    const uint64_t v = emu_exchange((uint64_t)(uint32_t)src, s_lane >= 0 ? (unsigned)s_lane : l);
    if (!((row_mask >> row) & 1)) return old;
    if (s_lane < 0) return bound_ctrl ? 0 : old;
    return (int)(uint32_t)v;
}

// ---- atomics and integer intrinsics --------------------------------------------------
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline float atomicAdd(float* p, float v) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    uint32_t cur = __atomic_load_n(q, __ATOMIC_SEQ_CST);
    for (;;) {
        float f;
        std::memcpy(&f, &cur, 4);
        f += v;
        uint32_t nb;
        std::memcpy(&nb, &f, 4);
        if (__atomic_compare_exchange_n(q, &cur, nb, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
            std::memcpy(&f, &cur, 4);
            return f;
        }
    }
}
inline unsigned atomicMax(unsigned* p, unsigned v) {
    unsigned cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
inline int atomicMax(int* p, int v) {
    int cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline uint64_t __umul64hi(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline size_t min(size_t a, size_t b) { return a < b ? a : b; }
inline size_t max(size_t a, size_t b) { return a > b ? a : b; }
inline float min(float a, float b) { return std::fmin(a, b); }
inline float max(float a, float b) { return std::fmax(a, b); }

// ---- runtime -------------------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum { hipStreamNonBlocking = 1 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
typedef struct EmuStream { int unused; }* hipStream_t;
typedef struct EmuEvent { std::chrono::steady_clock::time_point t; }* hipEvent_t;
struct hipDeviceProp_t { size_t sharedMemPerBlock = 160 * 1024; int multiProcessorCount = 256; char name[256] = "emu"; };

inline hipError_t hipMalloc(void** p, size_t bytes) {
    *p = std::malloc(bytes ? bytes : 1);
    if (!*p) return hipErrorOutOfMemory;
    std::memset(*p, 0xFF, bytes);
    return hipSuccess;
}
template <class T> inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc(reinterpret_cast<void**>(p), bytes); }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t = nullptr) {
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { return hipMemset(d, v, n); }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new EmuStream(); return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new EmuEvent(); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulated HIP error"; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) { *p = hipDeviceProp_t(); return hipSuccess; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }

// ---- kernel launch -------------------------------------------------------------------
inline std::atomic<long long>& emu_launch_count() {
    static std::atomic<long long> n{0};
    return n;
}

struct EmuStacks {
    std::vector<char*> v;
    ~EmuStacks() {
        for (char* p : v) std::free(p);
    }
};
inline EmuStacks& emu_stacks() {
    static EmuStacks s;
    return s;
}

inline void emu_fiber_entry(unsigned lo, unsigned hi) {
    const uintptr_t p = ((uintptr_t)hi << 32) | lo;
    EmuBlock* b = reinterpret_cast<EmuBlock*>(p);
    if (emu_annotate()) __sanitizer_finish_switch_fiber(nullptr, &b->sched_bottom, &b->sched_size);
    threadIdx = b->fibers[b->cur].tid;
    b->body();
    EmuFiber& f = b->fibers[b->cur];
    f.state = 3;
    if (emu_annotate()) __sanitizer_start_switch_fiber(nullptr, b->sched_bottom, b->sched_size);
    _longjmp(b->sched_jb, 1);
}

// Runs one workgroup: resumes runnable fibers in order; a barrier opens when every fiber
// still running has reached it (the workgroup barrier for all, a wave barrier for the
// fiber's wave). A round in which nothing can run while fibers remain is a divergent
// barrier: reported and fatal, like the hang it would be on a GPU.
inline void emu_run_block(EmuBlock& b, const char* name) {
    for (unsigned t = 0; t < b.nthreads; t++) {
        EmuFiber& f = b.fibers[t];
        f.state = 0;
        f.started = false;
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = kEmuStack;
        f.ctx.uc_link = nullptr;
        const uintptr_t p = reinterpret_cast<uintptr_t>(&b);
        makecontext(&f.ctx, (void (*)())emu_fiber_entry, 2, (unsigned)(p & 0xFFFFFFFFu), (unsigned)(p >> 32));
    }
    std::fill(b.wave_op.begin(), b.wave_op.end(), 0u);
    for (;;) {
        bool ran = false;
        for (unsigned t = 0; t < b.nthreads; t++) {
            EmuFiber& f = b.fibers[t];
            if (f.state != 0) continue;
            ran = true;
            b.cur = t;
            if (!_setjmp(b.sched_jb)) {
                if (emu_annotate()) __sanitizer_start_switch_fiber(&b.sched_fake, f.stack, kEmuStack);
                if (!f.started) {
                    f.started = true;
                    swapcontext(&b.sched, &f.ctx);
                } else {
                    _longjmp(f.jb, 1);
                }
            }
            if (emu_annotate()) __sanitizer_finish_switch_fiber(b.sched_fake, nullptr, nullptr);
        }
        unsigned live = 0, at_block = 0;
        for (unsigned t = 0; t < b.nthreads; t++) {
            if (b.fibers[t].state != 3) live++;
            if (b.fibers[t].state == 1) at_block++;
        }
        if (live == 0) return;
        bool released = false;
        if (at_block == live) {
            for (unsigned t = 0; t < b.nthreads; t++)
                if (b.fibers[t].state == 1) b.fibers[t].state = 0;
            released = true;
        }
        for (unsigned w = 0; w < b.nwaves; w++) {
            const unsigned t0 = w * 64, t1 = std::min(b.nthreads, t0 + 64);
            unsigned lw = 0, at = 0;
            for (unsigned t = t0; t < t1; t++) {
                if (b.fibers[t].state != 3) lw++;
                if (b.fibers[t].state == 2) at++;
            }
            if (lw > 0 && at == lw) {
                for (unsigned t = t0; t < t1; t++) b.fibers[t].state = 0;
                released = true;
            }
        }
        if (!ran && !released) {
            std::fprintf(stderr, "emu: divergent barrier in %s, workgroup (%u,%u,%u): ", name, blockIdx.x, blockIdx.y,
                         blockIdx.z);
            for (unsigned t = 0; t < b.nthreads; t++) std::fprintf(stderr, "%d", b.fibers[t].state);
            std::fprintf(stderr, "\n");
            std::abort();
        }
    }
}

template <class... KArgs, class... Args>
void emu_launch(const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem, hipStream_t,
                Args&&... args) {
    emu_launch_count()++;
    static const bool trace = std::getenv("EMU_TRACE") != nullptr;
    const auto t_start = std::chrono::steady_clock::now();
    const unsigned nthreads = block.x * block.y * block.z;
    if (nthreads == 0 || nthreads > 1024) {
        std::fprintf(stderr, "emu: bad workgroup size %u in %s\n", nthreads, name);
        std::abort();
    }
    const unsigned long long nblocks = (unsigned long long)grid.x * grid.y * grid.z;
    if (nblocks == 0) return;
    std::tuple<std::decay_t<KArgs>...> params(static_cast<std::decay_t<KArgs>>(args)...);
    std::vector<unsigned char> dyn(shmem ? shmem : 1);
    std::vector<char*>& stacks = emu_stacks().v;
    while (stacks.size() < nthreads) stacks.push_back(static_cast<char*>(std::malloc(kEmuStack)));
    EmuBlock b;
    b.nthreads = nthreads;
    b.nwaves = (nthreads + 63) / 64;
    b.fibers.resize(nthreads);
    for (unsigned t = 0; t < nthreads; t++) {
        b.fibers[t].stack = stacks[t];
        b.fibers[t].tid.x = t % block.x;
        b.fibers[t].tid.y = (t / block.x) % block.y;
        b.fibers[t].tid.z = t / (block.x * block.y);
    }
    b.xch[0].assign((size_t)b.nwaves * 64, 0);
    b.xch[1].assign((size_t)b.nwaves * 64, 0);
    b.wave_op.assign(nthreads, 0);
    b.and_votes.assign(nthreads, 1);
    b.dyn_smem = shmem ? dyn.data() : nullptr;
    b.body = [&]() { std::apply(kernel, params); };
    EmuBlock* saved = emu_blk;
    emu_blk = &b;
    blockDim = block;
    gridDim = grid;
    for (unsigned long long i = 0; i < nblocks; i++) {
        blockIdx.x = (unsigned)(i % grid.x);
        blockIdx.y = (unsigned)((i / grid.x) % grid.y);
        blockIdx.z = (unsigned)(i / ((unsigned long long)grid.x * grid.y));
        if (shmem) std::memset(dyn.data(), 0xFF, shmem);
        emu_run_block(b, name);
    }
    emu_blk = saved;
    if (trace)
        std::fprintf(stderr, "emu launch %s grid %ux%ux%u block %u shmem %zu: %.3f s\n", name, grid.x, grid.y, grid.z,
                     nthreads, shmem, std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count());
}

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
    emu_launch(#kernel, kernel, dim3(grid), dim3(block), (size_t)(shmem), stream, ##__VA_ARGS__)

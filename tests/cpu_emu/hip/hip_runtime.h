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
// allocations are guarded blocks in a reserved address range filled with 0xFF bytes, so
// reads of unwritten memory show up as NaN / 0xFFFFFFFF rather than as zeros, and host
// pointers handed to copies or kernels in place of device pointers are reported.
#pragma once

#include <setjmp.h>
#include <sys/mman.h>
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
#include <map>
#include <memory>
#include <mutex>
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
#define __align__(n) alignas(n)
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
inline uint2 make_uint2(unsigned a, unsigned b) { return {a, b}; }
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

// every lane's value, in lane order
inline void emu_gather(uint64_t v, uint64_t* out) {
    const unsigned t = emu_tid();
    std::vector<uint64_t>& x = emu_blk->xch[emu_blk->wave_op[t]++ & 1u];
    const size_t base = (size_t)emu_wave() * 64;
    x[base + emu_lane()] = v;
    emu_wave_sync();
    for (unsigned l = 0; l < 64; l++) out[l] = x[base + l];
}

// v_mfma_f32_16x16x16_bf16, as the CDNA3/CDNA4 instruction set documents its operand
// layout: lane l holds A[l % 16][4 (l / 16) + q], B[4 (l / 16) + q][l % 16] and
// C/D[4 (l / 16) + q][l % 16] for q = 0..3. Products and sums in fp32.
typedef short emu_bf16x4 __attribute__((vector_size(8)));
typedef float emu_f32x4 __attribute__((vector_size(16)));
inline emu_f32x4 __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(emu_bf16x4 a, emu_bf16x4 b, emu_f32x4 c, int, int, int) {
    uint64_t av, bv, A[64], B[64];
    std::memcpy(&av, &a, 8);
    std::memcpy(&bv, &b, 8);
    emu_gather(av, A);
    emu_gather(bv, B);
    auto el = [](uint64_t w, int q) {
        const uint32_t bits = (uint32_t)((w >> (16 * q)) & 0xFFFFu) << 16;
        float f;
        std::memcpy(&f, &bits, 4);
        return f;
    };
    const unsigned l = emu_lane();
    const int col = (int)(l % 16), kb = (int)(l / 16);
    emu_f32x4 d;
    for (int q = 0; q < 4; q++) {
        const int row = 4 * kb + q;
        float acc = c[q];
        for (int k = 0; k < 16; k++) acc += el(A[row + 16 * (k / 4)], k % 4) * el(B[col + 16 * (k / 4)], k % 4);
        d[q] = acc;
    }
    return d;
}

inline void __builtin_amdgcn_wave_barrier() { emu_wave_sync(); }
// A scheduling barrier: instructions before it issue before those after it. The lanes of a
// wave issue in lockstep on the GPU; here they are brought to the same point.
inline void __builtin_amdgcn_sched_barrier(int) { emu_wave_sync(); }
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

// ds_bpermute: lane `src`'s value within the lane's width-wide segment (src per lane)
template <class T> inline T __shfl(T v, int src, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of a wide type");
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    const unsigned l = emu_lane();
    const unsigned from = (l & ~(unsigned)(width - 1)) | ((unsigned)src & (unsigned)(width - 1));
    const uint64_t r = emu_exchange(b, from);
    T o;
    std::memcpy(&o, &r, sizeof(T));
    return o;
}

// v_rcp_f32 (an approximation to within 1 ulp; the exact quotient stands in for it)
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }

// v_readlane_b32: lane `src`'s value, for every lane (src must be the same in all lanes)
inline int __builtin_amdgcn_readlane(int v, int src) {
    return (int)(uint32_t)emu_exchange((uint64_t)(uint32_t)v, (unsigned)src & 63u);
}

// v_readfirstlane_b32: the value of the lowest lane that reaches the call
inline int __builtin_amdgcn_readfirstlane(int v) {
    const uint64_t live = __ballot(1);
    return __builtin_amdgcn_readlane(v, live ? __builtin_ctzll(live) : 0);
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
// HIP's scoped atomic builtin (float add only; the scope is ignored on the CPU). The
// address_space(1) qualifier of the device code is an attribute g++ ignores.
#define __HIP_MEMORY_SCOPE_AGENT 3
inline float __hip_atomic_fetch_add(float* p, float v, int, int) { return atomicAdd(p, v); }
inline double atomicAdd(double* p, double v) {
    const double old = *p;
    *p = old + v;
    return old;
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
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned atomicOr(unsigned* p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
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

// ---- stream capture (HIP graphs) ------------------------------------------------------
// A capture records launches (and copies) as closures holding their arguments by value; a
// graph launch replays them in order. Operations a capture must not contain here abort:
// memset nodes, synchronisation, allocation.
struct EmuGraph { std::vector<std::function<void()>> nodes; };
typedef EmuGraph* hipGraph_t;
typedef EmuGraph* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
inline EmuGraph*& emu_capture() {
    static EmuGraph* g = nullptr;
    return g;
}
inline void emu_no_capture(const char* what) {
    if (emu_capture()) {
        std::fprintf(stderr, "emu: %s during stream capture\n", what);
        std::abort();
    }
}

// ---- device memory -------------------------------------------------------------------
// Device allocations live in their own reserved address range, each right-aligned (256 B)
// against a PROT_NONE guard page and preceded by one, its tail past the requested size
// poisoned for AddressSanitizer; freed ranges become PROT_NONE and are never reused. Copies
// and fills check that their device side lies inside one live allocation and their host
// side outside every one, and every launch scans its arguments for host addresses: the
// pointer-kind mistakes a CPU run would otherwise accept and a GPU turns into a fault.
extern "C" void __asan_poison_memory_region(void const volatile* addr, size_t size) __attribute__((weak));

struct EmuDevMem {
    static constexpr uintptr_t kBase = 0x300000000000ull;
    static constexpr size_t kSpan = (size_t)1 << 40;
    static constexpr size_t kPage = 4096;
    uintptr_t base = 0, next = 0;
    std::map<uintptr_t, size_t> live;  // start -> requested bytes
    std::mutex mu;
    EmuDevMem() {
        void* r = mmap((void*)kBase, kSpan, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE | MAP_FIXED_NOREPLACE,
                       -1, 0);
        if (r == MAP_FAILED || (uintptr_t)r != kBase) {
            std::fprintf(stderr, "emu: cannot reserve the device address range\n");
            std::abort();
        }
        base = next = kBase + kPage;
    }
    bool inside(uintptr_t a) const { return a >= kBase && a < kBase + kSpan; }
    // the live allocation holding [a, a + n), or 0
    uintptr_t find(uintptr_t a, size_t n) {
        std::lock_guard<std::mutex> g(mu);
        auto it = live.upper_bound(a);
        if (it == live.begin()) return 0;
        --it;
        if (a < it->first || a + n > it->first + (it->second ? it->second : 1)) return 0;
        return it->first;
    }
};
inline EmuDevMem& emu_dev() {
    static EmuDevMem* m = new EmuDevMem();
    return *m;
}

[[noreturn]] inline void emu_fault(const char* what, const void* p, size_t n) {
    std::fprintf(stderr, "emu: GPU memory fault: %s (address %p, %zu bytes)\n", what, p, n);
    std::abort();
}

inline void emu_check_dev(const void* p, size_t n, const char* what) {
    if (n == 0) return;
    if (!emu_dev().find((uintptr_t)p, n)) emu_fault(what, p, n);
}
inline void emu_check_host(const void* p, size_t n, const char* what) {
    if (n == 0) return;
    if (emu_dev().inside((uintptr_t)p)) emu_fault(what, p, n);
}

inline hipError_t hipMalloc(void** p, size_t bytes) {
    emu_no_capture("hipMalloc");
    EmuDevMem& m = emu_dev();
    const size_t n = bytes ? bytes : 1;
    const size_t body = (n + 255) & ~(size_t)255;
    const size_t pages = (body + EmuDevMem::kPage - 1) / EmuDevMem::kPage;
    uintptr_t start;
    {
        std::lock_guard<std::mutex> g(m.mu);
        start = m.next;
        if (start + (pages + 1) * EmuDevMem::kPage > EmuDevMem::kBase + EmuDevMem::kSpan) return hipErrorOutOfMemory;
        m.next = start + (pages + 1) * EmuDevMem::kPage;  // + the guard page after it
    }
    if (mprotect((void*)start, pages * EmuDevMem::kPage, PROT_READ | PROT_WRITE) != 0) return hipErrorOutOfMemory;
    const uintptr_t end = start + pages * EmuDevMem::kPage;
    const uintptr_t q = end - body;
    // Fresh device memory holds what the last user left there: fill it with all ones by
    // default, or (MSPLAT_EMU_FILL) zeros or pseudo-random bytes, so that a result that depends
    // on memory read before it is written differs between fills.
    static const int fill_mode = [] {
        const char* f = std::getenv("MSPLAT_EMU_FILL");
        return !f ? 0 : std::strcmp(f, "zero") == 0 ? 1 : std::strcmp(f, "random") == 0 ? 2 : 0;
    }();
    if (fill_mode == 2) {
        static std::atomic<uint64_t> seed{0x9E3779B97F4A7C15ull};
        uint64_t x = seed.fetch_add(0x632BE59BD9B4E019ull);
        for (size_t i = 0; i < body; i += 8) {
            x ^= x << 13; x ^= x >> 7; x ^= x << 17;
            std::memcpy((char*)q + i, &x, std::min<size_t>(8, body - i));
        }
    } else {
        std::memset((void*)q, fill_mode == 1 ? 0x00 : 0xFF, body);
    }
    if (__asan_poison_memory_region) {
        if (q > start) __asan_poison_memory_region((void*)start, q - start);
        if (body > n) __asan_poison_memory_region((void*)(q + n), body - n);
    }
    {
        std::lock_guard<std::mutex> g(m.mu);
        m.live[q] = n;
    }
    *p = (void*)q;
    return hipSuccess;
}
template <class T> inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc(reinterpret_cast<void**>(p), bytes); }

// File-scope device arrays are host globals here. HIP_SYMBOL hands a symbol's host address
// and size to hipGetSymbolAddress, which maps it once to an emulated device allocation of
// the same size (guarded and 0xFF-filled like any other, so code that used the symbol's
// host name as a device address, read past its end, or relied on zero-initialised contents
// would fault or misbehave here as well).
struct EmuSymbol {
    const void* host;
    size_t bytes;
};
#define HIP_SYMBOL(x) (EmuSymbol{(const void*)&(x), sizeof(x)})
inline hipError_t hipGetSymbolAddress(void** p, EmuSymbol sym) {
    static std::mutex mu;
    static std::map<const void*, void*> mapped;
    std::lock_guard<std::mutex> g(mu);
    auto it = mapped.find(sym.host);
    if (it == mapped.end()) {
        void* d = nullptr;
        const hipError_t e = hipMalloc(&d, sym.bytes);
        if (e != hipSuccess) return e;
        it = mapped.emplace(sym.host, d).first;
    }
    *p = it->second;
    return hipSuccess;
}
inline hipError_t hipFree(void* p) {
    emu_no_capture("hipFree");
    if (!p) return hipSuccess;
    EmuDevMem& m = emu_dev();
    size_t n;
    {
        std::lock_guard<std::mutex> g(m.mu);
        auto it = m.live.find((uintptr_t)p);
        if (it == m.live.end()) emu_fault("hipFree of an address that is not a live allocation", p, 0);
        n = it->second;
        m.live.erase(it);
    }
    const uintptr_t lo = (uintptr_t)p & ~(uintptr_t)(EmuDevMem::kPage - 1);
    const uintptr_t hi = ((uintptr_t)p + ((n + 255) & ~(size_t)255) + EmuDevMem::kPage - 1) & ~(uintptr_t)(EmuDevMem::kPage - 1);
    madvise((void*)lo, hi - lo, MADV_DONTNEED);
    mprotect((void*)lo, hi - lo, PROT_NONE);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) {
    emu_no_capture("synchronous hipMemcpy");
    if (n == 0) return hipSuccess;
    switch (k) {
        case hipMemcpyHostToDevice: emu_check_dev(d, n, "host-to-device copy into non-device memory");
                                    emu_check_host(s, n, "host-to-device copy from device memory"); break;
        case hipMemcpyDeviceToHost: emu_check_dev(s, n, "device-to-host copy from non-device memory");
                                    emu_check_host(d, n, "device-to-host copy into device memory"); break;
        case hipMemcpyDeviceToDevice: emu_check_dev(d, n, "device-to-device copy into non-device memory");
                                      emu_check_dev(s, n, "device-to-device copy from non-device memory"); break;
        default: emu_check_host(d, n, "host copy into device memory"); emu_check_host(s, n, "host copy from device memory");
    }
    std::memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t = nullptr) {
    if (EmuGraph* g = emu_capture()) {
        if (k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToHost) {
            std::fprintf(stderr, "emu: host copy recorded into a graph\n");
            std::abort();
        }
        g->nodes.push_back([=]() {
            EmuGraph* saved = emu_capture();
            emu_capture() = nullptr;
            hipMemcpy(d, s, n, k);
            emu_capture() = saved;
        });
        return hipSuccess;
    }
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* d, int v, size_t n) {
    emu_no_capture("hipMemset (a memset node)");
    emu_check_dev(d, n, "memset of non-device memory");
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { return hipMemset(d, v, n); }

// A kernel argument word that is a mapped host address outside the device range. Words are
// read at 8-byte alignment from the argument tuple; ints and floats that happen to look like
// a canonical address are only reported when that address is mapped, which in practice means
// a real host pointer.
inline bool emu_host_address(uint64_t w) {
    if (w < 0x10007fff8000ull || w >= 0x800000000000ull) return false;  // below ASan's high memory
    if (emu_dev().inside((uintptr_t)w)) return false;
    const uintptr_t pg = (uintptr_t)w & ~(uintptr_t)4095;
    return msync((void*)pg, 4096, MS_ASYNC) == 0;
}

inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new EmuStream(); return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { emu_no_capture("hipStreamSynchronize"); return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { emu_no_capture("hipDeviceSynchronize"); return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    emu_no_capture("hipStreamBeginCapture");
    emu_capture() = new EmuGraph();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    *g = emu_capture();
    emu_capture() = nullptr;
    return *g ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) {
    *e = new EmuGraph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    emu_no_capture("hipGraphLaunch");
    for (auto& f : e->nodes) f();
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new EmuEvent(); return hipSuccess; }
enum { hipEventDefault = 0, hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned = 0) {
    *p = std::malloc(bytes ? bytes : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 63 };
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 256; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { emu_no_capture("hipEventSynchronize"); return hipSuccess; }
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

// Pointer arguments are checked directly; struct arguments word by word at 8-byte
// alignment (their padding is copied from initialised objects in this code base).
template <class T> void emu_check_arg(const char* name, int k, const T& a) {
    if constexpr (std::is_pointer_v<T>) {
        if (a && emu_host_address((uint64_t)(uintptr_t)a)) {
            std::fprintf(stderr, "emu: kernel %s argument %d is host address %p\n", name, k, (const void*)a);
            emu_fault("host pointer passed to a kernel", (const void*)a, 8);
        }
    } else if constexpr (std::is_class_v<T> && sizeof(T) >= 8 && alignof(T) >= 8) {
        uint64_t w[sizeof(T) / 8];
        std::memcpy(w, &a, sizeof(w));
        for (size_t i = 0; i < sizeof(T) / 8; i++)
            if (emu_host_address(w[i])) {
                std::fprintf(stderr, "emu: kernel %s argument %d bytes [%zu, %zu) hold host address %#llx\n", name, k,
                             8 * i, 8 * i + 8, (unsigned long long)w[i]);
                emu_fault("host pointer passed to a kernel", (void*)w[i], 8);
            }
    }
}

template <class... KArgs>
void emu_launch_impl(const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem,
                     const std::tuple<std::decay_t<KArgs>...>& params);

template <class... KArgs, class... Args>
void emu_launch(const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem, hipStream_t,
                Args&&... args) {
    std::tuple<std::decay_t<KArgs>...> params(static_cast<std::decay_t<KArgs>>(args)...);
    if (EmuGraph* g = emu_capture()) {
        g->nodes.push_back([=]() { emu_launch_impl(name, kernel, grid, block, shmem, params); });
        return;
    }
    emu_launch_impl(name, kernel, grid, block, shmem, params);
}

template <class... KArgs>
void emu_launch_impl(const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem,
                     const std::tuple<std::decay_t<KArgs>...>& params) {
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
    {
        int k = 0;
        std::apply([&](const auto&... a) { (emu_check_arg(name, k++, a), ...); }, params);
    }
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

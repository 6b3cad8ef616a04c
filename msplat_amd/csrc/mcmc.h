// mcmc.h — GPU-resident MCMC relocation / growth pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdexcept>
#include <string>
#include "kernels.h"

namespace msk {

struct McmcInfo {
    int num_active = 0, dead = 0, grown = 0, nonfinite = 0, culled = 0;
    uint32_t reloc_count_out = 0;
};

struct McmcScratch {
    void* buf[16] = {nullptr};
    size_t cap[16] = {0};
    void release();
};

// Hooks of the learned relocation proposal into a pass (all optional).
struct ReloHooks {
    // [N 4] network outputs for rows [0, N): relocation sources are drawn with weight
    // floor(sigmoid(o) exp(clamp(z0, -8, 8)) 2^20) instead of floor(sigmoid(o) 2^32), and a
    // relocated copy moves by exp(max log-scale) clamp(z[1:4], -3, 3) from its source.
    const float* mlp_out = nullptr;
    // [N] each: source and destination row of every relocated (dead) row, written when set.
    uint32_t* pair_src = nullptr;
    uint32_t* pair_dst = nullptr;
    int pairs = 0;  // out: number of pairs written
};

// One Model::mcmcAfterTrain pass on device-resident state with capacity for cap_max rows:
// relocates dead Gaussians onto opacity-weighted live ones (Eq. 9), then grows the active
// set by ~5% toward cap_max. `count` is the relocation counter (seeds), advanced per draw.
// Synchronises the stream twice (sizes and totals are read back).
McmcInfo mcmc_relocate_and_grow(hipStream_t s, const Params& P, int cap_max, float cull_radius, uint32_t& count,
                                McmcScratch& scratch, ReloHooks* hooks = nullptr);

// Eq. 9 on (opacity, ensemble size) pairs with unit scale (known-answer tests).
void eq9_cases(hipStream_t s, int n, const float* o_old, const int* cnt, float* o_new, float* s_new);

}  // namespace msk

// CPU emulation of msplat_amd/csrc/lane_ops.h (TEST INFRASTRUCTURE): the same swaps over
// the emulated wave (this directory precedes csrc on the emulation build's include path).
#pragma once
#include <hip/hip_runtime.h>

namespace msk {

inline void emu_swap_halves(float& x, float& y, unsigned span) {
    uint64_t X[64], Y[64];
    uint32_t xb, yb;
    std::memcpy(&xb, &x, 4);
    std::memcpy(&yb, &y, 4);
    emu_gather(xb, X);
    emu_gather(yb, Y);
    const unsigned l = emu_lane();
    const bool upper = (l / span) & 1u;  // odd half (span 32) or odd row (span 16)
    // x's odd parts take y's even parts; y's even parts take x's odd parts
    const uint32_t nx = upper ? (uint32_t)Y[l - span] : (uint32_t)X[l];
    const uint32_t ny = upper ? (uint32_t)Y[l] : (uint32_t)X[l + span];
    std::memcpy(&x, &nx, 4);
    std::memcpy(&y, &ny, 4);
}

inline void permlane32_swap(float& x, float& y) { emu_swap_halves(x, y, 32); }
inline void permlane16_swap(float& x, float& y) { emu_swap_halves(x, y, 16); }

}  // namespace msk

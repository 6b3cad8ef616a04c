// Cross-lane swaps of CDNA4 (gfx950) used by the raster backward's reductions.
//
// v_permlane32_swap_b32 x, y: lanes 32-63 of x trade places with lanes 0-31 of y.
// v_permlane16_swap_b32 x, y: the odd 16-lane rows of x trade places with the even rows of y.
// Both registers are read and written, so the instructions are written as inline asm with
// two read-write operands: the compiler's __builtin_amdgcn_permlane{16,32}_swap return the
// first register for both elements of their result pair (the second is lost). The s_nop
// gives the two wait states a VALU write needs before a permlane reads the register.
#pragma once
#include <hip/hip_runtime.h>

namespace msk {

__device__ __forceinline__ void permlane32_swap(float& x, float& y) {
    asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(x), "+v"(y));
}

__device__ __forceinline__ void permlane16_swap(float& x, float& y) {
    asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(x), "+v"(y));
}

}  // namespace msk

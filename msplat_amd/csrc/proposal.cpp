// proposal.cpp — host side of the learned relocation proposal: the kernel-level C ABI of
// the network (msplat_kernels.h). The trainer-side logic lives in Model.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/msplat_kernels.h"
#include "mlp.h"

extern "C" {

size_t msplat_mlp_packed_bytes(void) { return sizeof(msk::MlpPacked); }

int msplat_mlp_pack(const float* theta, void* packed) {
    if (!theta || !packed) return -1;
    msk::mlp_pack(nullptr, theta, reinterpret_cast<msk::MlpPacked*>(packed));
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

int msplat_mlp_forward(int n, const uint16_t* X, const void* packed, float* out) {
    if (n < 0 || (n > 0 && (!X || !packed || !out))) return -1;
    msk::mlp_forward(nullptr, n, X, reinterpret_cast<const msk::MlpPacked*>(packed), out);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

int msplat_mlp_backward(int n, const uint16_t* X, const float* targets, const void* packed, float* grad,
                        double* loss2) {
    if (n < 0 || (n > 0 && (!X || !targets || !packed || !grad || !loss2))) return -1;
    msk::mlp_backward(nullptr, n, X, targets, reinterpret_cast<const msk::MlpPacked*>(packed), grad, loss2);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

}  // extern "C"

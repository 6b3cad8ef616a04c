/* msplat_api_ext.h — additions to the msplat C API that have no counterpart in the
 * reference's msplat_c_api.h. The reference's MsplatConfig struct layout is ABI for its
 * Swift package and stays untouched; every new option goes through a new function here.
 */
#ifndef MSPLAT_API_EXT_H
#define MSPLAT_API_EXT_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include "msplat_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Model::prepareCam as a function (reference core/src/model.cpp:845-889). */
void msplat_camera_prepare(const float c2w[16], float fx, float fy, float cx, float cy, int width, int height,
                           MsplatCamera* out);

/* Seeded synthetic Gaussian set: kind 0 = uniform cube [-1,1]^3, kind 1 = 64-component
 * isotropic mixture. Returns 0, or -1 for bad arguments. */
int msplat_synth_gaussians(int kind, uint64_t seed, int n, int degree, float* means, float* log_scales, float* quats,
                           float* opac_logits, float* f_dc, float* f_rest);

/* Seeded cameras on a sphere looking at the origin (OpenGL camera-to-world, row-major). */
int msplat_synth_cameras(uint64_t seed, int count, float radius, float jitter, float* c2w);

/* Float RGB -> uint8 as the reference quantises ground truth (core/src/input_data.cpp:66-69). */
void msplat_quantize_rgb(const float* src, int64_t n, uint8_t* dst);

#ifdef __cplusplus
}
#endif
#endif

/* msplat_types.h — plain-C types shared by the kernel-level C ABI (msplat_kernels.h)
 * and by the test-side CPU restatement. No torch, no HIP types.
 */
#ifndef MSPLAT_TYPES_H
#define MSPLAT_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One pinhole view, already prepared the way Model::prepareCam does it
 * (reference core/src/model.cpp:845-889): viewmat is the row-major world->camera
 * matrix in the OpenCV convention (camera looks down +z), projmat = P(fov, near 0.001,
 * far 1000) * viewmat, cam_pos the camera centre in world space. */
typedef struct {
    float viewmat[16];
    float projmat[16];
    float fx, fy, cx, cy;
    float cam_pos[3];
    int32_t width, height;
} MsplatCamera;

/* Hyper-parameters of one optimisation step: the scalar arguments the reference
 * passes to msplat_train_step (reference core/metal/bindings.h:71-95) after
 * Model::fullIteration (core/src/model.cpp:913-958) computed them. */
typedef struct {
    float step_size[6];     /* lr_g / (1 - beta1^t), groups means, scales, quats, f_dc, f_rest, opacity */
    float bc2_sqrt[6];      /* sqrt(1 - beta2^t) */
    float beta1, beta2, eps;
    float ssim_weight;
    float loss_inv_n;       /* 1 / (H W 3) */
    float op_reg, sc_reg;   /* 0 before step 500 */
    float reg_invis_scale;  /* 1.0 = reference */
    float background[3];
    /* SGLD (reference core/src/model.cpp:971-974) */
    float noise_lr, xyz_lr, max_disp;
    uint32_t sgld_seed;     /* = training step */
    int32_t sgld_enable;
    int32_t degree;         /* SH degree of the stored coefficients */
    int32_t degrees_to_use; /* min(step / interval, degree) */
} MsplatStepParams;

#ifdef __cplusplus
}
#endif
#endif

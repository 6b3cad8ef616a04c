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
#include "msplat_c_api.h"

typedef struct MsplatEngine MsplatEngine;

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

/* Camera matrices from a vertical field of view (radians) and the render size, as the
 * reference's viewer builds them (core/src/msplat_api.mm:486-528). */
void msplat_camera_from_fov(const float c2w[16], float fov_y, int width, int height, MsplatCamera* out);

/* ---- Datasets ---------------------------------------------------------------------------
 * msplat_dataset_create reads scene folders (COLMAP / Nerfstudio / Polycam, msplat_io.h):
 * every camera's image decoded, resized and undistorted as Camera::loadImage does
 * (downscaleFactor > 1 shrinks by area), uint8 ground truth on the device, every
 * testEvery-th camera held out in eval mode, and the initial Gaussians seeded from the
 * point cloud (kNN scales x 0.1, opacity 0.5, random rotations, DC colour from the points).
 * It also accepts "synthetic:<preset>[:key=value,...]":
 *   presets  c0  (50k uniform-cube Gaussians, 16 cameras 256x256, f = 256)
 *            c1  (1.5M 64-component mixture, 64 jittered cameras 1600x1040, f = 1300)
 *   keys     n, hidden, cams, w, h, f, kind (0 cube / 1 mixture), seed, hidden_seed,
 *            cam_seed, radius, jitter, degree
 * Ground truth is rendered on the GPU from the hidden set at SH degree 3 over the default
 * background and quantised to uint8; the trainable set is the initial model. Cameras sit
 * on a sphere of the given radius (no normalisation; see input_data.cpp:115-169). */
MsplatDataset msplat_dataset_create_empty(void);
/* Adds a camera with host uint8 RGB ground truth [height, width, 3]; test = held out.
 * Returns its index within the train or test list, or -1. */
int msplat_dataset_add_camera(MsplatDataset ds, const float c2w[16], float fx, float fy, float cx, float cy, int width,
                              int height, const uint8_t* rgb, bool test);
/* Training mask of camera `index` of the train (test = false) or held-out list: host uint8
 * [height, width], white = use, black = ignore. Binarised as Camera::loadMask does
 * (reference input_data.cpp:81-111): a value of 128 or more (above one half) keeps the
 * pixel; an all-white mask is dropped. Steps on a masked camera pass the mask to the loss
 * kernels and normalise the loss by the kept pixels (model.cpp:929-934); a mask that keeps
 * no pixel trains as unmasked, as the reference's step does. Returns the count of kept
 * pixels, or -1. */
int64_t msplat_dataset_set_mask(MsplatDataset ds, int index, bool test, const uint8_t* mask);
/* Camera::loadMask for every camera of a scene-folder dataset that has a name relative to
 * its image folder (COLMAP): "<mask_dir>/<name>.png", else <name> with its extension
 * replaced by .png, area-resized to the camera's size, kept above one half (input_data.cpp:
 * 77-111); cameras without a file, or whose mask keeps every pixel, stay unmasked. The
 * reference's CLI does this for --mask-dir (cli/msplat.mm:153-163). Returns the number of
 * cameras masked, or -1. */
int64_t msplat_dataset_load_masks(MsplatDataset ds, const char* mask_dir);
/* Initial Gaussians (row-major host arrays; f_rest [n, (degree+1)^2-1, 3]). Returns 0 / -1. */
int msplat_dataset_set_init_gaussians(MsplatDataset ds, int n, int degree, const float* means, const float* log_scales,
                                      const float* quats, const float* opac_logits, const float* f_dc,
                                      const float* f_rest);

/* Camera `index` of the train (test = false) or held-out list: pose, intrinsics
 * {fx, fy, cx, cy} and size; rgb (optional) receives the uint8 ground truth [h, w, 3].
 * Returns 0 / -1. */
int msplat_dataset_camera_info(MsplatDataset ds, int index, bool test, float c2w[16], float intrinsics[4], int* width,
                               int* height, uint8_t* rgb);

/* ---- Multi-view batched steps (no counterpart in the reference, which trains one view per
 * step) ---------------------------------------------------------------------------------
 * One iteration over the next `views` training cameras of the shuffled order: every view
 * renders the same parameters, the gradients of the mean of the per-view losses are summed
 * over the views, and one Adam update, the regulariser and SGLD follow; the LR schedule and
 * MCMC relocation then advance by one iteration. Views may differ in size. views <= 64. */
MsplatStats msplat_trainer_step_batch(MsplatTrainer t, int views);
/* Cameras (train indices) and per-view losses of the most recent batched step; returns the
 * number of views, or -1. Either output may be NULL. Waits for the device when losses is set. */
int msplat_trainer_last_batch(MsplatTrainer t, int32_t* cameras, double* losses, int max_views);
/* ---- Learned relocation proposal (no counterpart in the reference) -------------------
 * msplat_trainer_set_option(t, "learned_relocation", 1) starts gathering per-Gaussian
 * statistics and training the proposal network on the outcomes of every relocation event
 * (plain ones included); "relocation_mode" 0 lets the held-out-loss judge pick plain or
 * learned per event (learned only while its held-out loss change is no worse), 1 forces
 * plain, 2 forces learned once the network has trained. out = {events, learned events,
 * plain events, training samples, network logistic loss on the newest unseen outcomes,
 * best constant predictor's, network offset loss on them, best constant offset's, training
 * rounds, last plain held-out change, last learned held-out change, last held-out loss}.
 * Returns 0, or -1 when the proposal is off. */
int msplat_trainer_proposal_info(MsplatTrainer t, double out[12]);
/* Reads (set = 0) or replaces (set = 1) the network's 5508 fp32 parameters (layout in
 * msplat_kernels.h). Returns 0 / -1. */
int msplat_trainer_proposal_params(MsplatTrainer t, float* theta, int set);
/* Copies a proposal buffer to host memory: "out" ([N 4] network outputs of the last learned
 * event), "features" ([N 16] bfloat16 of the last event), "stats" ([5 N]: visible views,
 * sum |grad2d|, largest radius, depth sum, depth square sum since the last event). Copies
 * min(size, bytes); returns the buffer's size in bytes, or -1. host may be NULL. */
int64_t msplat_trainer_proposal_read(MsplatTrainer t, const char* name, void* host, int64_t bytes);

/* Recorded-graph counters of the batched steps: out = {graphs recorded, steps replayed from
 * a recorded graph, launch width, launch height}. Batched steps are recorded once per launch
 * configuration and replayed when msplat_trainer_set_option(t, "batch_graph", 1) is set.
 * Returns 0 / -1. */
int msplat_trainer_batch_info(MsplatTrainer t, int64_t out[4]);
/* Launch autotuner of the batched steps (msplat_trainer_set_option(t, "autotune", 1)): the
 * candidates timed so far, five int32 each in cands[5k..5k+4] = {pixels per lane of the
 * persistent raster kernels (the raster sub-tile: 1 of a 256-lane workgroup per tile, or 4 of
 * one wave per tile), their resident workgroups, rank chunks of the tile histogram
 * (0 = default 768), waves per block-list bucket workgroup (0 = default 4), 1 = the step
 * replayed from a recorded graph / 0 = launched eagerly}, with its score in ms[k]: the mean
 * over its timed steps of the measured time / the time a least-squares line through all
 * timed steps (ms = a + b * pixels of the batch) gives for that step's pixels (1 = an
 * average step, lower is faster); once tuning has finished the chosen candidate in
 * chosen[0..4] (zeros before).
 * Stage one times 7 raster choices, stage two 4 list choices around the best, and, when
 * batch_graph is on, stage three the best so far launched eagerly. Returns the number of
 * timed candidates (at most max written), or -1. */
int msplat_trainer_autotune_info(MsplatTrainer t, int32_t* cands, float* ms, int max, int32_t chosen[5]);

/* ---- Trainer introspection (tests, benchmarks) --------------------------------------- */
/* The scalars Model::fullIteration would hand to the kernels for (camera, step, adam step). */
int msplat_trainer_step_params(MsplatTrainer t, int cameraIndex, int step, int adamStep, MsplatStepParams* out);
/* Train-camera index used by the most recent step (-1 before the first). */
int msplat_trainer_last_camera(MsplatTrainer t);
/* Mean loss of the most recent step (waits for the device). */
double msplat_trainer_last_loss(MsplatTrainer t);
/* Copies `count` floats of the active rows of group g (0 means, 1 scales, 2 quats, 3 f_dc,
 * 4 f_rest, 5 opacity); which = 0 parameters, 1 Adam exp_avg, 2 exp_avg_sq. Returns the
 * number of floats available, or -1. */
int64_t msplat_trainer_read_group(MsplatTrainer t, int group, int which, float* dst, int64_t count);
/* Runs Model::mcmcAfterTrain(step) now (its step gate applies). info = {num_active, dead,
 * grown, nonfinite, culled, relocation counter}. Returns 0 / -1. */
int msplat_trainer_run_mcmc(MsplatTrainer t, int step, int32_t info[6]);
/* Trainer options beyond MsplatConfig (whose layout is fixed). Returns 0, or -1 for an
 * unknown name or a failure.
 *   "sh_bf16"  1 = store f_dc / f_rest coefficients and their Adam moments in bfloat16
 *              (fp32 masters, fp32 arithmetic); 0 = all fp32 (default). Converts the
 *              current state; checkpoints stay fp32 on disk.
 *   "nan_inject_step"  the reference's MSPLAT_NAN_INJECT self-test (model.cpp:638-659): at
 *              the end of the relocation pass of that step, 100 draws among the active
 *              rows get NaN DC colours (-1: never; the environment variable
 *              MSPLAT_NAN_INJECT sets it when the trainer is created).
 * The other options (batching, proposal, raster geometry) are listed with the entry points
 * they affect. */
int msplat_trainer_set_option(MsplatTrainer t, const char* name, double value);
/* Adam step counter and current learning rates (lr[6]). */
int msplat_trainer_optim_state(MsplatTrainer t, int* adamSteps, float lr[6]);
/* The trainer's engine (borrowed: msplat_engine_destroy releases only the handle). */
MsplatEngine* msplat_trainer_engine(MsplatTrainer t);

/* Eq. 9 on the GPU for (opacity, ensemble size) pairs with unit scale: o_new[i] and the
 * scale coefficient s_new[i] (known-answer tests). Host arrays. Returns 0 / -1. */
int msplat_eq9_device(int n, const float* o_old, const int* counts, float* o_new, float* s_new);

#ifdef __cplusplus
}
#endif
#endif

/* msplat_c_api.h — the msplat C API, implemented for AMD Instinct MI355X (gfx950).
 *
 * Same entry points, argument meanings and struct layouts as the reference's
 * core/include/msplat_c_api.h:16-168, so Swift / nanobind / ctypes callers link against
 * this library unchanged. Differences a caller can observe:
 *   - C++ exceptions never cross this boundary: creation functions return NULL, integer
 *     functions return -1, metrics come back zeroed, and the reason goes to stderr
 *     (the reference lets exceptions escape, msplat_api.mm:265-330).
 *   - msplat_viewer_render_buffer returns a borrowed HIP device pointer (float RGB
 *     [H, W, 3]) instead of an MTLBuffer; msplat_viewer_command_queue returns the
 *     viewer's hipStream_t instead of an MTLCommandQueue.
 *   - msplat_set_metallib_path is accepted and ignored (kernels are linked in).
 *   - msplat_dataset_create understands "synthetic:<preset>[:key=value,...]" paths in
 *     addition to on-disk datasets (see msplat_api_ext.h).
 */
#ifndef MSPLAT_C_API_H
#define MSPLAT_C_API_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Training options. Field order and types are ABI (reference msplat_c_api.h:16-30). */
typedef struct {
    int iterations;
    int shDegree;
    int shDegreeInterval;
    float ssimWeight;
    int capMax;          /* MCMC: Gaussian budget (rows allocated on the device)   */
    float noiseLr;       /* MCMC: SGLD noise scale                                  */
    float opacityReg;    /* MCMC: opacity regulariser (active from step 500)        */
    float scaleReg;      /* MCMC: scale regulariser (active from step 500)          */
    float cullRadius;    /* MCMC: Gaussians beyond this radius are relocated (0 = off) */
    bool keepCrs;        /* export in the dataset's original coordinate frame       */
    float downscaleFactor;
    float bgColor[3];
} MsplatConfig;

/* Defaults of the reference (msplat_c_api.h:32-47). */
static inline MsplatConfig msplat_default_config(void) {
    MsplatConfig c;
    c.iterations = 30000;
    c.shDegree = 3;
    c.shDegreeInterval = 1000;
    c.ssimWeight = 0.2f;
    c.capMax = 1000000;
    c.noiseLr = 5e5f;
    c.opacityReg = 0.01f;
    c.scaleReg = 0.01f;
    c.cullRadius = 3.0f;
    c.keepCrs = false;
    c.downscaleFactor = 1.0f;
    c.bgColor[0] = 0.6130f;
    c.bgColor[1] = 0.0101f;
    c.bgColor[2] = 0.3984f;
    return c;
}

typedef struct {
    int iteration;
    int splatCount;
    float msPerStep; /* host wall time of the call (enqueue; the device may still be busy) */
} MsplatStats;

typedef struct {
    float psnr;
    float ssim;
    float l1;
    int numTest;
    int numGaussians;
} MsplatEvalMetrics;

/* Float RGB, HWC. `data` comes from malloc(); the caller releases it with free(). */
typedef struct {
    float* data;
    int width;
    int height;
} MsplatPixelBuffer;

/* Datasets: cameras plus uint8 ground truth resident on the GPU. A dataset must outlive
 * every trainer created from it. */
typedef void* MsplatDataset;

MsplatDataset msplat_dataset_create(const char* path, float downscaleFactor, bool evalMode, int testEvery);
void msplat_dataset_destroy(MsplatDataset ds);
int msplat_dataset_num_train(MsplatDataset ds);
int msplat_dataset_num_test(MsplatDataset ds);

/* Trainers: one per process is the reference's model (global caches); several work here,
 * each with its own HIP stream and buffers. Not thread-safe per handle. */
typedef void* MsplatTrainer;

MsplatTrainer msplat_trainer_create(MsplatDataset ds, MsplatConfig config);
void msplat_trainer_destroy(MsplatTrainer t);

MsplatStats msplat_trainer_step(MsplatTrainer t);
void msplat_trainer_train(MsplatTrainer t);
MsplatEvalMetrics msplat_trainer_evaluate(MsplatTrainer t);
MsplatPixelBuffer msplat_trainer_render(MsplatTrainer t, int cameraIndex, bool useTest);
MsplatPixelBuffer msplat_trainer_render_pose(MsplatTrainer t, const float camToWorld[16], int refCameraIndex);

/* RGBA8 render into caller memory of at least width*height*4 bytes; with outRGBA == NULL
 * only the size is reported. */
void msplat_trainer_render_pose_to_buffer(MsplatTrainer t, const float camToWorld[16], int refCameraIndex,
                                          uint8_t* outRGBA, int* outWidth, int* outHeight);
void msplat_trainer_export_ply(MsplatTrainer t, const char* path);
void msplat_trainer_export_splat(MsplatTrainer t, const char* path);
void msplat_trainer_save_checkpoint(MsplatTrainer t, const char* path);
int msplat_trainer_load_checkpoint(MsplatTrainer t, const char* path);
int msplat_trainer_splat_count(MsplatTrainer t);
int msplat_trainer_iteration(MsplatTrainer t);
void msplat_dataset_camera_pose(MsplatDataset ds, int cameraIndex, float camToWorld[16]);

void msplat_set_metallib_path(const char* path);
void msplat_sync(void);
void msplat_cleanup(void);

/* Render-only viewer over a Gaussian PLY (no optimiser state). fovY in radians. */
typedef void* MsplatViewer;

MsplatViewer msplat_viewer_create(const char* plyPath, int width, int height, float fovY);
void msplat_viewer_resize(MsplatViewer v, int width, int height);
void msplat_viewer_set_fov(MsplatViewer v, float fovY);
void msplat_viewer_set_background(MsplatViewer v, float r, float g, float b);
/* camToWorld: 4x4 row-major, OpenGL convention. outRGBA: width*height*4 bytes. */
void msplat_viewer_render(MsplatViewer v, const float camToWorld[16], uint8_t* outRGBA);
int msplat_viewer_splat_count(MsplatViewer v);
void msplat_viewer_get_size(MsplatViewer v, int* outWidth, int* outHeight);
/* Device pointer to the float RGB image, valid until the next render on this viewer.
 * Work is enqueued on the viewer's stream and not waited for. */
void* msplat_viewer_render_buffer(MsplatViewer v, const float camToWorld[16], int* outWidth, int* outHeight);
void msplat_viewer_bounds(MsplatViewer v, float boundsMin[3], float boundsMax[3]);
/* Centre of the densest cell of an 8x8x8 grid over the Gaussian means. */
void msplat_viewer_density_center(MsplatViewer v, float center[3]);
void* msplat_viewer_command_queue(MsplatViewer v);
void msplat_viewer_destroy(MsplatViewer v);

#ifdef __cplusplus
}
#endif

#endif /* MSPLAT_C_API_H */

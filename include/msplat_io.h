/* msplat_io.h — host-side dataset ingestion of the MI355X engine: image decoding, the float
 * image operations, scene folders (COLMAP / Nerfstudio / Polycam) and the seeding of
 * initial Gaussians from a point cloud. No device is touched by any of these functions.
 *
 * They replace the reference's loaders (core/include/loaders.hpp:9-54: imreadRGB,
 * resizeArea, undistortImage, loadColmap, loadNerfstudio, loadPolycam, readPly,
 * readColmapPoints), the dispatcher and camera preparation of core/src/input_data.cpp
 * (inputDataFromX 244-261, Camera::loadImage 22-70, Camera::loadMask 77-111,
 * autoScaleAndCenter 115-169) and the seeding of Model::Model (core/src/model.cpp:96-160 with
 * PointsTensor::scales, core/include/kdtree_tensor.hpp:27-40). msplat_dataset_create
 * (msplat_c_api.h) composes them for a folder path; these entry points expose each step.
 *
 * Every function returns 0 (or a count) on success and -1 on failure, with the reason on
 * stderr.
 */
#ifndef MSPLAT_IO_H
#define MSPLAT_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Decodes a PNG (any colour type and bit depth, Adam7) or baseline JPEG file to 8-bit RGB
 * [height, width, 3] (imreadRGB, image_io.cpp:13-53, without its division by 255). Grey is
 * replicated, alpha composited over black, 16-bit rounded to 8. rgb = NULL: sizes only. */
int msplat_image_decode(const char* path, int* width, int* height, uint8_t* rgb);

/* resizeArea (image_io.cpp:113-162) of float RGB [sh, sw, 3] to [dh, dw, 3]. */
int msplat_image_resize_area(const float* src, int sw, int sh, float* dst, int dw, int dh);

/* undistortImage (image_io.cpp:164-345): float RGB [h, w, 3] with distortion
 * {k1, k2, p1, p2, k3} to the cropped undistorted image in dst (at most [h, w, 3]); its
 * size and intrinsics {fx, fy, cx, cy} are written to out_w, out_h, intr. */
int msplat_image_undistort(const float* src, int w, int h, float fx, float fy, float cx, float cy, const float dist[5],
                           float* dst, int* out_w, int* out_h, float intr[4]);

/* A scene folder: Nerfstudio (transforms.json), COLMAP (cameras.bin, images.bin,
 * points3D.bin or .ply, directly or in sparse/0; images from colmap_images, else images/,
 * else the folder) or Polycam (keyframes/corrected_cameras/ or cameras.json), cameras
 * sorted by file name, positions centred and scaled into [-1, 1] (the points follow). */
typedef struct MsplatSceneImpl* MsplatScene;
MsplatScene msplat_scene_load(const char* path, const char* colmap_images);
void msplat_scene_free(MsplatScene s);
int msplat_scene_num_cameras(MsplatScene s);
/* Camera i: camera-to-world (row-major, OpenGL axes), {fx, fy, cx, cy}, distortion
 * {k1, k2, p1, p2, k3}, size (0 when the metadata has none), image path and the name
 * relative to the image folder (empty for the JSON formats); the strings are truncated to
 * their capacities. */
int msplat_scene_camera(MsplatScene s, int i, float c2w[16], float intr[4], float dist[5], int* width, int* height,
                        char* path, int path_cap, char* relative, int relative_cap);
/* Scale and translation of the centring: position_scene = (position - translation) * scale. */
int msplat_scene_transform(MsplatScene s, float* scale, float translation[3]);
int64_t msplat_scene_num_points(MsplatScene s);
int msplat_scene_points(MsplatScene s, float* xyz, uint8_t* rgb);
/* Camera::loadImage for camera i: decode, adopt the file's size, downscale by area
 * (downscale > 1), undistort, quantise; the camera's size and intrinsics change as the
 * reference's do. The image stays in the scene: msplat_scene_image copies it out. */
int msplat_scene_load_image(MsplatScene s, int i, float downscale);
int msplat_scene_image(MsplatScene s, int i, uint8_t* rgb);
/* Camera::loadMask for camera i from mask_dir into mask [height, width] (255 kept / 0):
 * returns the count of kept pixels, 0 when there is no mask file or the mask keeps every
 * pixel (no mask), -1 on error. */
int64_t msplat_scene_load_mask(MsplatScene s, int i, const char* mask_dir, uint8_t* mask);

/* PointsTensor::scales: mean distance from each of n points to its k - 1 nearest others. */
int msplat_knn_mean_distance(const float* xyz, int64_t n, int k, float* out);
/* The initial Gaussians Model::Model seeds from n points (xyz [n,3], rgb uint8 [n,3]):
 * means, log scales (isotropic, log(init_scale_mult x knn distance)), quaternions
 * (w, x, y, z) from std::mt19937(42), f_dc, zero f_rest [n, (degree+1)^2-1, 3] (may be NULL
 * for degree 0) and opacity logits of init_opacity. The reference uses 0.1 and 0.5. */
int msplat_seed_gaussians(const float* xyz, const uint8_t* rgb, int64_t n, int degree, float init_scale_mult,
                          float init_opacity, float* means, float* log_scales, float* quats, float* opac_logits,
                          float* f_dc, float* f_rest);

#ifdef __cplusplus
}
#endif

#endif

// ingest.h — scene folders on disk to cameras, images and initial Gaussians (reference
// core/src/input_data.cpp, core/src/loaders/load_{colmap,nerfstudio,polycam,ply}.cpp,
// core/include/kdtree_tensor.hpp, Model::Model in core/src/model.cpp:84-160). Host-only.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace msplat_hip {

struct SceneCamera {
    std::string path;          // image file
    std::string relative;      // name relative to the image folder (mask lookup)
    int width = 0, height = 0; // from the metadata (0: from the image)
    float fx = 0, fy = 0, cx = 0, cy = 0;
    float k1 = 0, k2 = 0, k3 = 0, p1 = 0, p2 = 0;
    float c2w[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // row-major, OpenGL axes
    bool distorted() const { return k1 != 0 || k2 != 0 || k3 != 0 || p1 != 0 || p2 != 0; }
};

struct ScenePoints {
    std::vector<float> xyz;    // [n 3]
    std::vector<uint8_t> rgb;  // [n 3]
    int64_t count() const { return (int64_t)(xyz.size() / 3); }
};

struct Scene {
    std::vector<SceneCamera> cams;
    ScenePoints points;
    float scale = 1.0f, translation[3] = {0.f, 0.f, 0.f};
};

// inputDataFromX (input_data.cpp:244-261): Nerfstudio when transforms.json exists, COLMAP
// when cameras.bin (directly or in sparse/0) does, Polycam when keyframes/corrected_cameras
// or cameras.json does. Every loader ends with auto_scale_and_center.
Scene load_scene(const std::string& root, const std::string& colmap_images = "");
Scene load_colmap(const std::string& root, const std::string& colmap_images = "");
Scene load_nerfstudio(const std::string& root);
Scene load_polycam(const std::string& root);
ScenePoints read_ply_points(const std::string& path);
ScenePoints read_colmap_points(const std::string& path);

// autoScaleAndCenter (input_data.cpp:115-169): camera positions minus their mean, then
// divided by the largest absolute coordinate; the points follow.
void auto_scale_and_center(Scene& s);

// Camera::loadImage (input_data.cpp:22-70): decode, adopt the file's size (rescaling the
// intrinsics when the metadata's differs), area-downscale by `downscale` (> 1), undistort
// when the camera has distortion, quantise to uint8. The camera's size and intrinsics
// are updated in place.
std::vector<uint8_t> load_camera_image(SceneCamera& cam, float downscale);

// Camera::loadMask (input_data.cpp:77-111): "<dir>/<relative>.png", else the relative name
// with its extension replaced by .png; resized to the camera's size by area; kept where
// the first channel is above 1/2 (255) else 0. Returns false when there is no file or the
// mask keeps every pixel (dropped); `kept` counts the kept pixels.
bool load_camera_mask(const SceneCamera& cam, const std::string& dir, std::vector<uint8_t>& mask, int64_t& kept);

// PointsTensor::scales (kdtree_tensor.hpp:27-40): the mean distance from each point to its
// k - 1 nearest other points (the k nearest including itself, the first dropped), with
// squared distances summed in float in axis order. Exact, on a uniform grid.
std::vector<float> knn_mean_distance(const float* xyz, int64_t n, int k = 4);

// The initial Gaussians of Model::Model (model.cpp:96-160) from a point cloud, written to
// caller arrays of n rows: means = xyz; log scales = log(init_scale_mult x knn distance),
// isotropic; quaternions drawn from std::mt19937(42) through a uniform float of one 32-bit
// draw (the draw converted to float, divided by 2^32); f_dc = (rgb / 255 - 1/2) / C0;
// f_rest = 0 ([n, (degree+1)^2 - 1, 3]); opacity logit = log(p / (1 - p)).
void seed_gaussians(const ScenePoints& pts, int degree, float init_scale_mult, float init_opacity, float* means,
                    float* log_scales, float* quats, float* opac, float* f_dc, float* f_rest);

}  // namespace msplat_hip

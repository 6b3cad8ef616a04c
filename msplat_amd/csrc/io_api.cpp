// io_api.cpp — the host-side ingestion C ABI (include/msplat_io.h) over image_io / ingest.
// Every entry point catches C++ exceptions and reports failure through its return value.
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/msplat_io.h"
#include "image_io.h"
#include "ingest.h"

using namespace msplat_hip;

struct MsplatSceneImpl {
    Scene scene;
    std::vector<std::vector<uint8_t>> images;  // per camera, after msplat_scene_load_image
};

#define IO_GUARD_BEGIN try {
#define IO_GUARD_END(ret)                                              \
    }                                                                  \
    catch (const std::exception& ex) {                                 \
        std::fprintf(stderr, "msplat: %s: %s\n", __func__, ex.what()); \
        return ret;                                                    \
    }

static void copy_str(const std::string& s, char* out, int cap) {
    if (!out || cap <= 0) return;
    const size_t n = std::min(s.size(), (size_t)cap - 1);
    std::memcpy(out, s.data(), n);
    out[n] = '\0';
}

extern "C" {

int msplat_image_decode(const char* path, int* width, int* height, uint8_t* rgb) {
    IO_GUARD_BEGIN
    if (!path || !width || !height) return -1;
    const std::vector<uint8_t> b = decode_image_file(path, *width, *height);
    if (rgb) std::memcpy(rgb, b.data(), b.size());
    return 0;
    IO_GUARD_END(-1)
}

int msplat_image_resize_area(const float* src, int sw, int sh, float* dst, int dw, int dh) {
    IO_GUARD_BEGIN
    if (!src || !dst || sw <= 0 || sh <= 0 || dw <= 0 || dh <= 0) return -1;
    HostImage in;
    in.width = sw;
    in.height = sh;
    in.rgb.assign(src, src + (size_t)sw * sh * 3);
    const HostImage out = resize_area(in, dw, dh);
    std::memcpy(dst, out.rgb.data(), out.rgb.size() * sizeof(float));
    return 0;
    IO_GUARD_END(-1)
}

int msplat_image_undistort(const float* src, int w, int h, float fx, float fy, float cx, float cy, const float dist[5],
                           float* dst, int* out_w, int* out_h, float intr[4]) {
    IO_GUARD_BEGIN
    if (!src || !dist || !dst || !out_w || !out_h || !intr || w <= 0 || h <= 0) return -1;
    HostImage in;
    in.width = w;
    in.height = h;
    in.rgb.assign(src, src + (size_t)w * h * 3);
    const Undistorted u = undistort_image(in, fx, fy, cx, cy, dist[0], dist[1], dist[2], dist[3], dist[4]);
    std::memcpy(dst, u.image.rgb.data(), u.image.rgb.size() * sizeof(float));
    *out_w = u.image.width;
    *out_h = u.image.height;
    intr[0] = u.fx; intr[1] = u.fy; intr[2] = u.cx; intr[3] = u.cy;
    return 0;
    IO_GUARD_END(-1)
}

MsplatScene msplat_scene_load(const char* path, const char* colmap_images) {
    IO_GUARD_BEGIN
    if (!path) return nullptr;
    std::unique_ptr<MsplatSceneImpl> s(new MsplatSceneImpl());
    s->scene = load_scene(path, colmap_images ? colmap_images : "");
    s->images.resize(s->scene.cams.size());
    return s.release();
    IO_GUARD_END(nullptr)
}

void msplat_scene_free(MsplatScene s) { delete s; }

int msplat_scene_num_cameras(MsplatScene s) { return s ? (int)s->scene.cams.size() : -1; }

int msplat_scene_camera(MsplatScene s, int i, float c2w[16], float intr[4], float dist[5], int* width, int* height,
                        char* path, int path_cap, char* relative, int relative_cap) {
    if (!s || i < 0 || i >= (int)s->scene.cams.size()) return -1;
    const SceneCamera& c = s->scene.cams[(size_t)i];
    if (c2w) std::memcpy(c2w, c.c2w, sizeof(c.c2w));
    if (intr) { intr[0] = c.fx; intr[1] = c.fy; intr[2] = c.cx; intr[3] = c.cy; }
    if (dist) { dist[0] = c.k1; dist[1] = c.k2; dist[2] = c.p1; dist[3] = c.p2; dist[4] = c.k3; }
    if (width) *width = c.width;
    if (height) *height = c.height;
    copy_str(c.path, path, path_cap);
    copy_str(c.relative, relative, relative_cap);
    return 0;
}

int msplat_scene_transform(MsplatScene s, float* scale, float translation[3]) {
    if (!s) return -1;
    if (scale) *scale = s->scene.scale;
    if (translation)
        for (int k = 0; k < 3; k++) translation[k] = s->scene.translation[k];
    return 0;
}

int64_t msplat_scene_num_points(MsplatScene s) { return s ? s->scene.points.count() : -1; }

int msplat_scene_points(MsplatScene s, float* xyz, uint8_t* rgb) {
    if (!s) return -1;
    const ScenePoints& p = s->scene.points;
    if (xyz) std::memcpy(xyz, p.xyz.data(), p.xyz.size() * sizeof(float));
    if (rgb) std::memcpy(rgb, p.rgb.data(), p.rgb.size());
    return 0;
}

int msplat_scene_load_image(MsplatScene s, int i, float downscale) {
    IO_GUARD_BEGIN
    if (!s || i < 0 || i >= (int)s->scene.cams.size()) return -1;
    s->images[(size_t)i] = load_camera_image(s->scene.cams[(size_t)i], downscale);
    return 0;
    IO_GUARD_END(-1)
}

int msplat_scene_image(MsplatScene s, int i, uint8_t* rgb) {
    if (!s || !rgb || i < 0 || i >= (int)s->images.size() || s->images[(size_t)i].empty()) return -1;
    std::memcpy(rgb, s->images[(size_t)i].data(), s->images[(size_t)i].size());
    return 0;
}

int64_t msplat_scene_load_mask(MsplatScene s, int i, const char* mask_dir, uint8_t* mask) {
    IO_GUARD_BEGIN
    if (!s || !mask_dir || i < 0 || i >= (int)s->scene.cams.size()) return -1;
    std::vector<uint8_t> m;
    int64_t kept = 0;
    if (!load_camera_mask(s->scene.cams[(size_t)i], mask_dir, m, kept)) return 0;
    if (mask) std::memcpy(mask, m.data(), m.size());
    return kept;
    IO_GUARD_END(-1)
}

int msplat_knn_mean_distance(const float* xyz, int64_t n, int k, float* out) {
    IO_GUARD_BEGIN
    if (!xyz || !out || n < 0 || k < 2) return -1;
    const std::vector<float> d = knn_mean_distance(xyz, n, k);
    std::memcpy(out, d.data(), d.size() * sizeof(float));
    return 0;
    IO_GUARD_END(-1)
}

int msplat_seed_gaussians(const float* xyz, const uint8_t* rgb, int64_t n, int degree, float init_scale_mult,
                          float init_opacity, float* means, float* log_scales, float* quats, float* opac_logits,
                          float* f_dc, float* f_rest) {
    IO_GUARD_BEGIN
    if (!xyz || !rgb || n < 0 || !means || !log_scales || !quats || !opac_logits || !f_dc) return -1;
    if (degree > 0 && !f_rest) return -1;
    ScenePoints p;
    p.xyz.assign(xyz, xyz + (size_t)n * 3);
    p.rgb.assign(rgb, rgb + (size_t)n * 3);
    seed_gaussians(p, degree, init_scale_mult, init_opacity, means, log_scales, quats, opac_logits, f_dc, f_rest);
    return 0;
    IO_GUARD_END(-1)
}

}  // extern "C"

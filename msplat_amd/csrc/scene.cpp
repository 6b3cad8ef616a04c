// scene.cpp — camera preparation and the seeded synthetic scenes the benchmarks use.
//
// msplat_camera_prepare is Model::prepareCam (reference core/src/model.cpp:845-889): the
// camera-to-world pose (4x4 row-major, OpenGL convention) becomes a world->camera view
// matrix in the OpenCV convention (columns 1 and 2 negated) and a projection matrix with
// near 0.001 / far 1000, all in float as the reference computes them.
//
// The synthetic generators produce the workloads of the benchmark configurations:
// trainable and hidden ground-truth Gaussian sets (uniform cube or a 64-component isotropic
// mixture) and pinhole cameras on a sphere looking at the origin. Randomness comes from a
// counter-based splitmix64 stream with double-precision Box-Muller, so every scene is a
// pure function of (kind, seed, size).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "msplat_types.h"
#include "msplat_api_ext.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
};

}  // namespace

extern "C" {

void msplat_camera_prepare(const float c2w[16], float fx, float fy, float cx, float cy, int width, int height,
                           MsplatCamera* out) {
    const float fovX = 2.0f * std::atan((float)width / (2.0f * fx));
    const float fovY = 2.0f * std::atan((float)height / (2.0f * fy));
    float R[3][3], Rinv[3][3], T[3], Tinv[3];
    for (int i = 0; i < 3; i++) {
        R[i][0] = c2w[i * 4 + 0];
        R[i][1] = -c2w[i * 4 + 1];
        R[i][2] = -c2w[i * 4 + 2];
        T[i] = c2w[i * 4 + 3];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rinv[i][j] = R[j][i];
    for (int i = 0; i < 3; i++) Tinv[i] = -(Rinv[i][0] * T[0] + Rinv[i][1] * T[1] + Rinv[i][2] * T[2]);
    const float vm[16] = {Rinv[0][0], Rinv[0][1], Rinv[0][2], Tinv[0], Rinv[1][0], Rinv[1][1], Rinv[1][2], Tinv[1],
                          Rinv[2][0], Rinv[2][1], Rinv[2][2], Tinv[2], 0, 0, 0, 1};
    const float t_p = 0.001f * std::tan(0.5f * fovY), r_p = 0.001f * std::tan(0.5f * fovX);
    const float pm[16] = {0.001f / r_p, 0, 0, 0, 0, 0.001f / t_p, 0, 0, 0, 0, (1000.0f + 0.001f) / (1000.0f - 0.001f),
                          -1000.0f * 0.001f / (1000.0f - 0.001f), 0, 0, 1, 0};
    float pvm[16] = {};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) pvm[i * 4 + j] += pm[i * 4 + k] * vm[k * 4 + j];
    std::memcpy(out->viewmat, vm, sizeof(vm));
    std::memcpy(out->projmat, pvm, sizeof(pvm));
    out->fx = fx;
    out->fy = fy;
    out->cx = cx;
    out->cy = cy;
    out->cam_pos[0] = T[0];
    out->cam_pos[1] = T[1];
    out->cam_pos[2] = T[2];
    out->width = width;
    out->height = height;
}

// Gaussian set: kind 0 = means uniform in [-1,1]^3; kind 1 = 64-component isotropic mixture
// (centres uniform in [-0.8,0.8]^3, sigma uniform in [0.05, 0.3], equal weights). log-scales
// N(-4.5, 0.5), quaternions normalised N(0,1)^4, opacity logits N(0,1), DC N(0, 0.3),
// higher SH N(0, 0.02). f_rest holds ((degree+1)^2 - 1) * 3 floats per Gaussian.
int msplat_synth_gaussians(int kind, uint64_t seed, int n, int degree, float* means, float* log_scales, float* quats,
                           float* opac_logits, float* f_dc, float* f_rest) {
    if (n < 0 || degree < 0 || degree > 3) return -1;
    Rng rng(seed * 1000003ull + (uint64_t)kind);
    const int nr = ((degree + 1) * (degree + 1) - 1) * 3;
    std::vector<double> cc, cs;
    if (kind == 1) {
        Rng crng(seed * 7919ull + 17ull);
        for (int k = 0; k < 64; k++) {
            for (int d = 0; d < 3; d++) cc.push_back(-0.8 + 1.6 * crng.uniform());
            cs.push_back(0.05 + 0.25 * crng.uniform());
        }
    }
    for (int i = 0; i < n; i++) {
        if (kind == 1) {
            int k = (int)(rng.uniform() * 64.0);
            if (k > 63) k = 63;
            for (int d = 0; d < 3; d++) means[3 * (size_t)i + d] = (float)(cc[3 * k + d] + cs[k] * rng.normal());
        } else {
            for (int d = 0; d < 3; d++) means[3 * (size_t)i + d] = (float)(-1.0 + 2.0 * rng.uniform());
        }
        for (int d = 0; d < 3; d++) log_scales[3 * (size_t)i + d] = (float)(-4.5 + 0.5 * rng.normal());
        double q[4], qn = 0;
        for (int d = 0; d < 4; d++) { q[d] = rng.normal(); qn += q[d] * q[d]; }
        qn = std::sqrt(qn > 1e-30 ? qn : 1e-30);
        for (int d = 0; d < 4; d++) quats[4 * (size_t)i + d] = (float)(q[d] / qn);
        opac_logits[i] = (float)rng.normal();
        for (int d = 0; d < 3; d++) f_dc[3 * (size_t)i + d] = (float)(0.3 * rng.normal());
        for (int d = 0; d < nr; d++) f_rest[(size_t)nr * i + d] = (float)(0.02 * rng.normal());
    }
    return 0;
}

// Cameras on a sphere of the given radius looking at the origin (OpenGL camera-to-world,
// +y up where possible). Directions follow a Fibonacci lattice; jitter > 0 perturbs the
// direction by up to `jitter` radians and the radius by up to 10 * jitter percent.
int msplat_synth_cameras(uint64_t seed, int count, float radius, float jitter, float* c2w) {
    if (count <= 0) return -1;
    Rng rng(seed * 2654435761ull + 99ull);
    const double ga = 3.14159265358979323846 * (3.0 - std::sqrt(5.0));
    for (int k = 0; k < count; k++) {
        double y = 1.0 - 2.0 * (k + 0.5) / count;
        double r = std::sqrt(std::max(0.0, 1.0 - y * y));
        double th = ga * k;
        double d[3] = {r * std::cos(th), y * 0.9, r * std::sin(th)};
        if (jitter > 0) {
            for (int a = 0; a < 3; a++) d[a] += jitter * (2.0 * rng.uniform() - 1.0);
        }
        double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; a++) d[a] /= dn;
        double rad = radius * (jitter > 0 ? (1.0 + 0.1 * jitter * (2.0 * rng.uniform() - 1.0)) : 1.0);
        double pos[3] = {d[0] * rad, d[1] * rad, d[2] * rad};
        // OpenGL: camera looks down -z, so z axis points from the origin to the camera.
        double z[3] = {d[0], d[1], d[2]};
        double up[3] = {0, 1, 0};
        double x[3] = {up[1] * z[2] - up[2] * z[1], up[2] * z[0] - up[0] * z[2], up[0] * z[1] - up[1] * z[0]};
        double xn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        if (xn < 1e-6) { x[0] = 1; x[1] = 0; x[2] = 0; xn = 1; }
        for (int a = 0; a < 3; a++) x[a] /= xn;
        double yv[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
        float* M = c2w + 16 * (size_t)k;
        for (int rr = 0; rr < 3; rr++) {
            M[rr * 4 + 0] = (float)x[rr];
            M[rr * 4 + 1] = (float)yv[rr];
            M[rr * 4 + 2] = (float)z[rr];
            M[rr * 4 + 3] = (float)pos[rr];
        }
        M[12] = 0; M[13] = 0; M[14] = 0; M[15] = 1;
    }
    return 0;
}

// Float RGB in [0,1] -> uint8 by round-half-away-from-zero of x * 255, clamped
// (reference core/src/input_data.cpp:66-69).
void msplat_quantize_rgb(const float* src, int64_t n, uint8_t* dst) {
    for (int64_t i = 0; i < n; i++) {
        float v = std::round(src[i] * 255.0f);
        v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
        dst[i] = (uint8_t)v;
    }
}

}  // extern "C"

extern "C" {

void msplat_camera_from_fov(const float c2w[16], float fov_y, int width, int height, MsplatCamera* out) {
    // intrinsics from the vertical field of view and the render size (viewer preparePose,
    // reference core/src/msplat_api.mm:486-528)
    const float fy = (float)height / (2.0f * std::tan(fov_y * 0.5f));
    const float aspect = (float)width / (float)height;
    const float fov_x = 2.0f * std::atan(std::tan(fov_y * 0.5f) * aspect);
    const float fx = (float)width / (2.0f * std::tan(fov_x * 0.5f));
    float R[3][3], Rinv[3][3], T[3], Tinv[3];
    for (int i = 0; i < 3; i++) {
        R[i][0] = c2w[i * 4 + 0];
        R[i][1] = -c2w[i * 4 + 1];
        R[i][2] = -c2w[i * 4 + 2];
        T[i] = c2w[i * 4 + 3];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rinv[i][j] = R[j][i];
    for (int i = 0; i < 3; i++) Tinv[i] = -(Rinv[i][0] * T[0] + Rinv[i][1] * T[1] + Rinv[i][2] * T[2]);
    const float vm[16] = {Rinv[0][0], Rinv[0][1], Rinv[0][2], Tinv[0], Rinv[1][0], Rinv[1][1], Rinv[1][2], Tinv[1],
                          Rinv[2][0], Rinv[2][1], Rinv[2][2], Tinv[2], 0, 0, 0, 1};
    const float t_p = 0.001f * std::tan(0.5f * fov_y), r_p = 0.001f * std::tan(0.5f * fov_x);
    const float pm[16] = {0.001f / r_p, 0, 0, 0, 0, 0.001f / t_p, 0, 0, 0, 0, (1000.0f + 0.001f) / (1000.0f - 0.001f),
                          -1000.0f * 0.001f / (1000.0f - 0.001f), 0, 0, 1, 0};
    float pvm[16] = {};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) pvm[i * 4 + j] += pm[i * 4 + k] * vm[k * 4 + j];
    std::memcpy(out->viewmat, vm, sizeof(vm));
    std::memcpy(out->projmat, pvm, sizeof(pvm));
    out->fx = fx;
    out->fy = fy;
    out->cx = (float)width * 0.5f;
    out->cy = (float)height * 0.5f;
    out->cam_pos[0] = T[0];
    out->cam_pos[1] = T[1];
    out->cam_pos[2] = T[2];
    out->width = width;
    out->height = height;
}

}  // extern "C"

// oracle.hpp — CPU restatement of the reference training step (TEST INFRASTRUCTURE).
//
// This file is the checker, not the product: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load the library built from it. It is a scalar, templated
// (T = float for the FP32 oracle, T = double for the FP64 variant) transcription of the
// Metal kernels dispatched on the reference's training path, in the reference's operation
// order, with OpenMP over Gaussians / tiles / pixels for the timed CPU baseline.
//
// Citations are to the reference repository (semhfe/msplat):
//   core/metal/msplat_metal.metal  — kernels ("metal:<line>")
//   core/metal/msplat_metal.mm     — dispatch order, K rule, capacity ("mm:<line>")
//   core/src/model.cpp             — host step, MCMC ("model:<line>")
//   core/src/ssim.cpp              — SSIM window ("ssim:<line>")
//
// Parity pins (what checks this restatement beyond itself) are listed per function and
// exercised by tests/test_oracle.py; functions with no pin say "parity unpinned".
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/msplat_types.h"
#include "../msplat_amd/csrc/ms_math.h"

namespace oracle {

static constexpr int BLOCK = 16;            // metal:5-6 (BLOCK_X = BLOCK_Y = 16)
static constexpr int CHUNK = 512;           // metal:2943, mm:1206
static constexpr int CAP_MULT = 48;         // mm:518 (capacity = 48 N)
static constexpr uint32_t BITONIC_TG_CAP = 2048u;  // metal:2242 (selects LDS vs device sort;
                                                   // the sorted result does not depend on it)

// ---- elementary functions: float uses the project-written ms_* (shared with the GPU so
// that integer outputs can be bit-identical), double uses libm.
template <class T> inline T t_exp(T x);
template <> inline float t_exp<float>(float x) { return ms_expf(x); }
template <> inline double t_exp<double>(double x) { return std::exp(x); }
template <class T> inline T t_log(T x);
template <> inline float t_log<float>(float x) { return ms_logf(x); }
template <> inline double t_log<double>(double x) { return std::log(x); }
template <class T> inline void t_sincos(T a, T* s, T* c);
template <> inline void t_sincos<float>(float a, float* s, float* c) { ms_sincosf(a, s, c); }
template <> inline void t_sincos<double>(double a, double* s, double* c) { *s = std::sin(a); *c = std::cos(a); }
template <class T> inline T t_fma(T a, T b, T c) { return std::fma(a, b, c); }
template <class T> inline bool t_finite(T x) { return std::isfinite(x); }
template <class T> inline T t_min(T a, T b) { return std::fmin(a, b); }   // Metal min/max: fmin/fmax
template <class T> inline T t_max(T a, T b) { return std::fmax(a, b); }
template <class T> inline T t_clamp(T x, T lo, T hi) { return t_min(t_max(x, lo), hi); }
template <class T> inline int t_f2i(T x) { return ms_f2i((float)x); }
template <> inline int t_f2i<double>(double x) {
    if (!(x == x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (-2147483647 - 1);
    return (int)x;
}

// SH constants, metal:14-29 (float literals, as in the reference's constant tables)
static const float SH_C0 = 0.28209479177387814f;
static const float SH_C1 = 0.4886025119029199f;
static const float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                               -1.0925484305920792f, 0.5462742152960396f};
static const float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                               0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                               -0.5900435899266435f};
inline int num_sh_bases(int degree) {  // metal:41-51
    return degree <= 0 ? 1 : degree == 1 ? 4 : degree == 2 ? 9 : degree == 3 ? 16 : 25;
}

// 1-D SSIM window, metal:3376-3380 (== createSSIMWindow(11, 1.5), ssim:7-24).
static const float GAUSS_1D[11] = {0.0010283801f, 0.0075987581f, 0.0360007721f, 0.1093606895f,
                                   0.2130055377f, 0.2660117249f, 0.2130055377f, 0.1093606895f,
                                   0.0360007721f, 0.0075987581f, 0.0010283801f};
static const float SSIM_C1 = 0.0001f, SSIM_C2 = 0.0009f;  // (0.01)^2, (0.03)^2

// ---------------------------------------------------------------------------------------
// 3x3 matrices, row-major m[r][c]. Metal's float3x3 is column-major; every use below is
// translated with M_metal[c][r] == m[r][c].
template <class T> struct M3 {
    T m[3][3];
};
template <class T> inline M3<T> mul(const M3<T>& a, const M3<T>& b) {
    M3<T> o;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) o.m[r][c] = a.m[r][0] * b.m[0][c] + a.m[r][1] * b.m[1][c] + a.m[r][2] * b.m[2][c];
    return o;
}
template <class T> inline M3<T> transpose(const M3<T>& a) {
    M3<T> o;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) o.m[r][c] = a.m[c][r];
    return o;
}

// quat_to_rotmat, metal:134-156. Quaternion stored (w, x, y, z); the norm is summed in the
// reference's order w_field^2 + x^2 + y^2 + z^2 where the float4's .w holds z.
template <class T> inline M3<T> quat_to_rotmat(const T q[4]) {
    T n2 = q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    T s = T(1) / std::sqrt(t_max(n2, T(1e-8f)));
    T w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
    M3<T> R;
    R.m[0][0] = T(1) - T(2) * (y * y + z * z);
    R.m[1][0] = T(2) * (x * y + w * z);
    R.m[2][0] = T(2) * (x * z - w * y);
    R.m[0][1] = T(2) * (x * y - w * z);
    R.m[1][1] = T(1) - T(2) * (x * x + z * z);
    R.m[2][1] = T(2) * (y * z + w * x);
    R.m[0][2] = T(2) * (x * z + w * y);
    R.m[1][2] = T(2) * (y * z - w * x);
    R.m[2][2] = T(1) - T(2) * (x * x + y * y);
    return R;
}

// scale_rot_to_cov3d, metal:200-213: M = R S (column c of R scaled by s_c), cov = M M^T.
template <class T> inline void cov3d_from(const T s[3], const T q[4], T cov[6]) {
    M3<T> R = quat_to_rotmat<T>(q);
    T M[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M[r][c] = R.m[r][c] * s[c];
    auto dotr = [&](int a, int b) { return M[a][0] * M[b][0] + M[a][1] * M[b][1] + M[a][2] * M[b][2]; };
    cov[0] = dotr(0, 0);
    cov[1] = dotr(1, 0);
    cov[2] = dotr(2, 0);
    cov[3] = dotr(1, 1);
    cov[4] = dotr(2, 1);
    cov[5] = dotr(2, 2);
}

// transform_4x3 / transform_4x4, metal:90-108
template <class T> inline void xform3(const float* m, const T p[3], T o[3]) {
    for (int r = 0; r < 3; r++) o[r] = T(m[r * 4 + 0]) * p[0] + T(m[r * 4 + 1]) * p[1] + T(m[r * 4 + 2]) * p[2] + T(m[r * 4 + 3]);
}
template <class T> inline void xform4(const float* m, const T p[3], T o[4]) {
    for (int r = 0; r < 4; r++) o[r] = T(m[r * 4 + 0]) * p[0] + T(m[r * 4 + 1]) * p[1] + T(m[r * 4 + 2]) * p[2] + T(m[r * 4 + 3]);
}

// project_cov3d_ewa (register overload), metal:266-306
template <class T>
inline void project_cov3d_ewa(const T cov[6], const float* vm, T fx, T fy, T tan_fovx, T tan_fovy, T pv[3], T out[3]) {
    T lim_x = T(1.3f) * tan_fovx, lim_y = T(1.3f) * tan_fovy;
    T px = pv[2] * t_min(lim_x, t_max(-lim_x, pv[0] / pv[2]));
    T py = pv[2] * t_min(lim_y, t_max(-lim_y, pv[1] / pv[2]));
    T rz = T(1) / pv[2];
    T rz2 = rz * rz;
    T j00 = fx * rz, j11 = fy * rz;
    T j20 = -fx * px * rz2, j21 = -fy * py * rz2;
    T t0[3], t1[3];
    for (int i = 0; i < 3; i++) {
        t0[i] = j00 * T(vm[0 + i]) + j20 * T(vm[8 + i]);
        t1[i] = j11 * T(vm[4 + i]) + j21 * T(vm[8 + i]);
    }
    T v00 = cov[0], v01 = cov[1], v02 = cov[2], v11 = cov[3], v12 = cov[4], v22 = cov[5];
    T tv0[3] = {t0[0] * v00 + t0[1] * v01 + t0[2] * v02, t0[0] * v01 + t0[1] * v11 + t0[2] * v12,
                t0[0] * v02 + t0[1] * v12 + t0[2] * v22};
    T tv1[3] = {t1[0] * v00 + t1[1] * v01 + t1[2] * v02, t1[0] * v01 + t1[1] * v11 + t1[2] * v12,
                t1[0] * v02 + t1[1] * v12 + t1[2] * v22};
    out[0] = (tv0[0] * t0[0] + tv0[1] * t0[1] + tv0[2] * t0[2]) + T(0.3f);
    out[1] = tv0[0] * t1[0] + tv0[1] * t1[1] + tv0[2] * t1[2];
    out[2] = (tv1[0] * t1[0] + tv1[1] * t1[1] + tv1[2] * t1[2]) + T(0.3f);
}

// compute_cov2d_bounds, metal:308-332
template <class T> inline bool cov2d_bounds(const T c[3], T conic[3], T* radius) {
    T det = c[0] * c[2] - c[1] * c[1];
    if (!(det > T(1e-6f))) return false;
    T inv_det = T(1) / det;
    conic[0] = c[2] * inv_det;
    conic[1] = -c[1] * inv_det;
    conic[2] = c[0] * inv_det;
    T b = T(0.5f) * (c[0] + c[2]);
    T disc = std::sqrt(t_max(T(0.1f), b * b - det));
    *radius = std::ceil(T(3) * std::sqrt(b + disc));
    return true;
}

// get_bbox / get_tile_bbox, metal:57-87
template <class T>
inline void tile_bbox(T cx, T cy, T rx, T ry, int tbx, int tby, int& x0, int& y0, int& x1, int& y1) {
    T tcx = cx / T(BLOCK), tcy = cy / T(BLOCK);
    T trx = rx / T(BLOCK), try_ = ry / T(BLOCK);
    x0 = std::min(std::max(0, t_f2i(tcx - trx)), tbx);
    x1 = std::min(std::max(0, t_f2i(tcx + trx + T(1))), tbx);
    y0 = std::min(std::max(0, t_f2i(tcy - try_)), tby);
    y1 = std::min(std::max(0, t_f2i(tcy + try_ + T(1))), tby);
}

// sh_coeffs_to_color, metal:620-687 (degrees <= 3: features_rest holds (deg+1)^2-1 bases)
template <class T> inline void sh_to_color(int deg, const T d[3], const T* dc, const T* rest, T col[3]) {
    for (int c = 0; c < 3; c++) col[c] = T(SH_C0) * dc[c];
    if (deg < 1) return;
    T x = d[0], y = d[1], z = d[2];
    T xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    for (int c = 0; c < 3; c++) {
        col[c] += T(SH_C1) * (-y * rest[0 * 3 + c] + z * rest[1 * 3 + c] - x * rest[2 * 3 + c]);
        if (deg < 2) continue;
        col[c] += (T(SH_C2[0]) * xy * rest[3 * 3 + c] + T(SH_C2[1]) * yz * rest[4 * 3 + c] +
                   T(SH_C2[2]) * (T(2) * zz - xx - yy) * rest[5 * 3 + c] + T(SH_C2[3]) * xz * rest[6 * 3 + c] +
                   T(SH_C2[4]) * (xx - yy) * rest[7 * 3 + c]);
        if (deg < 3) continue;
        col[c] += (T(SH_C3[0]) * y * (T(3) * xx - yy) * rest[8 * 3 + c] + T(SH_C3[1]) * xy * z * rest[9 * 3 + c] +
                   T(SH_C3[2]) * y * (T(4) * zz - xx - yy) * rest[10 * 3 + c] +
                   T(SH_C3[3]) * z * (T(2) * zz - T(3) * xx - T(3) * yy) * rest[11 * 3 + c] +
                   T(SH_C3[4]) * x * (T(4) * zz - xx - yy) * rest[12 * 3 + c] +
                   T(SH_C3[5]) * z * (xx - yy) * rest[13 * 3 + c] + T(SH_C3[6]) * x * (xx - T(3) * yy) * rest[14 * 3 + c]);
    }
}

// SH basis values for degrees 1..3 in the order of the fused SH backward (metal:1993-2033).
template <class T> inline void sh_basis(const T d[3], T b[15]) {
    T x = d[0], y = d[1], z = d[2];
    T xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    b[0] = -T(SH_C1) * y;
    b[1] = T(SH_C1) * z;
    b[2] = -T(SH_C1) * x;
    b[3] = T(SH_C2[0]) * xy;
    b[4] = T(SH_C2[1]) * yz;
    b[5] = T(SH_C2[2]) * (T(2) * zz - xx - yy);
    b[6] = T(SH_C2[3]) * xz;
    b[7] = T(SH_C2[4]) * (xx - yy);
    b[8] = T(SH_C3[0]) * y * (T(3) * xx - yy);
    b[9] = T(SH_C3[1]) * xy * z;
    b[10] = T(SH_C3[2]) * y * (T(4) * zz - xx - yy);
    b[11] = T(SH_C3[3]) * z * (T(2) * zz - T(3) * xx - T(3) * yy);
    b[12] = T(SH_C3[4]) * x * (T(4) * zz - xx - yy);
    b[13] = T(SH_C3[5]) * z * (xx - yy);
    b[14] = T(SH_C3[6]) * x * (xx - T(3) * yy);
}

template <class T> inline T sigmoid(T x) { return T(1) / (T(1) + t_exp<T>(-x)); }

// ---------------------------------------------------------------------------------------
// Forward projection outputs (one entry per Gaussian).
template <class T> struct Proj {
    std::vector<T> xys, depths, conics, colors, aabb, opac_sig;
    std::vector<int> radii, tiles_hit;
    void resize(int n) {
        xys.assign(2 * (size_t)n, T(0)); depths.assign(n, T(0)); conics.assign(3 * (size_t)n, T(0));
        colors.assign(3 * (size_t)n, T(0)); aabb.assign(2 * (size_t)n, T(0)); opac_sig.assign(n, T(0));
        radii.assign(n, 0); tiles_hit.assign(n, 0);
    }
};

// project_and_sh_forward_kernel, metal:1719-1834. Culled Gaussians report radii = 0 and
// num_tiles_hit = 0 (their other outputs are left zero here; the reference leaves them
// stale). Pins: closed forms (R orthonormal, conic * cov2d = I, EWA J W S W^T J^T by an
// independent numpy FP64 evaluation) and bit-exact agreement of the FP32 integer outputs
// with the FP64 variant away from rounding boundaries.
template <class T>
void project_forward(int N, const T* means, const T* scales, const T* quats, const T* f_dc, const T* f_rest,
                     const T* opac, int degree, int degrees_to_use, const MsplatCamera& cam, Proj<T>& out) {
    out.resize(N);
    const int W = cam.width, H = cam.height;
    const int tbx = (W + BLOCK - 1) / BLOCK, tby = (H + BLOCK - 1) / BLOCK;
    const int nb = num_sh_bases(degree) - 1;
    const T fx = cam.fx, fy = cam.fy, cx = cam.cx, cy = cam.cy;
    const T tan_fovx = T(0.5f) * T((float)W) / fx, tan_fovy = T(0.5f) * T((float)H) / fy;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        const T* p = means + 3 * (size_t)i;
        T pv[3];
        xform3<T>(cam.viewmat, p, pv);
        if (pv[2] <= T(0.01f)) continue;  // clip_near_plane, clip_thresh 0.01 (model:896)
        T s[3];
        for (int k = 0; k < 3; k++) s[k] = t_exp<T>(t_clamp(scales[3 * (size_t)i + k], T(-15), T(15)));
        T q[4] = {quats[4 * (size_t)i], quats[4 * (size_t)i + 1], quats[4 * (size_t)i + 2], quats[4 * (size_t)i + 3]};
        T cov[6];
        cov3d_from<T>(s, q, cov);
        T cov2d[3];
        T pvc[3] = {pv[0], pv[1], pv[2]};
        project_cov3d_ewa<T>(cov, cam.viewmat, fx, fy, tan_fovx, tan_fovy, pvc, cov2d);
        T conic[3], radius;
        if (!cov2d_bounds<T>(cov2d, conic, &radius)) continue;
        // project_pix, metal:335-344; ndc2pix metal:53-55 (0.5 W x + c - 0.5)
        T ph[4];
        xform4<T>(cam.projmat, p, ph);
        T rw = T(1) / (ph[3] + T(1e-6f));
        T ndx = ph[0] * rw, ndy = ph[1] * rw;
        T cxp = T(0.5f) * T((float)W) * ndx + cx - T(0.5f);
        T cyp = T(0.5f) * T((float)H) * ndy + cy - T(0.5f);
        T ax = std::ceil(T(3) * std::sqrt(cov2d[0]));
        T ay = std::ceil(T(3) * std::sqrt(cov2d[2]));
        int x0, y0, x1, y1;
        tile_bbox<T>(cxp, cyp, ax, ay, tbx, tby, x0, y0, x1, y1);
        int area = (int)((uint32_t)(x1 - x0) * (uint32_t)(y1 - y0));
        if (area <= 0) continue;
        // SH colour (metal:1805-1811)
        T d[3] = {p[0] - T(cam.cam_pos[0]), p[1] - T(cam.cam_pos[1]), p[2] - T(cam.cam_pos[2])};
        T inv = T(1) / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        T vd[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
        T col[3];
        sh_to_color<T>(degrees_to_use, vd, f_dc + 3 * (size_t)i, f_rest + (size_t)nb * 3 * i, col);
        // zombie gate, metal:1818-1833
        bool ok = t_finite(cxp) && t_finite(cyp) && t_finite(pv[2]) && t_finite(ax) && t_finite(ay) &&
                  t_finite(conic[0]) && t_finite(conic[1]) && t_finite(conic[2]) && t_finite(col[0]) &&
                  t_finite(col[1]) && t_finite(col[2]);
        if (!ok) continue;
        out.radii[i] = t_f2i(radius);
        out.tiles_hit[i] = area;
        out.depths[i] = pv[2];
        out.xys[2 * (size_t)i] = cxp;
        out.xys[2 * (size_t)i + 1] = cyp;
        out.aabb[2 * (size_t)i] = ax;
        out.aabb[2 * (size_t)i + 1] = ay;
        for (int k = 0; k < 3; k++) {
            out.conics[3 * (size_t)i + k] = conic[k];
            out.colors[3 * (size_t)i + k] = col[k];
        }
        out.opac_sig[i] = sigmoid<T>(opac[i]);  // pack-time sigmoid, metal:2563
    }
}

// ---------------------------------------------------------------------------------------
// Tile intersection lists: count (metal:2342-2365), prefix_sum_tiles (metal:2376-2435),
// scatter (metal:2440-2480) and hybrid_sort_pack (metal:2489-2627). The per-tile bitonic
// networks sort unique 64-bit keys (depth bits << 32 | id), so their output equals
// std::sort of the same keys (pin: textbook sort; Σ counts = Σ num_tiles_hit).
template <class T> struct Lists {
    int num_tiles = 0, tbx = 0, tby = 0;
    int64_t capacity = 0;
    std::vector<uint32_t> counts, exact_off, sort_off;  // sort_off has num_tiles+1 entries
    std::vector<int32_t> bins;                          // [num_tiles][2]
    std::vector<int32_t> ids;                           // exact offsets, I entries
    std::vector<T> xy_opac, conic, rgb;                 // packed, 3 per entry
    int64_t num_isect = 0, num_padded = 0;
    bool overflow = false;
};

inline uint32_t pad_count(uint32_t c) {  // metal:2248-2256
    if (c == 0u) return 0u;
    uint32_t x = std::max(c, 32u), p = 1u;
    while (p < x) p <<= 1;
    return p;
}

template <class T> inline uint32_t depth_bits(T d) {
    float f = (float)d;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

template <class T>
void build_lists(int N, const Proj<T>& pj, int W, int H, Lists<T>& L) {
    L.tbx = (W + BLOCK - 1) / BLOCK;
    L.tby = (H + BLOCK - 1) / BLOCK;
    L.num_tiles = L.tbx * L.tby;
    L.capacity = (int64_t)N * CAP_MULT;
    const int NT = L.num_tiles;
    L.counts.assign(NT, 0u);
    std::vector<std::vector<uint64_t>> keys(NT);
    for (int i = 0; i < N; i++) {
        if (pj.radii[i] <= 0) continue;  // gaussian_is_active, metal:2267-2276
        T d = pj.depths[i];
        if (!t_finite(d) || d <= T(0)) continue;
        int x0, y0, x1, y1;
        tile_bbox<T>(pj.xys[2 * (size_t)i], pj.xys[2 * (size_t)i + 1], pj.aabb[2 * (size_t)i], pj.aabb[2 * (size_t)i + 1],
                     L.tbx, L.tby, x0, y0, x1, y1);
        uint64_t key = ((uint64_t)depth_bits(d) << 32) | (uint64_t)(uint32_t)i;
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) {
                L.counts[y * L.tbx + x]++;
                keys[y * L.tbx + x].push_back(key);
            }
    }
    L.exact_off.assign(NT, 0u);
    L.sort_off.assign(NT + 1, 0u);
    L.bins.assign(2 * (size_t)NT, 0);
    uint64_t ex = 0, pd = 0;
    for (int t = 0; t < NT; t++) {
        L.exact_off[t] = (uint32_t)ex;
        L.sort_off[t] = (uint32_t)pd;
        L.bins[2 * t] = (int32_t)ex;
        L.bins[2 * t + 1] = (int32_t)(ex + L.counts[t]);
        ex += L.counts[t];
        pd += pad_count(L.counts[t]);
    }
    L.sort_off[NT] = (uint32_t)pd;
    L.num_isect = (int64_t)ex;
    L.num_padded = (int64_t)pd;
    L.overflow = (int64_t)pd > L.capacity;  // the reference would drop keys here (metal:2474-2477)
    L.ids.assign((size_t)ex, 0);
    L.xy_opac.assign(3 * (size_t)ex, T(0));
    L.conic.assign(3 * (size_t)ex, T(0));
    L.rgb.assign(3 * (size_t)ex, T(0));
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < NT; t++) {
        std::vector<uint64_t>& k = keys[t];
        std::sort(k.begin(), k.end());
        size_t base = L.exact_off[t];
        for (size_t j = 0; j < k.size(); j++) {
            uint32_t g = (uint32_t)(k[j] & 0xFFFFFFFFu);
            size_t e = base + j;
            L.ids[e] = (int32_t)g;
            T xy0 = pj.xys[2 * (size_t)g], xy1 = pj.xys[2 * (size_t)g + 1], op = pj.opac_sig[g];
            const T* cn = &pj.conics[3 * (size_t)g];
            const T* cl = &pj.colors[3 * (size_t)g];
            bool ok = t_finite(xy0) && t_finite(xy1) && t_finite(op) && t_finite(cn[0]) && t_finite(cn[1]) &&
                      t_finite(cn[2]) && t_finite(cl[0]) && t_finite(cl[1]) && t_finite(cl[2]);
            if (!ok) continue;  // pack gate: transparent entry, metal:2570-2578
            L.xy_opac[3 * e] = xy0; L.xy_opac[3 * e + 1] = xy1; L.xy_opac[3 * e + 2] = op;
            for (int c = 0; c < 3; c++) { L.conic[3 * e + c] = cn[c]; L.rgb[3 * e + c] = cl[c]; }
        }
    }
}

// K_max rule, mm:1204-1224
inline uint32_t k_max_rule(int64_t capacity, int num_tiles) {
    if (num_tiles >= 400) return 1;
    uint32_t avg = (uint32_t)(capacity / std::max(1, num_tiles));
    uint32_t cons = avg * 6u;
    uint32_t K = (cons + CHUNK - 1) / CHUNK;
    if (K < 2) K = 2;
    uint32_t abs_max = (uint32_t)((capacity + CHUNK - 1) / CHUNK);
    if (K > abs_max) K = abs_max;
    return K;
}

// ---------------------------------------------------------------------------------------
// Rasterisation forward. Per-pixel front-to-back compositing; the list walk per pixel is
// exactly the reference loop (batching in 64-entry LDS tiles does not change the order).
template <class T> struct Raster {
    int W = 0, H = 0;
    uint32_t K = 1;
    std::vector<T> img, final_T;
    std::vector<int32_t> final_idx;
    // chunked path (K > 1)
    std::vector<T> chunk_T, chunk_C, prefix_T, after_C;
    std::vector<int32_t> chunk_idx;
};

template <class T> inline T sigma_of(const T* cn, T dx, T dy) {
    return t_fma(T(0.5f), t_fma(cn[0], dx * dx, cn[2] * dy * dy), cn[1] * dx * dy);
}

// One pixel over entries [lo, hi) starting from T = 1: nd_rasterize_forward_kernel
// (metal:545-606) and rasterize_forward_chunked_kernel (metal:3024-3067) share this walk.
template <class T>
inline void composite(const Lists<T>& L, int lo, int hi, T px, T py, T& Tout, T C[3], int& last) {
    T Tr = T(1);
    C[0] = C[1] = C[2] = T(0);
    last = lo - 1;
    for (int e = lo; e < hi; e++) {
        const T* xo = &L.xy_opac[3 * (size_t)e];
        const T* cn = &L.conic[3 * (size_t)e];
        T dx = xo[0] - px, dy = xo[1] - py;
        T sigma = sigma_of(cn, dx, dy);
        if (!t_finite(sigma) || sigma < T(0) || sigma >= T(5.55f)) continue;
        T alpha = t_min(T(0.999f), xo[2] * t_exp<T>(-sigma));
        if (alpha < T(1.f / 255.f)) continue;
        T next_T = Tr * (T(1) - alpha);
        if (next_T <= T(1e-4f)) { last = e - 1; break; }
        T vis = alpha * Tr;
        for (int c = 0; c < 3; c++) C[c] = t_fma(t_max(L.rgb[3 * (size_t)e + c] + T(0.5f), T(0)), vis, C[c]);
        Tr = next_T;
        last = e;
    }
    Tout = Tr;
}

template <class T> inline T saturate(T x) { return t_min(t_max(x, T(0)), T(1)); }

template <class T>
void raster_forward(const Lists<T>& L, int W, int H, const float bg[3], uint32_t K, Raster<T>& R) {
    R.W = W; R.H = H; R.K = K;
    const size_t P = (size_t)W * H;
    R.img.assign(3 * P, T(0));
    R.final_T.assign(P, T(0));
    R.final_idx.assign(P, 0);
    if (K <= 1) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < H; i++)
            for (int j = 0; j < W; j++) {
                int t = (i / BLOCK) * L.tbx + (j / BLOCK);
                T Tr, C[3];
                int last;
                composite(L, L.bins[2 * t], L.bins[2 * t + 1], T((float)j), T((float)i), Tr, C, last);
                size_t pid = (size_t)i * W + j;
                R.final_T[pid] = Tr;
                R.final_idx[pid] = last;
                for (int c = 0; c < 3; c++) R.img[3 * pid + c] = saturate(t_fma(T(bg[c]), Tr, C[c]));
            }
        return;
    }
    // chunked forward (metal:2960-3076) + merge (metal:3081-3134)
    R.chunk_T.assign((size_t)K * P, T(1));
    R.chunk_C.assign((size_t)K * P * 3, T(0));
    R.chunk_idx.assign((size_t)K * P, -1);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            int t = (i / BLOCK) * L.tbx + (j / BLOCK);
            int s = L.bins[2 * t], e = L.bins[2 * t + 1];
            size_t pid = (size_t)i * W + j;
            for (uint32_t k = 0; k < K; k++) {
                int cs = s + (int)(k * CHUNK), ce = std::min(s + (int)((k + 1) * CHUNK), e);
                size_t o = (size_t)k * P + pid;
                if (cs >= ce) { R.chunk_T[o] = T(1); R.chunk_idx[o] = -1; continue; }
                T Tr, C[3];
                int last;
                composite(L, cs, ce, T((float)j), T((float)i), Tr, C, last);
                R.chunk_T[o] = Tr;
                for (int c = 0; c < 3; c++) R.chunk_C[3 * o + c] = C[c];
                R.chunk_idx[o] = last;
            }
            T Tr = T(1), C[3] = {T(0), T(0), T(0)};
            int last = -1;
            uint32_t cutoff = K;
            for (uint32_t k = 0; k < K; k++) {
                size_t o = (size_t)k * P + pid;
                int cf = R.chunk_idx[o];
                if (cf < 0 && k > 0) break;
                T cT = R.chunk_T[o];
                for (int c = 0; c < 3; c++) C[c] = t_fma(R.chunk_C[3 * o + c], Tr, C[c]);
                Tr *= cT;
                if (cf >= 0) last = cf;
                if (Tr <= T(1e-4f)) { cutoff = k + 1; break; }
            }
            for (uint32_t k = cutoff; k < K; k++) R.chunk_idx[(size_t)k * P + pid] = -1;
            R.final_T[pid] = Tr;
            R.final_idx[pid] = last;
            for (int c = 0; c < 3; c++) R.img[3 * pid + c] = saturate(t_fma(T(bg[c]), Tr, C[c]));
        }
}

// ---------------------------------------------------------------------------------------
// Loss-value bookkeeping of ssim_fused_v_fwd_h_bwd (metal:3598-3680): the SSIM and L1 terms
// of the pixel at item i = dy*26 + dx of a 16x16 threadgroup's 16x26 derivative sweep are
// accumulated by thread i mod 256, and a thread adds them to the loss only when its own
// pixel lies inside the image. Every in-image thread adds the constant w once. So in
// threadgroups cut by the right/bottom border some in-image pixels contribute only w.
// (No effect on gradients, and none when W and H are multiples of 16.)
inline bool ssim_loss_counted(int gx, int gy, int W, int H) {
    int tx = gx / 16, ty = gy / 16;
    int i = (gy - ty * 16) * 26 + (gx - tx * 16 + 5);
    int tr = i < 256 ? i : i - 256;
    return tx * 16 + (tr % 16) < W && ty * 16 + (tr / 16) < H;
}

// ---------------------------------------------------------------------------------------
// Fused SSIM + L1 loss forward/backward: ssim_h_fwd (metal:3388-3450),
// ssim_fused_v_fwd_h_bwd (metal:3563-3681), ssim_v_bwd (metal:3764-3850).
// Zero padding; the derivative fields at columns outside the image come from zeroed
// statistics (f1 = 0, f2 = -2/C2, f3 = 2/C2) and reach pixels within 5 columns of the
// left/right borders through the horizontal transpose convolution (reference behaviour,
// reproduced). Returns the loss sum in double. Optional training mask [H W] (metal:
// 3602-3653, 3815-3834): a pixel whose mask value is below 128 has its derivative fields and
// its SSIM and L1 loss terms multiplied by zero (its loss still adds w) and its own L1
// gradient dropped. Pins: SSIM map against an independent numpy zero-padded SSIM;
// gradient by FP64 finite differences away from side borders, and the border term against
// its closed form; the masked loss and gradient against the same numpy formulation with
// masked pixels' terms removed (finite differences of that objective).
template <class T>
double ssim_loss(int W, int H, const T* rend, const uint8_t* gt, T ssim_w, T inv_n, std::vector<T>& v_rend,
                 const uint8_t* mask = nullptr) {
    const size_t P = (size_t)W * H;
    auto M = [&](int y, int x) -> T { return (mask && mask[(size_t)y * W + x] < 128) ? T(0) : T(1); };
    std::vector<T> hb(15 * P);
    auto X = [&](int y, int x, int c) -> T { return T((float)gt[((size_t)y * W + x) * 3 + c] * (1.0f / 255.0f)); };
    auto Y = [&](int y, int x, int c) -> T { return rend[((size_t)y * W + x) * 3 + c]; };
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int c = 0; c < 3; c++) {
                T mx = 0, my = 0, sx = 0, sy = 0, cxy = 0;
                for (int k = 0; k < 11; k++) {
                    int gx = x - 5 + k;
                    T w = T(GAUSS_1D[k]);
                    T gv = 0, rv = 0;
                    if (gx >= 0 && gx < W) { gv = X(y, gx, c); rv = Y(y, gx, c); }
                    // fused multiply-adds, in the device kernel's order (its results are
                    // compared bit for bit)
                    const T wg = w * gv, wr = w * rv;
                    mx = std::fma(w, gv, mx); my = std::fma(w, rv, my);
                    sx = std::fma(wg, gv, sx); sy = std::fma(wr, rv, sy); cxy = std::fma(wg, rv, cxy);
                }
                size_t o = ((size_t)y * W + x) * 15 + c * 5;
                hb[o] = mx; hb[o + 1] = my; hb[o + 2] = sx; hb[o + 3] = sy; hb[o + 4] = cxy;
            }
    // derivative fields at (row y in [0,H), column x in [-5, W+5)), then H transpose conv
    const int WP = W + 10;
    std::vector<T> f(3 * (size_t)H * WP * 3);  // [y][x+5][c][3]
    std::vector<double> row_loss(H, 0.0);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        double rl = 0.0;
        for (int xx = -5; xx < W + 5; xx++) {
            T ssim_sum_c[3] = {0, 0, 0}, l1_c[3] = {0, 0, 0};
            for (int c = 0; c < 3; c++) {
                T m[5] = {0, 0, 0, 0, 0};
                if (xx >= 0 && xx < W) {
                    for (int k = 0; k < 11; k++) {
                        int gy = y - 5 + k;
                        T w = T(GAUSS_1D[k]);
                        for (int q = 0; q < 5; q++) {
                            T hv = (gy >= 0 && gy < H) ? hb[((size_t)gy * W + xx) * 15 + c * 5 + q] : T(0);
                            m[q] = std::fma(w, hv, m[q]);
                        }
                    }
                }
                T* fo = &f[(((size_t)y * WP + (xx + 5)) * 3 + c) * 3];
                if (!(t_finite(m[0]) && t_finite(m[1]) && t_finite(m[2]) && t_finite(m[3]) && t_finite(m[4]))) {
                    fo[0] = fo[1] = fo[2] = T(0);
                    continue;
                }
                T mu_x = m[0], mu_y = m[1];
                T sxx = t_max(m[2] - mu_x * mu_x, T(0)), syy = t_max(m[3] - mu_y * mu_y, T(0));
                T sxy = m[4] - mu_x * mu_y;
                T A = T(2) * mu_x * mu_y + T(SSIM_C1), B = T(2) * sxy + T(SSIM_C2);
                T Cd = mu_x * mu_x + mu_y * mu_y + T(SSIM_C1), D = sxx + syy + T(SSIM_C2);
                T iCD = T(1) / (Cd * D);
                T dmu = T(2) * B * (mu_x * Cd - A * mu_y) / (Cd * Cd * D);
                T dsyq = -A * B * iCD / D, dsxy = T(2) * A * iCD;
                const T mv = (xx >= 0 && xx < W) ? M(y, xx) : T(1);
                fo[0] = mv * (dmu - T(2) * mu_y * dsyq - mu_x * dsxy);
                fo[1] = mv * (T(2) * dsyq);
                fo[2] = mv * dsxy;
                if (xx >= 0 && xx < W) {
                    ssim_sum_c[c] = mv * (A * B) / (Cd * D);
                    T d = X(y, xx, c) - Y(y, xx, c);
                    l1_c[c] = mv * std::fabs(d);
                }
            }
            if (xx >= 0 && xx < W) {
                T ss = 0, l1 = 0;
                for (int c = 0; c < 3; c++) { ss += ssim_sum_c[c]; l1 += l1_c[c]; }
                T pl = ssim_loss_counted(xx, y, W, H) ? ssim_w * (T(1) - ss / T(3)) + (T(1) - ssim_w) * l1 / T(3)
                                                      : ssim_w;
                rl += (double)pl;
            }
        }
        row_loss[y] = rl;
    }
    double loss = 0.0;
    for (int y = 0; y < H; y++) loss += row_loss[y];
    std::vector<T> dh(9 * P);  // [y][x][c][3]
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int c = 0; c < 3; c++)
                for (int q = 0; q < 3; q++) {
                    T h = 0;
                    for (int k = 0; k < 11; k++) {
                        int xx = x - 5 + k;
                        h = std::fma(T(GAUSS_1D[10 - k]), f[(((size_t)y * WP + (xx + 5)) * 3 + c) * 3 + q], h);
                    }
                    dh[(((size_t)y * W + x) * 3 + c) * 3 + q] = h;
                }
    v_rend.assign(3 * P, T(0));
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int c = 0; c < 3; c++) {
                T cf[3] = {0, 0, 0};
                for (int k = 0; k < 11; k++) {
                    int gy = y - 5 + k;
                    T w = T(GAUSS_1D[10 - k]);
                    for (int q = 0; q < 3; q++) {
                        T hv = (gy >= 0 && gy < H) ? dh[(((size_t)gy * W + x) * 3 + c) * 3 + q] : T(0);
                        cf[q] = std::fma(w, hv, cf[q]);
                    }
                }
                T rv = Y(y, x, c), gv = X(y, x, c);
                T v_ssim = cf[0] + rv * cf[1] + gv * cf[2];
                T v_l1 = M(y, x) * ((gv > rv) ? T(-1) : ((gv < rv) ? T(1) : T(0)));
                T v = inv_n * (-ssim_w * v_ssim + (T(1) - ssim_w) * v_l1);
                if (!t_finite(v)) v = T(0);
                v_rend[((size_t)y * W + x) * 3 + c] = v;
            }
    return loss;
}

// ---------------------------------------------------------------------------------------
// Rasterisation backward: rasterize_backward_kernel (metal:931-1135) for K = 1 and
// compute_chunk_prefix_suffix + rasterize_backward_chunked (metal:3139-3366) for K > 1.
// Per-(pixel, entry) terms are summed per entry in pixel order, then per Gaussian in
// entry order (deterministic; the reference sums with float atomics in arbitrary order).
// Pin: FP64 finite differences of the composited image (tests/test_oracle.py) except
// where the alpha = 0.999 clamp is active (skipped in backward, reference behaviour).
template <class T> struct RasterGrads {
    std::vector<T> v_xy, v_conic, v_rgb, v_opac;
};

template <class T>
inline void pixel_backward(const Lists<T>& L, int lo, int hi, int bin_final, T px, T py, T Tstart, const T buf0[3],
                           T T_final, const float bg[3], const T v_out[3], T* acc /* [entries][9] */) {
    T Tr = Tstart;
    T buffer[3] = {buf0[0], buf0[1], buf0[2]};
    T tfb[3] = {T_final * T(bg[0]), T_final * T(bg[1]), T_final * T(bg[2])};
    for (int e = hi - 1; e >= lo; e--) {
        if (e > bin_final) continue;
        const T* xo = &L.xy_opac[3 * (size_t)e];
        const T* cn = &L.conic[3 * (size_t)e];
        T opac = xo[2];
        T dx = xo[0] - px, dy = xo[1] - py;
        T sigma = sigma_of(cn, dx, dy);
        if (sigma < T(0) || sigma >= T(5.55f)) continue;
        T vis = t_exp<T>(-sigma);
        T alpha = t_min(T(0.999f), opac * vis);
        if (alpha < T(1.f / 255.f)) continue;
        if (!(alpha < T(0.999f))) continue;
        T ra = T(1) / (T(1) - alpha);
        Tr *= ra;
        T fac = alpha * Tr;
        T rgb[3], tmp[3];
        for (int c = 0; c < 3; c++) {
            rgb[c] = t_max(L.rgb[3 * (size_t)e + c] + T(0.5f), T(0));
            tmp[c] = t_fma(rgb[c], Tr, t_fma(-buffer[c], ra, -ra * tfb[c]));
        }
        T v_alpha = tmp[0] * v_out[0] + tmp[1] * v_out[1] + tmp[2] * v_out[2];
        for (int c = 0; c < 3; c++) buffer[c] = t_fma(rgb[c], fac, buffer[c]);
        T v_sigma = -alpha * v_alpha;
        T* a = acc + 9 * (size_t)(e - lo);
        for (int c = 0; c < 3; c++)
            if (L.rgb[3 * (size_t)e + c] + T(0.5f) >= T(0)) a[c] += fac * v_out[c];
        a[3] += (T(0.5f) * v_sigma) * (dx * dx);
        a[4] += (T(0.5f) * v_sigma) * (dx * dy);
        a[5] += (T(0.5f) * v_sigma) * (dy * dy);
        a[6] += v_sigma * t_fma(cn[0], dx, cn[1] * dy);
        a[7] += v_sigma * t_fma(cn[1], dx, cn[2] * dy);
        a[8] += -v_sigma * (T(1) - opac);
    }
}

template <class T>
void raster_backward(const Lists<T>& L, Raster<T>& R, const float bg[3], const T* v_out_img, int N, RasterGrads<T>& G) {
    const int W = R.W, H = R.H;
    const size_t P = (size_t)W * H;
    const int NT = L.num_tiles;
    std::vector<T> acc(9 * (size_t)L.num_isect, T(0));
    if (R.K <= 1) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < NT; t++) {
            int lo = L.bins[2 * t], hi = L.bins[2 * t + 1];
            if (hi <= lo) continue;
            int ty = t / L.tbx, tx = t % L.tbx;
            T buf0[3] = {0, 0, 0};
            for (int i = ty * BLOCK; i < std::min(H, (ty + 1) * BLOCK); i++)
                for (int j = tx * BLOCK; j < std::min(W, (tx + 1) * BLOCK); j++) {
                    size_t pid = (size_t)i * W + j;
                    pixel_backward(L, lo, hi, R.final_idx[pid], T((float)j), T((float)i), R.final_T[pid], buf0,
                                   R.final_T[pid], bg, &v_out_img[3 * pid], &acc[9 * (size_t)lo]);
                }
        }
    } else {
        const uint32_t K = R.K;
        R.prefix_T.assign((size_t)K * P, T(0));
        R.after_C.assign((size_t)K * P * 3, T(0));
        for (size_t pid = 0; pid < P; pid++) {  // compute_chunk_prefix_suffix_kernel
            T pT = 1;
            for (uint32_t k = 0; k < K; k++) { R.prefix_T[k * P + pid] = pT; pT *= R.chunk_T[k * P + pid]; }
            T aC[3] = {0, 0, 0};
            for (int k = (int)K - 1; k >= 0; k--) {
                size_t o = (size_t)k * P + pid;
                for (int c = 0; c < 3; c++) R.after_C[3 * o + c] = aC[c];
                T pk = R.prefix_T[o];
                for (int c = 0; c < 3; c++) aC[c] += pk * R.chunk_C[3 * o + c];
            }
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < NT; t++) {
            int lo = L.bins[2 * t], hi = L.bins[2 * t + 1];
            if (hi <= lo) continue;
            int ty = t / L.tbx, tx = t % L.tbx;
            for (uint32_t k = 0; k < K; k++) {
                int cs = lo + (int)(k * CHUNK), ce = std::min(lo + (int)((k + 1) * CHUNK), hi);
                if (cs >= ce) break;
                for (int i = ty * BLOCK; i < std::min(H, (ty + 1) * BLOCK); i++)
                    for (int j = tx * BLOCK; j < std::min(W, (tx + 1) * BLOCK); j++) {
                        size_t pid = (size_t)i * W + j;
                        size_t o = (size_t)k * P + pid;
                        int bf = R.chunk_idx[o];
                        if (bf < cs) continue;
                        T Ts = R.prefix_T[o] * R.chunk_T[o];
                        T b0[3] = {R.after_C[3 * o], R.after_C[3 * o + 1], R.after_C[3 * o + 2]};
                        pixel_backward(L, cs, ce, bf, T((float)j), T((float)i), Ts, b0, R.final_T[pid], bg,
                                       &v_out_img[3 * pid], &acc[9 * (size_t)cs]);
                    }
            }
        }
    }
    G.v_xy.assign(2 * (size_t)N, T(0));
    G.v_conic.assign(3 * (size_t)N, T(0));
    G.v_rgb.assign(3 * (size_t)N, T(0));
    G.v_opac.assign((size_t)N, T(0));
    for (int64_t e = 0; e < L.num_isect; e++) {
        const T* a = &acc[9 * (size_t)e];
        size_t g = (size_t)L.ids[e];
        for (int c = 0; c < 3; c++) G.v_rgb[3 * g + c] += a[c];
        for (int c = 0; c < 3; c++) G.v_conic[3 * g + c] += a[3 + c];
        G.v_xy[2 * g] += a[6];
        G.v_xy[2 * g + 1] += a[7];
        G.v_opac[g] += a[8];
    }
}

// ---------------------------------------------------------------------------------------
// Projection backward: project_pix_vjp (metal:1252-1270; the perspective divide rw is held
// fixed, as in the reference), cov2d_to_conic_vjp (metal:1289-1300),
// project_cov3d_ewa_vjp (metal:1375-1443), scale_rot_to_cov3d_vjp (metal:1536-1569),
// quat_to_rotmat_vjp (metal:1445-1490; gradient w.r.t. the normalised components, applied
// to the stored quaternion), log-scale chain rule (metal:1955-1958).
// Pins: FP64 central finite differences of a forward with rw frozen (means), of the
// conic (scales), and the tangential part of the quaternion gradient.
template <class T>
void proj_backward_one(const T p[3], const T ls[3], const T q[4], const MsplatCamera& cam, const T conic[3],
                       const T v_xy[2], const T v_conic[3], T v_mean[3], T v_scale[3], T v_quat[4]) {
    const float* pm = cam.projmat;
    const float* vm = cam.viewmat;
    const int W = cam.width, H = cam.height;
    T ph[4];
    xform4<T>(pm, p, ph);
    T rw = T(1) / (ph[3] + T(1e-6f));
    T vnx = T(0.5f) * T((float)W) * v_xy[0], vny = T(0.5f) * T((float)H) * v_xy[1];
    T vp[3] = {vnx * rw, vny * rw, T(0)};
    v_mean[0] = T(pm[0]) * vp[0] + T(pm[4]) * vp[1] + T(pm[8]) * vp[2];
    v_mean[1] = T(pm[1]) * vp[0] + T(pm[5]) * vp[1] + T(pm[9]) * vp[2];
    v_mean[2] = T(pm[2]) * vp[0] + T(pm[6]) * vp[1] + T(pm[10]) * vp[2];
    // v_depth is never written by the reference and stays 0 (mm:1506): the z term adds 0.
    // cov2d_to_conic_vjp: v_Sigma = -X G X
    T X[2][2] = {{conic[0], conic[1]}, {conic[1], conic[2]}};
    T G[2][2] = {{v_conic[0], v_conic[1]}, {v_conic[1], v_conic[2]}};
    T P[2][2], Q[2][2];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) P[r][c] = (-X[r][0]) * G[0][c] + (-X[r][1]) * G[1][c];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) Q[r][c] = P[r][0] * X[0][c] + P[r][1] * X[1][c];
    T vc2[3] = {Q[0][0], Q[0][1] + Q[1][0], Q[1][1]};
    // recompute cov3d
    T s[3];
    for (int k = 0; k < 3; k++) s[k] = t_exp<T>(t_clamp(ls[k], T(-15), T(15)));
    T cov[6];
    cov3d_from<T>(s, q, cov);
    T fx = cam.fx, fy = cam.fy;
    T tan_fovx = T(0.5f) * T((float)W) / fx, tan_fovy = T(0.5f) * T((float)H) / fy;
    T pv[3];
    xform3<T>(vm, p, pv);
    T lim_x = T(1.3f) * tan_fovx, lim_y = T(1.3f) * tan_fovy;
    T pvx = pv[2] * t_min(lim_x, t_max(-lim_x, pv[0] / pv[2]));
    T pvy = pv[2] * t_min(lim_y, t_max(-lim_y, pv[1] / pv[2]));
    T rz = T(1) / pv[2], rz2 = rz * rz;
    M3<T> Wm, J, V, vcov;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Wm.m[r][c] = T(vm[r * 4 + c]);
    J.m[0][0] = fx * rz; J.m[0][1] = 0; J.m[0][2] = -fx * pvx * rz2;
    J.m[1][0] = 0; J.m[1][1] = fy * rz; J.m[1][2] = -fy * pvy * rz2;
    J.m[2][0] = 0; J.m[2][1] = 0; J.m[2][2] = 0;
    V.m[0][0] = cov[0]; V.m[0][1] = cov[1]; V.m[0][2] = cov[2];
    V.m[1][0] = cov[1]; V.m[1][1] = cov[3]; V.m[1][2] = cov[4];
    V.m[2][0] = cov[2]; V.m[2][1] = cov[4]; V.m[2][2] = cov[5];
    vcov.m[0][0] = vc2[0]; vcov.m[0][1] = T(0.5f) * vc2[1]; vcov.m[0][2] = 0;
    vcov.m[1][0] = T(0.5f) * vc2[1]; vcov.m[1][1] = vc2[2]; vcov.m[1][2] = 0;
    vcov.m[2][0] = 0; vcov.m[2][1] = 0; vcov.m[2][2] = 0;
    M3<T> Tm = mul(J, Wm);
    M3<T> Tt = transpose(Tm);
    M3<T> vV = mul(mul(Tt, vcov), Tm);
    M3<T> vT1 = mul(mul(vcov, Tm), transpose(V));
    M3<T> vT2 = mul(mul(transpose(vcov), Tm), V);
    M3<T> vT;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) vT.m[r][c] = vT1.m[r][c] + vT2.m[r][c];
    T vcov3[6] = {vV.m[0][0], vV.m[1][0] + vV.m[0][1], vV.m[2][0] + vV.m[0][2],
                  vV.m[1][1], vV.m[2][1] + vV.m[1][2], vV.m[2][2]};
    M3<T> vJ = mul(vT, transpose(Wm));
    T fx_rz2 = fx * rz2, fy_rz2 = fy * rz2, rz3 = rz2 * rz;
    T vt[3] = {-fx_rz2 * vJ.m[0][2], -fy_rz2 * vJ.m[1][2],
               -fx_rz2 * vJ.m[0][0] + T(2) * fx * pvx * rz3 * vJ.m[0][2] - fy_rz2 * vJ.m[1][1] +
                   T(2) * fy * pvy * rz3 * vJ.m[1][2]};
    for (int i = 0; i < 3; i++) v_mean[i] += vt[0] * Wm.m[0][i] + vt[1] * Wm.m[1][i] + vt[2] * Wm.m[2][i];
    // scale_rot_to_cov3d_vjp
    M3<T> vVs;
    vVs.m[0][0] = vcov3[0]; vVs.m[0][1] = T(0.5f) * vcov3[1]; vVs.m[0][2] = T(0.5f) * vcov3[2];
    vVs.m[1][0] = T(0.5f) * vcov3[1]; vVs.m[1][1] = vcov3[3]; vVs.m[1][2] = T(0.5f) * vcov3[4];
    vVs.m[2][0] = T(0.5f) * vcov3[2]; vVs.m[2][1] = T(0.5f) * vcov3[4]; vVs.m[2][2] = vcov3[5];
    M3<T> R = quat_to_rotmat<T>(q);
    M3<T> M;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M.m[r][c] = R.m[r][c] * s[c];
    M3<T> vVs2;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) vVs2.m[r][c] = T(2) * vVs.m[r][c];
    M3<T> vM = mul(vVs2, M);
    for (int i = 0; i < 3; i++) v_scale[i] = R.m[0][i] * vM.m[0][i] + R.m[1][i] * vM.m[1][i] + R.m[2][i] * vM.m[2][i];
    M3<T> vR;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) vR.m[r][c] = vM.m[r][c] * s[c];
    T n2 = q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    T sn = T(1) / std::sqrt(t_max(n2, T(1e-8f)));
    T w = q[0] * sn, x = q[1] * sn, y = q[2] * sn, z = q[3] * sn;
    // v_R[c][r] (Metal, column-major) == vR.m[r][c]
    auto g = [&](int c, int r) { return vR.m[r][c]; };
    v_quat[0] = T(2) * (x * (g(1, 2) - g(2, 1)) + y * (g(2, 0) - g(0, 2)) + z * (g(0, 1) - g(1, 0)));
    v_quat[1] = T(2) * (-T(2) * x * (g(1, 1) + g(2, 2)) + y * (g(0, 1) + g(1, 0)) + z * (g(0, 2) + g(2, 0)) +
                        w * (g(1, 2) - g(2, 1)));
    v_quat[2] = T(2) * (x * (g(0, 1) + g(1, 0)) - T(2) * y * (g(0, 0) + g(2, 2)) + z * (g(1, 2) + g(2, 1)) +
                        w * (g(2, 0) - g(0, 2)));
    v_quat[3] = T(2) * (x * (g(0, 2) + g(2, 0)) + y * (g(1, 2) + g(2, 1)) - T(2) * z * (g(0, 0) + g(1, 1)) +
                        w * (g(0, 1) - g(1, 0)));
    for (int k = 0; k < 3; k++) v_scale[k] *= s[k];
}

// adam_update_element, metal:1838-1847 (also fused_adam_kernel metal:1706-1712)
template <class T> inline void adam(T& p, T& m, T& v, T g, T step, T b1, T b2, T bc2s, T eps) {
    T mn = t_fma(b1, m, (T(1) - b1) * g);
    T vn = t_fma(b2, v, (T(1) - b2) * g * g);
    p -= step * mn / (std::sqrt(vn) / bc2s + eps);
    m = mn;
    v = vn;
}

// PCG hash + unit mapping, metal:3862-3870 (integer-exact).
inline uint32_t pcg_hash32(uint32_t seed, uint32_t idx) {
    uint32_t x = seed * 747796405u + idx * 2891336453u;
    x = ((x >> ((x >> 28) + 4)) ^ x) * 277803737u;
    return (x >> 22) ^ x;
}
template <class T> inline T pcg_unit(uint32_t h) { return (T((float)(h >> 9)) + T(0.5f)) * T(1.0f / (float)(1u << 23)); }

// sgld_noise_gen_kernel, metal:3871-3890 (Box-Muller on 4 hashed uniforms)
template <class T> inline void sgld_noise3(uint32_t seed, uint32_t idx, T n[3]) {
    uint32_t base = idx * 4u;
    T u0 = pcg_unit<T>(pcg_hash32(seed, base)), u1 = pcg_unit<T>(pcg_hash32(seed, base + 1));
    T u2 = pcg_unit<T>(pcg_hash32(seed, base + 2)), u3 = pcg_unit<T>(pcg_hash32(seed, base + 3));
    T r01 = std::sqrt(T(-2) * t_log<T>(u0)), r23 = std::sqrt(T(-2) * t_log<T>(u2));
    T a01 = T(6.283185307179586f) * u1, a23 = T(6.283185307179586f) * u3;
    T s01, c01, s23, c23;
    t_sincos<T>(a01, &s01, &c01);
    t_sincos<T>(a23, &s23, &c23);
    n[0] = r01 * c01;
    n[1] = r01 * s01;
    n[2] = r23 * c23;
}

// sgld_noise_kernel, metal:3892-3967: the mean moves by R diag(exp(2 s)) R^T (noise_lr *
// xyz_lr * w(o) * n), w = sigmoid(-100 ((1 - sigmoid(o)) - 0.995)), capped at length max_disp.
// Pins (tests/test_oracle_pins.py): identity rotation and isotropic scale give s^2 * ns * n;
// a 90-degree turn about z maps the local y scale onto world x; w = 1/2 at sigmoid(o) = 0.005;
// a capped move has length max_disp along the uncapped direction; rotating the frame by Q
// (q -> q_Q q, n -> Q n) rotates the move by Q.
template <class T>
inline void sgld_apply(T mean[3], const T ls[3], const T q[4], T opac_logit, const T n[3], T noise_lr, T xyz_lr, T max_disp) {
    T op_sig = T(1) / (T(1) + t_exp<T>(-opac_logit));
    T weight = T(1) / (T(1) + t_exp<T>(T(-100) * ((T(1) - op_sig) - T(0.995f))));
    T ns = noise_lr * xyz_lr * weight;
    T n0 = n[0] * ns, n1 = n[1] * ns, n2 = n[2] * ns;
    T qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    T ql = std::sqrt(t_max(qw * qw + qx * qx + qy * qy + qz * qz, T(1e-8f)));
    qw /= ql; qx /= ql; qy /= ql; qz /= ql;
    T sx = t_exp<T>(t_clamp(ls[0], T(-15), T(15))), sy = t_exp<T>(t_clamp(ls[1], T(-15), T(15))),
      sz = t_exp<T>(t_clamp(ls[2], T(-15), T(15)));
    T rt0 = (T(1) - T(2) * (qy * qy + qz * qz)) * n0 + T(2) * (qx * qy + qw * qz) * n1 + T(2) * (qx * qz - qw * qy) * n2;
    T rt1 = T(2) * (qx * qy - qw * qz) * n0 + (T(1) - T(2) * (qx * qx + qz * qz)) * n1 + T(2) * (qy * qz + qw * qx) * n2;
    T rt2 = T(2) * (qx * qz + qw * qy) * n0 + T(2) * (qy * qz - qw * qx) * n1 + (T(1) - T(2) * (qx * qx + qy * qy)) * n2;
    rt0 *= sx * sx; rt1 *= sy * sy; rt2 *= sz * sz;
    T v0 = (T(1) - T(2) * (qy * qy + qz * qz)) * rt0 + T(2) * (qx * qy - qw * qz) * rt1 + T(2) * (qx * qz + qw * qy) * rt2;
    T v1 = T(2) * (qx * qy + qw * qz) * rt0 + (T(1) - T(2) * (qx * qx + qz * qz)) * rt1 + T(2) * (qy * qz - qw * qx) * rt2;
    T v2 = T(2) * (qx * qz - qw * qy) * rt0 + T(2) * (qy * qz + qw * qx) * rt1 + (T(1) - T(2) * (qx * qx + qy * qy)) * rt2;
    if (!t_finite(v0) || !t_finite(v1) || !t_finite(v2)) return;
    T len2 = v0 * v0 + v1 * v1 + v2 * v2;
    if (!t_finite(len2)) return;
    if (len2 > max_disp * max_disp) {
        T inv = max_disp / std::sqrt(len2);
        v0 *= inv; v1 *= inv; v2 *= inv;
    }
    mean[0] += v0; mean[1] += v1; mean[2] += v2;
}

// mcmc_reg_grad_kernel, metal:3987-4026: gradients of the regulariser
//   duty * (op_reg * mean_j sigmoid(o_j) + sc_reg * mean_{j,k} exp(clamp(s_jk, -15, 15)))
// added to Gaussian i's opacity-logit and log-scale gradients (the mean over the N active
// rows, and over 3N scale entries). duty is 1 for a row whose DC second moments sum to 0
// (never visible yet) and reg_invis_scale otherwise. The clamp's own derivative is not
// applied (exp of the clamped value, as the reference writes it).
// Pin: FP64 central differences of that objective (tests/test_oracle_pins.py).
template <class T>
inline void reg_grad(int N, T opac_logit, const T ls[3], const T dc_sq[3], T op_reg, T sc_reg, T reg_invis_scale,
                     T& v_opac, T v_scale[3]) {
    T vdc = dc_sq[0] + dc_sq[1] + dc_sq[2];
    T duty = (vdc == T(0)) ? T(1) : reg_invis_scale;
    T invN = duty / T((float)N);
    T sig = T(1) / (T(1) + t_exp<T>(-opac_logit));
    v_opac += op_reg * sig * (T(1) - sig) * invN;
    T inv3N = invN * T(1.0f / 3.0f);
    for (int k = 0; k < 3; k++) v_scale[k] += sc_reg * t_exp<T>(t_clamp(ls[k], T(-15), T(15))) * inv3N;
}

// ---------------------------------------------------------------------------------------
// One full training step (msplat_train_step, mm:1708-1736, then the SGLD calls of
// model:971-974). State arrays are updated in place. Optional outputs are filled when the
// corresponding pointer in Debug is non-null.
template <class T> struct State {
    int N = 0, degree = 0;
    T *means, *scales, *quats, *f_dc, *f_rest, *opac;
    T* m[6];
    T* v[6];
    // SH storage in bfloat16 (FP32 only): the forward reads f_dc / f_rest rounded to bf16
    // (the device keeps a bf16 shadow of the fp32 master), and the Adam moments of groups
    // 3 and 4 are stored rounded to bf16 after every update. Masters stay fp32.
    int sh_bf16 = 0;
};

template <class T> inline T sh_store(T x, int bf16) { return x; }
template <> inline float sh_store<float>(float x, int bf16) { return bf16 ? ms_bf16_round(x) : x; }

template <class T> inline const T* sh_view(const State<T>& S, const T* src, size_t n, std::vector<T>& tmp) {
    if (!S.sh_bf16) return src;
    tmp.resize(n);
    for (size_t k = 0; k < n; k++) tmp[k] = sh_store<T>(src[k], 1);
    return tmp.data();
}

template <class T> struct StepOut {
    Proj<T> proj;
    Lists<T> lists;
    Raster<T> raster;
    std::vector<T> v_rend;
    RasterGrads<T> rg;
    std::vector<T> v_mean, v_scale, v_quat, v_opac_total;
    double loss_sum = 0.0;
    uint32_t nan_adam[6] = {0, 0, 0, 0, 0, 0};
};

template <class T>
void train_step(State<T>& S, const MsplatCamera& cam, const uint8_t* gt, const MsplatStepParams& hp, StepOut<T>& O,
                const uint8_t* mask = nullptr) {
    const int N = S.N, W = cam.width, H = cam.height;
    const int nb = num_sh_bases(S.degree) - 1;
    std::vector<T> dc_h, rest_h;
    const T* f_dc_fwd = sh_view(S, S.f_dc, 3 * (size_t)N, dc_h);
    const T* f_rest_fwd = sh_view(S, S.f_rest, (size_t)N * nb * 3, rest_h);
    project_forward<T>(N, S.means, S.scales, S.quats, f_dc_fwd, f_rest_fwd, S.opac, S.degree, hp.degrees_to_use, cam,
                       O.proj);
    build_lists<T>(N, O.proj, W, H, O.lists);
    uint32_t K = k_max_rule(O.lists.capacity, O.lists.num_tiles);
    raster_forward<T>(O.lists, W, H, hp.background, K, O.raster);
    O.loss_sum = ssim_loss<T>(W, H, O.raster.img.data(), gt, T(hp.ssim_weight), T(hp.loss_inv_n), O.v_rend, mask);
    raster_backward<T>(O.lists, O.raster, hp.background, O.v_rend.data(), N, O.rg);
    O.v_mean.assign(3 * (size_t)N, T(0));
    O.v_scale.assign(3 * (size_t)N, T(0));
    O.v_quat.assign(4 * (size_t)N, T(0));
    O.v_opac_total.assign((size_t)N, T(0));
    const T b1 = hp.beta1, b2 = hp.beta2, eps = hp.eps;
    uint32_t nan_cnt[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(static) reduction(+ : nan_cnt[:6])
    for (int i = 0; i < N; i++) {
        T* p = S.means + 3 * (size_t)i;
        T* ls = S.scales + 3 * (size_t)i;
        T* q = S.quats + 4 * (size_t)i;
        T vm[3] = {0, 0, 0}, vs[3] = {0, 0, 0}, vq[4] = {0, 0, 0, 0};
        if (O.proj.radii[i] > 0) {
            proj_backward_one<T>(p, ls, q, cam, &O.proj.conics[3 * (size_t)i], &O.rg.v_xy[2 * (size_t)i],
                                 &O.rg.v_conic[3 * (size_t)i], vm, vs, vq);
            // fused SH backward + Adam (metal:1963-2041): visible Gaussians only
            T pc[3] = {p[0] - T(cam.cam_pos[0]), p[1] - T(cam.cam_pos[1]), p[2] - T(cam.cam_pos[2])};
            T inv = T(1) / std::sqrt(pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
            T vd[3] = {pc[0] * inv, pc[1] * inv, pc[2] * inv};
            const T* vc = &O.rg.v_rgb[3 * (size_t)i];
            if (t_finite(vc[0]) && t_finite(vc[1]) && t_finite(vc[2])) {
                for (int c = 0; c < 3; c++) {
                    size_t k = 3 * (size_t)i + c;
                    adam<T>(S.f_dc[k], S.m[3][k], S.v[3][k], T(SH_C0) * vc[c], T(hp.step_size[3]), b1, b2,
                            T(hp.bc2_sqrt[3]), eps);
                    S.m[3][k] = sh_store<T>(S.m[3][k], S.sh_bf16);
                    S.v[3][k] = sh_store<T>(S.v[3][k], S.sh_bf16);
                }
                if (hp.degrees_to_use >= 1) {
                    T bs[15];
                    sh_basis<T>(vd, bs);
                    int nuse = hp.degrees_to_use >= 3 ? 15 : hp.degrees_to_use == 2 ? 8 : 3;
                    for (int b = 0; b < nuse; b++)
                        for (int c = 0; c < 3; c++) {
                            size_t k = (size_t)nb * 3 * i + b * 3 + c;
                            adam<T>(S.f_rest[k], S.m[4][k], S.v[4][k], bs[b] * vc[c], T(hp.step_size[4]), b1, b2,
                                    T(hp.bc2_sqrt[4]), eps);
                            S.m[4][k] = sh_store<T>(S.m[4][k], S.sh_bf16);
                            S.v[4][k] = sh_store<T>(S.v[4][k], S.sh_bf16);
                        }
                }
            } else {
                nan_cnt[3]++;
            }
        }
        T vo = O.rg.v_opac[i];
        // mcmc_reg_grad_kernel, metal:3987-4026 (dispatched when op_reg or sc_reg > 0)
        if (hp.op_reg > 0.0f || hp.sc_reg > 0.0f)
            reg_grad<T>(N, S.opac[i], ls, S.v[3] + 3 * (size_t)i, T(hp.op_reg), T(hp.sc_reg), T(hp.reg_invis_scale),
                        vo, vs);
        // gradients as fed to Adam (photometric + regulariser)
        for (int k = 0; k < 3; k++) { O.v_mean[3 * (size_t)i + k] = vm[k]; O.v_scale[3 * (size_t)i + k] = vs[k]; }
        for (int k = 0; k < 4; k++) O.v_quat[4 * (size_t)i + k] = vq[k];
        O.v_opac_total[i] = vo;
        // fused_adam_kernel x groups 0 (means), 1 (scales), 2 (quats), 5 (opacity), all N
        auto adam_g = [&](int grp, T& pp, T& mm, T& vv, T g) {
            if (!t_finite(g)) { nan_cnt[grp]++; return; }
            adam<T>(pp, mm, vv, g, T(hp.step_size[grp]), b1, b2, T(hp.bc2_sqrt[grp]), eps);
        };
        for (int k = 0; k < 3; k++) adam_g(0, p[k], S.m[0][3 * (size_t)i + k], S.v[0][3 * (size_t)i + k], vm[k]);
        for (int k = 0; k < 3; k++) adam_g(1, ls[k], S.m[1][3 * (size_t)i + k], S.v[1][3 * (size_t)i + k], vs[k]);
        for (int k = 0; k < 4; k++) adam_g(2, q[k], S.m[2][4 * (size_t)i + k], S.v[2][4 * (size_t)i + k], vq[k]);
        adam_g(5, S.opac[i], S.m[5][i], S.v[5][i], vo);
        // SGLD after the step's Adam (model:971-974)
        if (hp.sgld_enable) {
            T n[3];
            sgld_noise3<T>(hp.sgld_seed, (uint32_t)i, n);
            sgld_apply<T>(p, ls, q, S.opac[i], n, T(hp.noise_lr), T(hp.xyz_lr), T(hp.max_disp));
        }
    }
    for (int k = 0; k < 6; k++) O.nan_adam[k] = nan_cnt[k];
}

// Multi-view batched step (a feature of this build; the reference trains one view per
// step). Every view renders the same parameters; its raster, SSIM and projection gradients
// are summed over the views in view order (hp[v].loss_inv_n carries 1 / (3 H_v W_v B), so
// the objective is the mean of the per-view losses; with a training mask, 1 / (3 U_v B) for
// U_v unmasked pixels), and one update follows: Adam on the
// DC and rest coefficients of Gaussians that received a finite colour gradient in at least
// one view (bases up to degrees_to_use), the regulariser, Adam on means / scales / quats /
// opacity over all N, then SGLD. With one view it is train_step.
template <class T> struct BatchOut {
    std::vector<StepOut<T>> views;
    double loss_mean = 0.0;
    uint32_t nan_adam[6] = {0, 0, 0, 0, 0, 0};
};

template <class T>
void train_step_batch(State<T>& S, const MsplatCamera* cams, const uint8_t* const* gts, const MsplatStepParams* hps,
                      int B, BatchOut<T>& O, const uint8_t* const* masks = nullptr) {
    const int N = S.N;
    const int nb = num_sh_bases(S.degree) - 1;
    const MsplatStepParams& hp = hps[0];
    std::vector<T> dc_h, rest_h;
    const T* f_dc_fwd = sh_view(S, S.f_dc, 3 * (size_t)N, dc_h);
    const T* f_rest_fwd = sh_view(S, S.f_rest, (size_t)N * nb * 3, rest_h);
    std::vector<T> vm((size_t)N * 3, T(0)), vs((size_t)N * 3, T(0)), vq((size_t)N * 4, T(0)), vo((size_t)N, T(0));
    std::vector<T> vdc((size_t)N * 3, T(0)), vrest((size_t)N * nb * 3, T(0));
    std::vector<uint8_t> vis((size_t)N, 0);
    O.views.resize(B);
    O.loss_mean = 0.0;
    uint32_t nan3 = 0;
    const int nuse = hp.degrees_to_use >= 3 ? 15 : hp.degrees_to_use == 2 ? 8 : hp.degrees_to_use == 1 ? 3 : 0;
    for (int v = 0; v < B; v++) {
        StepOut<T>& V = O.views[v];
        const MsplatCamera& cam = cams[v];
        const int W = cam.width, H = cam.height;
        project_forward<T>(N, S.means, S.scales, S.quats, f_dc_fwd, f_rest_fwd, S.opac, S.degree, hp.degrees_to_use,
                           cam, V.proj);
        build_lists<T>(N, V.proj, W, H, V.lists);
        const uint32_t K = k_max_rule(V.lists.capacity, V.lists.num_tiles);
        raster_forward<T>(V.lists, W, H, hp.background, K, V.raster);
        V.loss_sum = ssim_loss<T>(W, H, V.raster.img.data(), gts[v], T(hp.ssim_weight), T(hps[v].loss_inv_n), V.v_rend,
                                  masks ? masks[v] : nullptr);
        raster_backward<T>(V.lists, V.raster, hp.background, V.v_rend.data(), N, V.rg);
        O.loss_mean += V.loss_sum / ((double)W * H) / B;
#pragma omp parallel for schedule(static) reduction(+ : nan3)
        for (int i = 0; i < N; i++) {
            if (V.proj.radii[i] <= 0) continue;
            const T* p = S.means + 3 * (size_t)i;
            const T* ls = S.scales + 3 * (size_t)i;
            const T* q = S.quats + 4 * (size_t)i;
            T m3[3] = {0, 0, 0}, s3[3] = {0, 0, 0}, q4[4] = {0, 0, 0, 0};
            proj_backward_one<T>(p, ls, q, cam, &V.proj.conics[3 * (size_t)i], &V.rg.v_xy[2 * (size_t)i],
                                 &V.rg.v_conic[3 * (size_t)i], m3, s3, q4);
            for (int k = 0; k < 3; k++) { vm[3 * (size_t)i + k] += m3[k]; vs[3 * (size_t)i + k] += s3[k]; }
            for (int k = 0; k < 4; k++) vq[4 * (size_t)i + k] += q4[k];
            vo[i] += V.rg.v_opac[i];
            const T* vc = &V.rg.v_rgb[3 * (size_t)i];
            if (!(t_finite(vc[0]) && t_finite(vc[1]) && t_finite(vc[2]))) { nan3++; continue; }
            vis[i] = 1;
            for (int c = 0; c < 3; c++) vdc[3 * (size_t)i + c] += T(SH_C0) * vc[c];
            if (nuse > 0) {
                T pc[3] = {p[0] - T(cam.cam_pos[0]), p[1] - T(cam.cam_pos[1]), p[2] - T(cam.cam_pos[2])};
                T inv = T(1) / std::sqrt(pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
                T vd[3] = {pc[0] * inv, pc[1] * inv, pc[2] * inv};
                T bs[15];
                sh_basis<T>(vd, bs);
                for (int b = 0; b < nuse; b++)
                    for (int c = 0; c < 3; c++) vrest[(size_t)nb * 3 * i + b * 3 + c] += bs[b] * vc[c];
            }
        }
    }
    const T b1 = hp.beta1, b2 = hp.beta2, eps = hp.eps;
    uint32_t nan_cnt[6] = {0, 0, 0, nan3, 0, 0};
#pragma omp parallel for schedule(static) reduction(+ : nan_cnt[:6])
    for (int i = 0; i < N; i++) {
        T* p = S.means + 3 * (size_t)i;
        T* ls = S.scales + 3 * (size_t)i;
        T* q = S.quats + 4 * (size_t)i;
        if (vis[i]) {
            for (int c = 0; c < 3; c++) {
                const size_t k = 3 * (size_t)i + c;
                adam<T>(S.f_dc[k], S.m[3][k], S.v[3][k], vdc[k], T(hp.step_size[3]), b1, b2, T(hp.bc2_sqrt[3]), eps);
                S.m[3][k] = sh_store<T>(S.m[3][k], S.sh_bf16);
                S.v[3][k] = sh_store<T>(S.v[3][k], S.sh_bf16);
            }
            for (int b = 0; b < nuse; b++)
                for (int c = 0; c < 3; c++) {
                    const size_t k = (size_t)nb * 3 * i + b * 3 + c;
                    adam<T>(S.f_rest[k], S.m[4][k], S.v[4][k], vrest[k], T(hp.step_size[4]), b1, b2,
                            T(hp.bc2_sqrt[4]), eps);
                    S.m[4][k] = sh_store<T>(S.m[4][k], S.sh_bf16);
                    S.v[4][k] = sh_store<T>(S.v[4][k], S.sh_bf16);
                }
        }
        T gvo = vo[i];
        T gvs[3] = {vs[3 * (size_t)i], vs[3 * (size_t)i + 1], vs[3 * (size_t)i + 2]};
        if (hp.op_reg > 0.0f || hp.sc_reg > 0.0f)
            reg_grad<T>(N, S.opac[i], ls, S.v[3] + 3 * (size_t)i, T(hp.op_reg), T(hp.sc_reg), T(hp.reg_invis_scale),
                        gvo, gvs);
        auto adam_g = [&](int grp, T& pp, T& mm, T& vv, T g) {
            if (!t_finite(g)) { nan_cnt[grp]++; return; }
            adam<T>(pp, mm, vv, g, T(hp.step_size[grp]), b1, b2, T(hp.bc2_sqrt[grp]), eps);
        };
        for (int k = 0; k < 3; k++) adam_g(0, p[k], S.m[0][3 * (size_t)i + k], S.v[0][3 * (size_t)i + k], vm[3 * (size_t)i + k]);
        for (int k = 0; k < 3; k++) adam_g(1, ls[k], S.m[1][3 * (size_t)i + k], S.v[1][3 * (size_t)i + k], gvs[k]);
        for (int k = 0; k < 4; k++) adam_g(2, q[k], S.m[2][4 * (size_t)i + k], S.v[2][4 * (size_t)i + k], vq[4 * (size_t)i + k]);
        adam_g(5, S.opac[i], S.m[5][i], S.v[5][i], gvo);
        if (hp.sgld_enable) {
            T n[3];
            sgld_noise3<T>(hp.sgld_seed, (uint32_t)i, n);
            sgld_apply<T>(p, ls, q, S.opac[i], n, T(hp.noise_lr), T(hp.xyz_lr), T(hp.max_disp));
        }
    }
    for (int k = 0; k < 6; k++) O.nan_adam[k] = nan_cnt[k];
}

// Forward only (msplat_render, mm:1099-1116)
template <class T>
void render(const State<T>& S, const MsplatCamera& cam, int degrees_to_use, const float bg[3], StepOut<T>& O) {
    std::vector<T> dc_h, rest_h;
    const int nb = num_sh_bases(S.degree) - 1;
    const T* f_dc_fwd = sh_view(S, S.f_dc, 3 * (size_t)S.N, dc_h);
    const T* f_rest_fwd = sh_view(S, S.f_rest, (size_t)S.N * nb * 3, rest_h);
    project_forward<T>(S.N, S.means, S.scales, S.quats, f_dc_fwd, f_rest_fwd, S.opac, S.degree, degrees_to_use, cam,
                       O.proj);
    build_lists<T>(S.N, O.proj, cam.width, cam.height, O.lists);
    uint32_t K = k_max_rule(O.lists.capacity, O.lists.num_tiles);
    raster_forward<T>(O.lists, cam.width, cam.height, bg, K, O.raster);
}

}  // namespace oracle

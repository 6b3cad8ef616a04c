// kernels.hip — gfx950 kernels of the training step.
//
// One wave = 64 lanes. Per-Gaussian kernels stream AoS parameter rows; tile kernels run one
// 256-thread workgroup (4 waves) per 16x16 binning tile so every tile list is read from
// HBM once per pass and broadcast from LDS; per-entry gradient sums use DPP wave
// reductions and one global atomic per (tile, entry). Float arithmetic of every stage that
// feeds an integer output (projection -> tile rectangles / sort keys, compositing ->
// contributor indices) follows the reference's operation order exactly, in IEEE-754
// single precision without contraction, so those integers match a CPU evaluation of the
// same formulas bit for bit.
//
// Reference kernels are cited as metal:<line> (core/metal/msplat_metal.metal).
#include <atomic>
#include <climits>
#include <stdexcept>
#include <string>
#include <utility>

#include "kernels.h"
#include "ms_math.h"
#include "lane_ops.h"

namespace msk {

__constant__ float c_gauss1d[11] = {0.0010283801f, 0.0075987581f, 0.0360007721f, 0.1093606895f,
                                    0.2130055377f, 0.2660117249f, 0.2130055377f, 0.1093606895f,
                                    0.0360007721f, 0.0075987581f, 0.0010283801f};

#define SH_C0 0.28209479177387814f
#define SH_C1 0.4886025119029199f
#define SH_C2_0 1.0925484305920792f
#define SH_C2_1 (-1.0925484305920792f)
#define SH_C2_2 0.31539156525252005f
#define SH_C2_3 (-1.0925484305920792f)
#define SH_C2_4 0.5462742152960396f
#define SH_C3_0 (-0.5900435899266435f)
#define SH_C3_1 2.890611442640554f
#define SH_C3_2 (-0.4570457994644658f)
#define SH_C3_3 0.3731763325901154f
#define SH_C3_4 (-0.4570457994644658f)
#define SH_C3_5 1.445305721320277f
#define SH_C3_6 (-0.5900435899266435f)

static inline unsigned cdiv(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------------------------------
// Wave64 helpers.
__device__ __forceinline__ float dpp_row_shr1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_row_shr2(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_row_shr4(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_row_shr8(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_bcast15(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xA, 0xF, false));
}
__device__ __forceinline__ float dpp_bcast31(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xC, 0xF, false));
}
// Inclusive wave sum: the total ends in lane 63.
__device__ __forceinline__ float wave_sum_to_lane63(float v) {
    v += dpp_row_shr1(v);
    v += dpp_row_shr2(v);
    v += dpp_row_shr4(v);
    v += dpp_row_shr8(v);
    v += dpp_bcast15(v);
    v += dpp_bcast31(v);
    return v;
}

// Sums of four 64-lane vectors a, b, c, d at once: half-waves swapped and added
// (v_permlane32_swap), then rows (v_permlane16_swap), then within each 16-lane row: ten
// instructions for four sums instead of 24. Every swap trades whole halves / rows, so each
// row ends up holding one vector's partial sums; lane 15 of the row has that vector's total,
// and wave_sum4_tag() (the same swaps applied to the tags 0, 1, 2, 3) says which vector it is.
__device__ __forceinline__ int wave_sum4_tag() {
    float a = 0.f, b = 1.f, c = 2.f, d = 3.f;
    permlane32_swap(a, b);
    permlane32_swap(c, d);
    permlane16_swap(a, c);
    return (int)a;
}

__device__ __forceinline__ float wave_sum4_rows(float a, float b, float c, float d) {
    permlane32_swap(a, b);
    const float ab = a + b;  // one vector per half-wave
    permlane32_swap(c, d);
    float cd = c + d;
    float abr = ab;
    permlane16_swap(abr, cd);
    float v = abr + cd;      // one vector per row
    v += dpp_row_shr1(v);
    v += dpp_row_shr2(v);
    v += dpp_row_shr4(v);
    v += dpp_row_shr8(v);
    return v;
}

// Float atomic add to a device (global) address, issued as a global_atomic_add_f32 rather
// than the flat atomic an unqualified pointer gets (the flat form also waits on the LDS
// counter and checks the address aperture). Relaxed, device scope, no return value.
typedef __attribute__((address_space(1))) float global_float;
__device__ __forceinline__ void gadd(float* p, float v) {
    __hip_atomic_fetch_add((global_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 1 / d for d in [0.001, 1] (1 - alpha in the raster backward): v_rcp_f32 (within 1 ulp)
// refined by one Newton step, instead of the correctly rounded division sequence (about ten
// instructions). The backward's gradients are compared at 1e-4, not bit for bit.
__device__ __forceinline__ float recip_nr(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, r, 1.0f), r, r);
}

// Block-wide exclusive scan of uint32 over blockDim.x (multiple of 64, <= 1024) threads.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_wave, uint32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds_wave[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; w++) { uint32_t t = lds_wave[w]; lds_wave[w] = acc; acc += t; }
        lds_wave[nw] = acc;
    }
    __syncthreads();
    uint32_t r = x - v + lds_wave[wid];
    *total = lds_wave[nw];
    __syncthreads();
    return r;
}

// Next work item of the calling workgroup from a queue head in global memory (all threads
// of the workgroup call it; the value is the same for all).
__device__ __forceinline__ uint32_t next_item(uint32_t* ctr) {
    __shared__ uint32_t s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(ctr, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    __syncthreads();
    return it;
}

__global__ void k_fill_zero(uint32_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}

// Up to four ranges in one launch (blockIdx.y picks the range).
struct ZeroRanges {
    uint32_t* p[4];
    size_t n[4];
};
__global__ void k_fill_zero4(ZeroRanges r) {
    uint32_t* p = r.p[blockIdx.y];
    const size_t n = r.n[blockIdx.y];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}

void fill_zero4(hipStream_t s, void* p0, size_t b0, void* p1, size_t b1, void* p2, size_t b2, void* p3, size_t b3) {
    ZeroRanges r{{(uint32_t*)p0, (uint32_t*)p1, (uint32_t*)p2, (uint32_t*)p3}, {b0 / 4, b1 / 4, b2 / 4, b3 / 4}};
    size_t most = 0;
    for (int k = 0; k < 4; k++) {
        if (!r.p[k]) r.n[k] = 0;
        most = std::max(most, r.n[k]);
    }
    if (most == 0) return;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(2048, (most + 255) / 256));
    hipLaunchKernelGGL(k_fill_zero4, dim3(blocks, 4), dim3(256), 0, s, r);
}

void fill_zero(hipStream_t s, void* p, size_t bytes) {
    if (!p || bytes == 0) return;
    size_t n = bytes / 4;
    unsigned blocks = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, s, (uint32_t*)p, n);
}

// Plain 16-byte copy: the reference point that HBM traffic counters are scaled against.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

void copy16(hipStream_t s, const void* src, void* dst, size_t bytes) {
    const size_t n = bytes / 16;
    if (n == 0) return;
    const unsigned blocks = (unsigned)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
}

// ------------------------------------------------------------------------------------
// A5: project_and_sh_forward (metal:1719-1834). One thread per Gaussian; the block's
// features_rest rows are staged through LDS with 16-byte loads (rows are 180 B apart).
constexpr int PROJ_BLOCK = 128;

__device__ __forceinline__ void quat_rot(const float q[4], float R[3][3]) {
    // quat_to_rotmat, metal:134-156 (R[r][c] row-major)
    float n2 = q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    float s = 1.0f / sqrtf(fmaxf(n2, 1e-8f));
    float w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
    R[0][0] = 1.f - 2.f * (y * y + z * z);
    R[1][0] = 2.f * (x * y + w * z);
    R[2][0] = 2.f * (x * z - w * y);
    R[0][1] = 2.f * (x * y - w * z);
    R[1][1] = 1.f - 2.f * (x * x + z * z);
    R[2][1] = 2.f * (y * z + w * x);
    R[0][2] = 2.f * (x * z + w * y);
    R[1][2] = 2.f * (y * z - w * x);
    R[2][2] = 1.f - 2.f * (x * x + y * y);
}

__device__ __forceinline__ void cov3d_of(const float s[3], const float q[4], float cov[6]) {
    float R[3][3];
    quat_rot(q, R);
    float M[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) M[r][c] = R[r][c] * s[c];
    cov[0] = M[0][0] * M[0][0] + M[0][1] * M[0][1] + M[0][2] * M[0][2];
    cov[1] = M[1][0] * M[0][0] + M[1][1] * M[0][1] + M[1][2] * M[0][2];
    cov[2] = M[2][0] * M[0][0] + M[2][1] * M[0][1] + M[2][2] * M[0][2];
    cov[3] = M[1][0] * M[1][0] + M[1][1] * M[1][1] + M[1][2] * M[1][2];
    cov[4] = M[2][0] * M[1][0] + M[2][1] * M[1][1] + M[2][2] * M[1][2];
    cov[5] = M[2][0] * M[2][0] + M[2][1] * M[2][1] + M[2][2] * M[2][2];
}

__device__ __forceinline__ float sh_ld(const float* p, int k) { return p[k]; }
__device__ __forceinline__ float sh_ld(const uint16_t* p, int k) { return ms_bf16_float(p[k]); }

// SHT = float (fp32 SH) or uint16_t (bf16 shadows of the fp32 masters).
template <typename SHT>
__global__ __launch_bounds__(PROJ_BLOCK) void k_project(
    int N, const float* __restrict__ means, const float* __restrict__ scales, const float* __restrict__ quats,
    const SHT* __restrict__ f_dc, const SHT* __restrict__ f_rest, const float* __restrict__ opac, int nb3,
    int deg_use, const ViewConsts* __restrict__ vc, int32_t* __restrict__ radii_o, int32_t* __restrict__ tiles_o,
    float* __restrict__ depth_o, ushort4* __restrict__ rect_o, float2* __restrict__ aabb_o, float4* __restrict__ rA,
    float4* __restrict__ rB, float2* __restrict__ rC, uint32_t* __restrict__ nanc, GradBufs zero) {
    extern __shared__ uint4 s_raw[];  // PROJ_BLOCK * nb3 coefficients when deg_use > 0
    SHT* s_rest = reinterpret_cast<SHT*>(s_raw);
    const int i0 = blockIdx.x * PROJ_BLOCK;
    const int i = i0 + threadIdx.x;
    if (i == 0) {  // row N: the transparent record that list entries past the last Gaussian read
        rA[N] = make_float4(0.f, 0.f, 0.f, 0.f);
        rB[N] = make_float4(0.f, 0.f, 0.f, 0.f);
        rC[N] = make_float2(0.f, 0.f);
    }
    if (deg_use > 0 && nb3 > 0) {
        // the block's rows are contiguous and start 16-byte aligned (i0 * nb3 * sizeof(SHT)
        // is a multiple of 16 for PROJ_BLOCK = 128)
        const int rows = min(PROJ_BLOCK, N - i0);
        const int nbytes = rows * nb3 * (int)sizeof(SHT);
        const uint8_t* src = reinterpret_cast<const uint8_t*>(f_rest + (size_t)i0 * nb3);
        const int n16 = nbytes >> 4;
        const uint4* src16 = reinterpret_cast<const uint4*>(src);
        for (int k = threadIdx.x; k < n16; k += PROJ_BLOCK) s_raw[k] = src16[k];
        const uint16_t* src2 = reinterpret_cast<const uint16_t*>(src);
        uint16_t* dst2 = reinterpret_cast<uint16_t*>(s_raw);
        for (int k = (n16 << 3) + threadIdx.x; k < (nbytes >> 1); k += PROJ_BLOCK) dst2[k] = src2[k];
        __syncthreads();
    }
    if (i >= N) return;
    if (zero.g) {  // a training step: the raster backward's accumulators start at zero
#pragma unroll
        for (int k = 0; k < G2D; k++) zero.g[G2D * (size_t)i + k] = 0.f;
    }
    const MsplatCamera cam = vc->cam;
    const int tbx = vc->tbx, tby = vc->tby;

    int32_t rad = 0, area = 0;
    float depth = 0.f;
    ushort4 rect = make_ushort4(0, 0, 0, 0);
    float2 aabb = make_float2(0.f, 0.f);
    float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 C = make_float2(0.f, __builtin_bit_cast(float, (uint32_t)i));

    const float p0 = means[3 * (size_t)i], p1 = means[3 * (size_t)i + 1], p2 = means[3 * (size_t)i + 2];
    const float* vm = cam.viewmat;
    const float pvz = vm[8] * p0 + vm[9] * p1 + vm[10] * p2 + vm[11];
    if (pvz > 0.01f) {  // clip_near_plane (metal:159-170), clip_thresh 0.01
        const float pvx = vm[0] * p0 + vm[1] * p1 + vm[2] * p2 + vm[3];
        const float pvy = vm[4] * p0 + vm[5] * p1 + vm[6] * p2 + vm[7];
        float s[3], q[4];
#pragma unroll
        for (int k = 0; k < 3; k++) s[k] = ms_expf_core(fminf(fmaxf(scales[3 * (size_t)i + k], -15.f), 15.f));
        const float4 q4 = reinterpret_cast<const float4*>(quats)[i];
        q[0] = q4.x; q[1] = q4.y; q[2] = q4.z; q[3] = q4.w;
        float cov[6];
        cov3d_of(s, q, cov);
        // project_cov3d_ewa (metal:266-306)
        const float W = (float)cam.width, H = (float)cam.height;
        const float fx = cam.fx, fy = cam.fy;
        const float tan_fovx = 0.5f * W / fx, tan_fovy = 0.5f * H / fy;
        const float lim_x = 1.3f * tan_fovx, lim_y = 1.3f * tan_fovy;
        const float px = pvz * fminf(lim_x, fmaxf(-lim_x, pvx / pvz));
        const float py = pvz * fminf(lim_y, fmaxf(-lim_y, pvy / pvz));
        const float rz = 1.f / pvz;
        const float rz2 = rz * rz;
        const float j00 = fx * rz, j11 = fy * rz;
        const float j20 = -fx * px * rz2, j21 = -fy * py * rz2;
        float t0[3], t1[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            t0[k] = j00 * vm[0 + k] + j20 * vm[8 + k];
            t1[k] = j11 * vm[4 + k] + j21 * vm[8 + k];
        }
        const float v00 = cov[0], v01 = cov[1], v02 = cov[2], v11 = cov[3], v12 = cov[4], v22 = cov[5];
        const float tv00 = t0[0] * v00 + t0[1] * v01 + t0[2] * v02;
        const float tv01 = t0[0] * v01 + t0[1] * v11 + t0[2] * v12;
        const float tv02 = t0[0] * v02 + t0[1] * v12 + t0[2] * v22;
        const float tv10 = t1[0] * v00 + t1[1] * v01 + t1[2] * v02;
        const float tv11 = t1[0] * v01 + t1[1] * v11 + t1[2] * v12;
        const float tv12 = t1[0] * v02 + t1[1] * v12 + t1[2] * v22;
        const float c0 = (tv00 * t0[0] + tv01 * t0[1] + tv02 * t0[2]) + 0.3f;
        const float c1 = tv00 * t1[0] + tv01 * t1[1] + tv02 * t1[2];
        const float c2 = (tv10 * t1[0] + tv11 * t1[1] + tv12 * t1[2]) + 0.3f;
        // compute_cov2d_bounds (metal:308-332)
        const float det = c0 * c2 - c1 * c1;
        if (det > 1e-6f) {
            const float inv_det = 1.f / det;
            const float cn0 = c2 * inv_det, cn1 = -c1 * inv_det, cn2 = c0 * inv_det;
            const float b = 0.5f * (c0 + c2);
            const float disc = sqrtf(fmaxf(0.1f, b * b - det));
            const float radius = ceilf(3.f * sqrtf(b + disc));
            // project_pix (metal:335-344)
            const float* pm = cam.projmat;
            const float hx = pm[0] * p0 + pm[1] * p1 + pm[2] * p2 + pm[3];
            const float hy = pm[4] * p0 + pm[5] * p1 + pm[6] * p2 + pm[7];
            const float hw = pm[12] * p0 + pm[13] * p1 + pm[14] * p2 + pm[15];
            const float rw = 1.f / (hw + 1e-6f);
            const float ndx = hx * rw, ndy = hy * rw;
            const float cxp = 0.5f * W * ndx + cam.cx - 0.5f;
            const float cyp = 0.5f * H * ndy + cam.cy - 0.5f;
            const float ax = ceilf(3.f * sqrtf(c0));
            const float ay = ceilf(3.f * sqrtf(c2));
            // get_tile_bbox (metal:57-87)
            const float tcx = cxp / 16.f, tcy = cyp / 16.f, trx = ax / 16.f, try_ = ay / 16.f;
            const int x0 = min(max(0, ms_f2i(tcx - trx)), tbx);
            const int x1 = min(max(0, ms_f2i(tcx + trx + 1.f)), tbx);
            const int y0 = min(max(0, ms_f2i(tcy - try_)), tby);
            const int y1 = min(max(0, ms_f2i(tcy + try_ + 1.f)), tby);
            const int a = (int)((uint32_t)(x1 - x0) * (uint32_t)(y1 - y0));
            if (a > 0) {
                // SH colour (metal:1805-1811; sh_coeffs_to_color metal:620-687)
                const float d0 = p0 - cam.cam_pos[0], d1 = p1 - cam.cam_pos[1], d2 = p2 - cam.cam_pos[2];
                const float inv = 1.f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
                const float x = d0 * inv, y = d1 * inv, z = d2 * inv;
                float col[3];
#pragma unroll
                for (int c = 0; c < 3; c++) col[c] = SH_C0 * sh_ld(f_dc + 3 * (size_t)i, c);
                if (deg_use >= 1) {
                    const SHT* rs = s_rest + threadIdx.x * nb3;
                    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        col[c] += SH_C1 * (-y * sh_ld(rs, 0 * 3 + c) + z * sh_ld(rs, 1 * 3 + c) - x * sh_ld(rs, 2 * 3 + c));
                        if (deg_use < 2) continue;
                        col[c] += (SH_C2_0 * xy * sh_ld(rs, 3 * 3 + c) + SH_C2_1 * yz * sh_ld(rs, 4 * 3 + c) +
                                   SH_C2_2 * (2.f * zz - xx - yy) * sh_ld(rs, 5 * 3 + c) + SH_C2_3 * xz * sh_ld(rs, 6 * 3 + c) +
                                   SH_C2_4 * (xx - yy) * sh_ld(rs, 7 * 3 + c));
                        if (deg_use < 3) continue;
                        col[c] += (SH_C3_0 * y * (3.f * xx - yy) * sh_ld(rs, 8 * 3 + c) + SH_C3_1 * xy * z * sh_ld(rs, 9 * 3 + c) +
                                   SH_C3_2 * y * (4.f * zz - xx - yy) * sh_ld(rs, 10 * 3 + c) +
                                   SH_C3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy) * sh_ld(rs, 11 * 3 + c) +
                                   SH_C3_4 * x * (4.f * zz - xx - yy) * sh_ld(rs, 12 * 3 + c) +
                                   SH_C3_5 * z * (xx - yy) * sh_ld(rs, 13 * 3 + c) +
                                   SH_C3_6 * x * (xx - 3.f * yy) * sh_ld(rs, 14 * 3 + c));
                    }
                }
                // zombie gate (metal:1818-1833)
                const bool ok_center = ms_finite(cxp) && ms_finite(cyp);
                const bool ok_depth = ms_finite(pvz);
                const bool ok_aabb = ms_finite(ax) && ms_finite(ay);
                const bool ok_conic = ms_finite(cn0) && ms_finite(cn1) && ms_finite(cn2);
                const bool ok_rgb = ms_finite(col[0]) && ms_finite(col[1]) && ms_finite(col[2]);
                if (ok_center && ok_depth && ok_aabb && ok_conic && ok_rgb) {
                    rad = ms_f2i(radius);
                    area = a;
                    depth = pvz;
                    // only listed Gaussians (radius > 0) keep a non-empty rectangle
                    if (rad > 0)
                        rect = make_ushort4((unsigned short)x0, (unsigned short)y0, (unsigned short)x1, (unsigned short)y1);
                    aabb = make_float2(ax, ay);
                    const float op = 1.f / (1.f + ms_expf(-opac[i]));  // pack-time sigmoid, metal:2563
                    if (ms_finite(op)) {
                        A = make_float4(cxp, cyp, op, cn0);
                        B = make_float4(cn1, cn2, col[0], col[1]);
                        C.x = col[2];
                    } else {
                        atomicAdd(&nanc[13], 1u);  // pack gate: transparent entry (metal:2570-2578)
                        atomicAdd(&nanc[14], 1u);
                    }
                } else {
                    atomicAdd(&nanc[7], 1u);
                    if (!ok_center) atomicAdd(&nanc[8], 1u);
                    if (!ok_depth) atomicAdd(&nanc[9], 1u);
                    if (!ok_aabb) atomicAdd(&nanc[10], 1u);
                    if (!ok_conic) atomicAdd(&nanc[11], 1u);
                    if (!ok_rgb) atomicAdd(&nanc[12], 1u);
                }
            }
        }
    }
    radii_o[i] = rad;
    tiles_o[i] = area;
    depth_o[i] = depth;
    rect_o[i] = rect;
    aabb_o[i] = aabb;
    rA[i] = A;
    rB[i] = B;
    rC[i] = C;
}

void project_forward(hipStream_t s, const Params& P, const ViewConsts* vc, int degrees_to_use, ProjOut& o,
                     uint32_t* nan_counts, const GradBufs* zero) {
    const GradBufs z = zero ? *zero : GradBufs{nullptr};
    if (P.N <= 0) {  // no Gaussians: only the transparent record (row 0)
        fill_zero4(s, o.rec.A, 16, o.rec.B, 16, o.rec.C, 8, nullptr, 0);
        return;
    }
    const int nb3 = (P.degree <= 0 ? 0 : (P.degree == 1 ? 3 : P.degree == 2 ? 8 : 15)) * 3;
    const int du = degrees_to_use < P.degree ? degrees_to_use : P.degree;
    const size_t esz = P.sh_bf16 ? 2 : 4;
    size_t lds = (du > 0 && nb3 > 0) ? (size_t)PROJ_BLOCK * nb3 * esz : 0;
    if (P.sh_bf16)
        hipLaunchKernelGGL(k_project<uint16_t>, dim3(cdiv(P.N, PROJ_BLOCK)), dim3(PROJ_BLOCK), lds, s, P.N, P.means,
                           P.scales, P.quats, P.f_dc_h, P.f_rest_h, P.opac, nb3, du, vc, o.radii, o.tiles_hit,
                           o.depths, o.rect, o.aabb, o.rec.A, o.rec.B, o.rec.C, nan_counts, z);
    else
        hipLaunchKernelGGL(k_project<float>, dim3(cdiv(P.N, PROJ_BLOCK)), dim3(PROJ_BLOCK), lds, s, P.N, P.means,
                           P.scales, P.quats, P.f_dc, P.f_rest, P.opac, nb3, du, vc, o.radii, o.tiles_hit, o.depths,
                           o.rect, o.aabb, o.rec.A, o.rec.B, o.rec.C, nan_counts, z);
}

Geom geom_of(int W, int H) {
    Geom g;
    g.W = W;
    g.H = H;
    g.tbx = (W + TILE - 1) / TILE;
    g.tby = (H + TILE - 1) / TILE;
    g.T = g.tbx * g.tby;
    return g;
}

// ------------------------------------------------------------------------------------
// A6-A9: per-tile lists sorted by (depth, id). The reference counts per-tile
// intersections (metal:2342-2365), scans them (metal:2376-2435), scatters 64-bit
// (depth bits << 32 | id) keys into padded buckets (metal:2440-2480) and bitonic-sorts each
// bucket with inline packing (metal:2489-2627). The same lists are built here as follows.
//  1. One stable radix sort of the N Gaussians by depth bits gives the global (depth, id)
//     order; r is a Gaussian's rank in it.
//  2. The ranks are cut into G-wide chunks, one workgroup each. A workgroup counts its
//     chunk's tile hits in an LDS histogram (no global atomics), one row of a chunk x tile
//     matrix; column sums give the tile counts, and one scan the tiles' list offsets.
//  3. Level 1: the same chunks list every Gaussian once per 4x4-tile block its rectangle
//     meets, with the mask of the block's tiles it covers, in per-block lists: a block
//     histogram and column scan give every (chunk, block) pair its offset, and the bucket
//     kernel writes each chunk's pairs there, walking the chunk's ranks in increasing order
//     (per wave, after a per-wave count), so each block list is in rank order.
//  4. Level 2: one workgroup per tile keeps, in order, the entries of its block's list
//     whose mask has the tile's bit, and writes them front to back as the tile's list:
//     sorted by rank, with no sort pass, and every line of the list written by one
//     workgroup (writing each hit straight into its tile's list from its chunk scatters
//     4-byte stores over the whole list buffer; the L2 then writes back partial lines,
//     about 8x the list's bytes).
// The resulting per-tile id lists are the reference's sorted lists exactly.

__device__ __forceinline__ uint32_t pad_count(uint32_t c) {  // metal:2248-2256
    if (c == 0u) return 0u;
    uint32_t x = c < 32u ? 32u : c;
    return 1u << (32 - __clz((int)(x - 1u)));
}

constexpr int HIST_THREADS = 512;
constexpr int MAX_CHUNK = 1984;  // Gaussians per chunk at most (a chunk-local rank fits 16 bits)

// srect [N]: tile rectangles in depth-rank order (empty for invisible Gaussians), so the
// chunk kernels read them contiguously.
// mat [chunks][T1]: row stride T1 = the launch geometry's tile count.
__global__ __launch_bounds__(HIST_THREADS) void k_tile_hist(int N, int G, int T1, const ushort4* __restrict__ srect,
                                                            const ViewConsts* __restrict__ vc,
                                                            uint32_t* __restrict__ mat) {
    extern __shared__ uint32_t hist[];
    const int tbx = vc->tbx, T = vc->T;
    for (int t = threadIdx.x; t < T; t += HIST_THREADS) hist[t] = 0u;
    __syncthreads();
    const int r0 = blockIdx.x * G, r1 = min(N, r0 + G);
    for (int r = r0 + threadIdx.x; r < r1; r += HIST_THREADS) {
        const ushort4 q = srect[r];
        for (int y = q.y; y < q.w; y++)
            for (int x = q.x; x < q.z; x++) atomicAdd(&hist[y * tbx + x], 1u);
    }
    __syncthreads();
    uint32_t* row = mat + (size_t)blockIdx.x * T1;
    for (int t = threadIdx.x; t < T; t += HIST_THREADS) row[t] = hist[t];
}

// Column sums of the chunk x tile matrix: the tile counts. A workgroup takes 64 columns; its
// 8 waves each own one eighth of the chunks (partial sums combined through LDS), so the chain
// of dependent loads per lane is an eighth of the column.
constexpr int COLSCAN_PARTS = 8;
__global__ __launch_bounds__(64 * COLSCAN_PARTS) void k_col_sum(const ViewConsts* __restrict__ vc, int nblk, int T1,
                                                                const uint32_t* __restrict__ mat,
                                                                uint32_t* __restrict__ counts) {
    __shared__ uint32_t part_sum[COLSCAN_PARTS][64];
    const int T = vc->T;
    const int c = (int)(threadIdx.x & 63), part = (int)(threadIdx.x >> 6);
    const int t = (int)blockIdx.x * 64 + c;
    const int per = (nblk + COLSCAN_PARTS - 1) / COLSCAN_PARTS;
    const int b0 = min(nblk, part * per), b1 = min(nblk, b0 + per);
    uint32_t sum = 0;
    if (t < T) {
        int b = b0;
        for (; b + 8 <= b1; b += 8) {  // eight independent loads in flight
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = mat[(size_t)(b + u) * T1 + t];
#pragma unroll
            for (int u = 0; u < 8; u++) sum += v[u];
        }
        for (; b < b1; b++) sum += mat[(size_t)b * T1 + t];
    }
    part_sum[part][c] = sum;
    __syncthreads();
    if (part != 0 || t >= T) return;
    uint32_t total = 0;
    for (int q = 0; q < COLSCAN_PARTS; q++) total += part_sum[q][c];
    counts[t] = total;
}

// Work items of the persistent raster kernels: (view << 20) | tile, ordered by decreasing
// tile list length so that the longest tiles start first and the launch ends on short ones.
constexpr int WORK_TILE_BITS = 20;
constexpr int WQ_THREADS = 1024, WQ_BINS = 256;
// Queue bin of a tile list of `len` entries: 8-entry bins, longest first, all lists of 2040
// entries or more in the first bin.
__device__ __forceinline__ int wq_bin(uint32_t len) {
    return WQ_BINS - 1 - (int)min(len >> 3, (uint32_t)(WQ_BINS - 1));
}

// queue != nullptr (a single unchunked view): the workgroup also writes the raster work
// queue of view 0 from the bins it has just written (what k_work_queue does for a batch of
// views), so the single-view step needs no launch of its own for it.
__global__ __launch_bounds__(1024) void k_prefix(const ViewConsts* __restrict__ vc, const uint32_t* __restrict__ counts,
                                                 uint32_t* __restrict__ exact_off, int2* __restrict__ bins,
                                                 uint32_t* __restrict__ totals, uint32_t cap,
                                                 uint32_t* __restrict__ queue, uint32_t* __restrict__ n_work,
                                                 uint32_t* __restrict__ ctr) {
    __shared__ uint32_t lw[17];
    const int T = vc->T;
    constexpr int PER = 8;  // consecutive tiles per thread: one block scan per 8192 tiles
    uint32_t run_ex = 0, my_pd = 0, mx = 0;
    for (int c0 = 0; c0 < T; c0 += 1024 * PER) {
        const int t0 = c0 + (int)threadIdx.x * PER;
        uint32_t v[PER], sum = 0u;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            v[q] = t0 + q < T ? counts[t0 + q] : 0u;
            sum += v[q];
            my_pd += pad_count(v[q]);
            mx = max(mx, v[q]);
        }
        uint32_t tot_e;
        uint32_t run = run_ex + block_exclusive_scan(sum, lw, &tot_e);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = t0 + q;
            if (t < T) {
                exact_off[t] = run;
                // never past the list buffers (the reference scatter drops overflow, metal:2474-2477)
                bins[t] = make_int2((int)min(run, cap), (int)min(run + v[q], cap));
            }
            run += v[q];
        }
        run_ex += tot_e;
    }
    __shared__ uint32_t smax, spd;
    if (threadIdx.x == 0) { smax = 0; spd = 0; }
    __syncthreads();
    atomicMax(&smax, mx);
    atomicAdd(&spd, my_pd);
    __syncthreads();
    if (threadIdx.x == 0) {
        totals[0] = run_ex;
        totals[1] = spd;      // what the reference's padded buckets would occupy
        totals[2] = run_ex > cap ? 1u : 0u;
        totals[3] = smax;
    }
    if (queue == nullptr) return;
    // the work queue: tiles binned by list length (the bins written above are visible to
    // the workgroup after its barriers), bin starts by a scan, tiles scattered to them
    __shared__ uint32_t qcnt[WQ_BINS];
    if (threadIdx.x < WQ_BINS) qcnt[threadIdx.x] = 0u;
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += 1024) {
        const int2 r = bins[t];
        atomicAdd(&qcnt[wq_bin((uint32_t)max(0, r.y - r.x))], 1u);
    }
    __syncthreads();
    uint32_t qtot;
    const uint32_t c = threadIdx.x < WQ_BINS ? qcnt[threadIdx.x] : 0u;
    const uint32_t start = block_exclusive_scan(c, lw, &qtot);
    if (threadIdx.x < WQ_BINS) qcnt[threadIdx.x] = start;
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += 1024) {
        const int2 r = bins[t];
        queue[atomicAdd(&qcnt[wq_bin((uint32_t)max(0, r.y - r.x))], 1u)] = (uint32_t)t;  // view 0
    }
    if (threadIdx.x == 0) {
        *n_work = (uint32_t)T;
        ctr[0] = 0u;
        ctr[1] = 0u;
    }
}

// Block lists (level 1): the tiles are grouped into BLK x BLK blocks, and every Gaussian
// whose rectangle meets a block is listed once for it, in rank order, with a 16-bit mask of
// the block's tiles its rectangle covers. A Gaussian meets about 4x fewer blocks than tiles,
// so this list is about 4x shorter than the tile lists, and a chunk's segment of a block's
// list is about 4x longer than its segment of a tile's. When the tile lists overflow the
// capacity, a Gaussian is listed for a block only with the bits of the block's tiles whose
// lists end within the capacity (and not at all when no such bit is left): the level-1 list
// then holds at most as many entries as those tiles, so it never passes the capacity.
constexpr int BLK = 4;

struct BlockGeom {
    int nbx, nby, NB;
};
__device__ __forceinline__ BlockGeom block_geom(int tbx, int tby) {
    BlockGeom b;
    b.nbx = (tbx + BLK - 1) / BLK;
    b.nby = (tby + BLK - 1) / BLK;
    b.NB = b.nbx * b.nby;
    return b;
}

// Bits (BLK ly + lx) of block (bx, by)'s tiles inside the rectangle [x0, x1) x [y0, y1).
__device__ __forceinline__ uint32_t block_cover(ushort4 q, int bx, int by) {
    const int lx0 = max((int)q.x - bx * BLK, 0), lx1 = min((int)q.z - bx * BLK, BLK);
    const int ly0 = max((int)q.y - by * BLK, 0), ly1 = min((int)q.w - by * BLK, BLK);
    if (lx1 <= lx0 || ly1 <= ly0) return 0u;
    const uint32_t row = ((1u << (lx1 - lx0)) - 1u) << lx0;
    uint32_t m = 0u;
    for (int ly = ly0; ly < ly1; ly++) m |= row << (ly * BLK);
    return m;
}

// Bits of block (bx, by)'s tiles whose lists end within the capacity.
__device__ __forceinline__ uint32_t block_complete(int bx, int by, int tbx, int tby,
                                                   const uint32_t* __restrict__ exact_off,
                                                   const uint32_t* __restrict__ counts, uint32_t cap) {
    uint32_t m = 0u;
    for (int ly = 0; ly < BLK; ly++)
        for (int lx = 0; lx < BLK; lx++) {
            const int tx = bx * BLK + lx, ty = by * BLK + ly;
            if (tx >= tbx || ty >= tby) continue;
            const int t = ty * tbx + tx;
            if (exact_off[t] + counts[t] <= cap) m |= 1u << (ly * BLK + lx);
        }
    return m;
}

// The listed bits of rank rectangle q for block (bx, by): its cover, restricted to the tiles
// whose lists fit when some list overflows; 0 means the pair is not listed.
__device__ __forceinline__ uint32_t block_bits(ushort4 q, int bx, int by, bool any_over, int tbx, int tby,
                                               const uint32_t* __restrict__ exact_off,
                                               const uint32_t* __restrict__ counts, uint32_t cap) {
    uint32_t m = block_cover(q, bx, by);
    if (any_over && m != 0u) m &= block_complete(bx, by, tbx, tby, exact_off, counts, cap);
    return m;
}

// Per chunk, the number of listed pairs of each block (one row of the chunk x block matrix,
// row stride NBL = the launch geometry's block count).
__global__ __launch_bounds__(HIST_THREADS) void k_block_hist(int N, int G, int NBL, const ushort4* __restrict__ srect,
                                                             const ViewConsts* __restrict__ vc,
                                                             const uint32_t* __restrict__ exact_off,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ totals, uint32_t cap,
                                                             uint32_t* __restrict__ bmat) {
    extern __shared__ uint32_t hist[];
    const int tbx = vc->tbx, tby = vc->tby;
    const BlockGeom bg = block_geom(tbx, tby);
    for (int b = threadIdx.x; b < bg.NB; b += HIST_THREADS) hist[b] = 0u;
    __syncthreads();
    const bool any_over = totals[0] > cap;
    const int r0 = blockIdx.x * G, r1 = min(N, r0 + G);
    for (int r = r0 + threadIdx.x; r < r1; r += HIST_THREADS) {
        const ushort4 q = srect[r];
        if (q.z <= q.x || q.w <= q.y) continue;
        const int bx0 = q.x / BLK, bx1 = (q.z - 1) / BLK, by0 = q.y / BLK, by1 = (q.w - 1) / BLK;
        for (int by = by0; by <= by1; by++)
            for (int bx = bx0; bx <= bx1; bx++)
                if (!any_over || block_bits(q, bx, by, true, tbx, tby, exact_off, counts, cap) != 0u)
                    atomicAdd(&hist[by * bg.nbx + bx], 1u);
    }
    __syncthreads();
    uint32_t* row = bmat + (size_t)blockIdx.x * NBL;
    for (int b = threadIdx.x; b < bg.NB; b += HIST_THREADS) row[b] = hist[b];
}

// Exclusive scan of each block's column over the chunks (in place); the column total is the
// block's list length. A workgroup takes 16 columns; its 16 thread groups each own one
// sixteenth of the chunks (column sums first, combined through LDS, then the rewrite), with
// eight loads in flight per thread.
constexpr int BSCAN_COLS = 16, BSCAN_PARTS = 16;
__global__ __launch_bounds__(BSCAN_COLS * BSCAN_PARTS) void k_block_scan(const ViewConsts* __restrict__ vc, int nblk,
                                                                         int NBL, uint32_t* __restrict__ bmat,
                                                                         uint32_t* __restrict__ btot) {
    __shared__ uint32_t part_sum[BSCAN_PARTS][BSCAN_COLS];
    const int NB = block_geom(vc->tbx, vc->tby).NB;
    const int c = (int)(threadIdx.x % BSCAN_COLS), part = (int)(threadIdx.x / BSCAN_COLS);
    const int b = (int)blockIdx.x * BSCAN_COLS + c;
    const int per = (nblk + BSCAN_PARTS - 1) / BSCAN_PARTS;
    const int k0 = min(nblk, part * per), k1 = min(nblk, k0 + per);
    uint32_t sum = 0;
    if (b < NB) {
        int k = k0;
        for (; k + 8 <= k1; k += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = bmat[(size_t)(k + u) * NBL + b];
#pragma unroll
            for (int u = 0; u < 8; u++) sum += v[u];
        }
        for (; k < k1; k++) sum += bmat[(size_t)k * NBL + b];
    }
    part_sum[part][c] = sum;
    __syncthreads();
    uint32_t run = 0, total = 0;
    for (int q = 0; q < BSCAN_PARTS; q++) {
        const uint32_t v = part_sum[q][c];
        if (q < part) run += v;
        total += v;
    }
    if (b >= NB) return;
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = bmat[(size_t)(k + u) * NBL + b];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            bmat[(size_t)(k + u) * NBL + b] = run;
            run += v[u];
        }
    }
    for (; k < k1; k++) {
        const size_t o = (size_t)k * NBL + b;
        const uint32_t v = bmat[o];
        bmat[o] = run;
        run += v;
    }
    if (part == 0) btot[b] = total;
}

// Level-1 bucket kernel: one wave per chunk (WAVES consecutive chunks per workgroup, each
// wave with its own block cursors in LDS), chunk c writing its listed pairs of block b from
// boff[b] + (its column offset) on. The wave takes its chunk's ranks 64 at a time, numbers
// their (rank, block) pairs in rank-major order (a wave scan of the ranks' block counts),
// and then places 64 pairs per step, one per lane: the lane finds its pair's rank by a
// binary search over the scan (lane shuffles), lanes with the same block find each other
// by ballots over the block number's bits, the lowest of them advances the block's cursor
// by their count, and each writes at the old cursor plus its place among them. Steps run
// in pair order and a step orders a block's pairs by lane, so every block list is in rank
// order. Lists are stored as two arrays, Gaussian ids and 16-bit tile masks. boff (the
// exclusive prefix of the block lengths) is computed by every wave in LDS; wave 0 of
// workgroup 0 also writes it out (NB + 1 values) for the compaction kernel.
static size_t bucket_lds_bytes(int NB, int waves) {
    const size_t t4 = ((size_t)NB + 3) & ~(size_t)3;
    return t4 * 4 * (size_t)waves;
}

// 4 waves per workgroup when the LDS allows (the default and a request of 4), 2 on request
// or when 4 do not fit, else 1. The choice is by the kernel's whole dynamic LDS (one cursor
// array per wave), so a launch never asks for more than `limit`; 0 when even one wave does
// not fit.
int bucket_waves(int NB, size_t limit, int request) {
    if ((request == 0 || request == 4) && bucket_lds_bytes(NB, 4) <= limit) return 4;
    if (request != 1 && bucket_lds_bytes(NB, 2) <= limit) return 2;
    return bucket_lds_bytes(NB, 1) <= limit ? 1 : 0;
}

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_block_bucket(int N, int G, int nblk, int NBL,
                                                             const uint32_t* __restrict__ sorted_ids,
                                                             const ushort4* __restrict__ srect,
                                                             const ViewConsts* __restrict__ vc,
                                                             const uint32_t* __restrict__ bmat,
                                                             const uint32_t* __restrict__ btot,
                                                             const uint32_t* __restrict__ exact_off,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ totals,
                                                             uint32_t* __restrict__ l1_ids,
                                                             uint16_t* __restrict__ l1_mask,
                                                             uint32_t* __restrict__ boff_out, uint32_t cap) {
    extern __shared__ uint32_t sm[];
    const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
    // consecutive chunks on one XCD (workgroups go to the 8 XCDs round robin)
    const int per = (int)(gridDim.x >> 3);
    const int wg = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    const int chunk = wg * WAVES + wid;
    if (chunk >= nblk) return;  // wave-uniform; no workgroup barrier follows
    const int tbx = vc->tbx, tby = vc->tby;
    const BlockGeom bgm = block_geom(tbx, tby);
    const int NB = bgm.NB, nbx = bgm.nbx;
    uint32_t* cs = sm + (size_t)wid * ((NB + 3) & ~3);  // this wave's cursors
    const uint32_t* row = bmat + (size_t)chunk * NBL;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // cursors: block list start (prefix of the block lengths) + this chunk's column offset
    {
        uint32_t run = 0u;
        for (int b0 = 0; b0 < NB; b0 += 64) {
            const int b = b0 + lane;
            const uint32_t v = b < NB ? btot[b] : 0u;
            uint32_t x = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (b < NB) {
                cs[b] = run + x - v + row[b];
                if (chunk == 0) boff_out[b] = run + x - v;
            }
            run += __shfl(x, 63, 64);
        }
        if (chunk == 0 && lane == 0) boff_out[NB] = run;
    }
    const bool any_over = totals[0] > cap;
    const int nbits = NB > 1 ? 32 - __clz(NB - 1) : 1;  // bits of a block number
    const int r0 = chunk * G, r1 = min(N, r0 + G);
    for (int base = r0; base < r1; base += 64) {
        const int r = base + lane;
        ushort4 q = make_ushort4(0, 0, 0, 0);
        uint32_t gid = 0u;
        if (r < r1) {
            q = srect[r];
            gid = sorted_ids[r];
        }
        const bool has = q.z > q.x && q.w > q.y;
        const int bx0 = q.x / BLK, by0 = q.y / BLK;
        const int wd = has ? (q.z - 1) / BLK - bx0 + 1 : 0;
        const int area = has ? wd * ((q.w - 1) / BLK - by0 + 1) : 0;
        // rank-major numbering of the 64 ranks' pairs
        int pre = area;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(pre, o, 64);
            if (lane >= o) pre += y;
        }
        const int total = __shfl(pre, 63, 64);
        pre -= area;  // exclusive
        const uint32_t qlo = (uint32_t)q.x | ((uint32_t)q.y << 16), qhi = (uint32_t)q.z | ((uint32_t)q.w << 16);
        for (int w0 = 0; w0 < total; w0 += 64) {
            const int p = w0 + lane;
            const bool valid = p < total;
            // the pair's rank: the last lane whose first pair is at or before p (lanes with
            // no pair share their successor's start, so the last one found has pairs)
            int i = 0;
#pragma unroll
            for (int st = 32; st >= 1; st >>= 1) {
                const int cand = i + st;
                if (__shfl(pre, cand, 64) <= p) i = cand;
            }
            const int k = p - __shfl(pre, i, 64);
            const uint32_t lo = (uint32_t)__shfl((int)qlo, i, 64), hi = (uint32_t)__shfl((int)qhi, i, 64);
            const uint32_t g = (uint32_t)__shfl((int)gid, i, 64);
            const int wdi = max(__shfl(wd, i, 64), 1);
            const ushort4 qi = make_ushort4((unsigned short)(lo & 0xFFFFu), (unsigned short)(lo >> 16),
                                            (unsigned short)(hi & 0xFFFFu), (unsigned short)(hi >> 16));
            const int dy = k / wdi;
            const int bx = qi.x / BLK + (k - dy * wdi), by = qi.y / BLK + dy;
            const int b = by * nbx + bx;
            const uint32_t m = valid ? block_bits(qi, bx, by, any_over, tbx, tby, exact_off, counts, cap) : 0u;
            const bool on = m != 0u;
            // lanes placing a pair of the same block
            uint64_t peers = __ballot(on);
            for (int bit = 0; bit < nbits; bit++) {
                const bool set = ((b >> bit) & 1) != 0;
                const uint64_t bal = __ballot(on && set);
                peers &= set ? bal : ~bal;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & below);
            uint32_t start = 0u;
            if (on && rank == 0u) start = atomicAdd(&cs[b], (uint32_t)__popcll(peers));
            const int leader = on ? (int)__builtin_ctzll(peers) : lane;
            start = (uint32_t)__shfl((int)start, leader, 64);
            if (on) {
                const uint32_t slot = start + rank;
                if (slot < cap) {
                    l1_ids[slot] = g;
                    l1_mask[slot] = (uint16_t)m;
                }
            }
        }
    }
}

// Level 2: one workgroup per tile takes its block's list in rank order and keeps the
// entries whose mask has the tile's bit (8 masks and ids per thread in flight per round),
// compacted by wave ballots (a wave's kept entries
// are numbered by popcount, the 4 waves' counts meet in LDS), so the tile's list is written
// front to back by one workgroup. Workgroups are numbered so that each XCD (workgroups
// x, x + 8, ...) takes a contiguous run of blocks, and a block's 16 tiles run back to back
// on one XCD, where the block list they all read stays in L2. A tile whose list crosses
// the capacity holds the empty record N in its slots below it; the tiles past it have
// empty, clamped bins.
constexpr int COMPACT_THREADS = 256, COMPACT_PER = 8;
__global__ __launch_bounds__(COMPACT_THREADS) void k_tile_compact(const ViewConsts* __restrict__ vc,
                                                                  const uint32_t* __restrict__ boff,
                                                                  const uint32_t* __restrict__ l1_ids,
                                                                  const uint16_t* __restrict__ l1_mask,
                                                                  const uint32_t* __restrict__ exact_off,
                                                                  const uint32_t* __restrict__ counts,
                                                                  uint32_t* __restrict__ ids, uint32_t cap,
                                                                  uint32_t empty_id) {
    constexpr int NW = COMPACT_THREADS / 64;
    __shared__ uint32_t wcount[2][COMPACT_PER][NW];  // two rounds' counts: one barrier per round
    const int tbx = vc->tbx, tby = vc->tby;
    const BlockGeom bgm = block_geom(tbx, tby);
    // blocks per XCD, from the view's own block count (a recorded launch may be sized for a
    // larger view: its spare workgroups exit, and the view's blocks still spread over all
    // eight XCDs)
    const int per = (bgm.NB + 7) / 8;
    const int j = (int)(blockIdx.x >> 3);
    if (j >= per * BLK * BLK) return;
    const int b = (int)(blockIdx.x & 7) * per + j / (BLK * BLK), local = j % (BLK * BLK);
    if (b >= bgm.NB) return;
    const int tx = (b % bgm.nbx) * BLK + local % BLK, ty = (b / bgm.nbx) * BLK + local / BLK;
    if (tx >= tbx || ty >= tby) return;
    const int t = ty * tbx + tx;
    const uint32_t off = exact_off[t], cnt = counts[t];
    if (cnt == 0u) return;
    const int tid = (int)threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (off + cnt > cap) {
        for (uint32_t p = off + (uint32_t)tid; p < cap; p += COMPACT_THREADS) ids[p] = empty_id;
        return;
    }
    const uint32_t lo = boff[b], n = boff[b + 1] - lo;
    const uint16_t* masks = l1_mask + lo;
    const uint32_t* lids = l1_ids + lo;
    uint32_t* out = ids + off;
    const uint32_t bit = 1u << local;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t run = 0u;
    int par = 0;
    for (uint32_t base = 0u; base < n; base += COMPACT_THREADS * COMPACT_PER) {
        // masks and ids are read together (one wait per round rather than a mask wait and
        // then an id wait)
        uint32_t mk[COMPACT_PER], v[COMPACT_PER];
#pragma unroll
        for (int k = 0; k < COMPACT_PER; k++) {
            const uint32_t e = base + (uint32_t)(k * COMPACT_THREADS + tid);
            mk[k] = e < n ? (uint32_t)masks[e] : 0u;
            v[k] = e < n ? lids[e] : 0u;
        }
        uint32_t rank[COMPACT_PER];
        bool keep[COMPACT_PER];
#pragma unroll
        for (int k = 0; k < COMPACT_PER; k++) keep[k] = (mk[k] & bit) != 0u;
#pragma unroll
        for (int k = 0; k < COMPACT_PER; k++) {
            const uint64_t bal = __ballot(keep[k]);
            rank[k] = (uint32_t)__popcll(bal & below);
            if (lane == 0) wcount[par][k][wid] = (uint32_t)__popcll(bal);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < COMPACT_PER; k++) {  // entries in list order: round k, then wave, then lane
            uint32_t before = 0u, total = 0u;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const uint32_t c = wcount[par][k][w];
                before += w < wid ? c : 0u;
                total += c;
            }
            const uint32_t p = run + before + rank[k];
            if (keep[k] && p < cnt) out[p] = v[k];
            run += total;
        }
        par ^= 1;
    }
}

// Chunks of at most MAX_CHUNK ranks; a smaller scene is cut finer (down to MIN_CHUNK ranks,
// up to about one chunk per resident bucket workgroup) so that the bucket kernel's serial
// walk per wave stays short: 50k Gaussians make 391 chunks of 128, not 26 of 1924.
constexpr int MIN_CHUNK = 128, CHUNK_TARGET = 768;
static int hist_blocks(int N, int target = 0) {
    const int by_max = (N + MAX_CHUNK - 1) / MAX_CHUNK;
    const int by_min = (N + MIN_CHUNK - 1) / MIN_CHUNK;
    const int nb = std::max(by_max, std::min(by_min, target > 0 ? target : CHUNK_TARGET));
    return nb < 1 ? 1 : nb;
}

// The level-1 (block list) kernels cut the ranks finer: one wave per chunk and a few bytes
// of LDS per block, so thousands of chunks fit at once (about 2048 chunks, at least 64
// ranks each).
constexpr int L1_MIN_CHUNK = 64, L1_CHUNK_TARGET = 2048;
static int l1_blocks(int N) {
    const int by_max = (N + MAX_CHUNK - 1) / MAX_CHUNK;
    const int by_min = (N + L1_MIN_CHUNK - 1) / L1_MIN_CHUNK;
    const int nb = std::max(by_max, std::min(by_min, L1_CHUNK_TARGET));
    return nb < 1 ? 1 : nb;
}

// Stable LSD radix sort of (key, value) u32 pairs, 8-bit digits, 4 passes: per-workgroup LDS
// digit histograms, one scan of the digit-major (digit, workgroup) count matrix, and a
// scatter that ranks each item among equal digits of earlier items with wave ballots. Only
// kernels: no memset or host synchronisation, so a step that uses it can be captured.
constexpr int RS_THREADS = 256, RS_ITEMS = 8, RS_TILE = RS_THREADS * RS_ITEMS;

static int rs_blocks(int N) { return (N + RS_TILE - 1) / RS_TILE; }

// Sort key of Gaussian i: its depth's bits (positive floats order as their bits), invisible
// Gaussians last.
__device__ __forceinline__ uint32_t depth_key(const int32_t* __restrict__ radii, const float* __restrict__ depths,
                                              int i) {
    return radii[i] > 0 ? __builtin_bit_cast(uint32_t, depths[i]) : 0xFFFFFFFFu;
}

// FIRST: the keys are made from (radii, depths) instead of read from `keys`.
template <bool FIRST>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(int N, const uint32_t* __restrict__ keys, int shift,
                                                        uint32_t* __restrict__ counts, int nblk,
                                                        const int32_t* __restrict__ radii,
                                                        const float* __restrict__ depths) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE + threadIdx.x;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const int i = base + k * RS_THREADS;
        if (i < N) atomicAdd(&h[((FIRST ? depth_key(radii, depths, i) : keys[i]) >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Per digit d (one workgroup each): exclusive scan of its row of per-workgroup counts, in
// place, and the digit's total.
__global__ __launch_bounds__(RS_THREADS) void k_rs_rowscan(uint32_t* __restrict__ counts, int nblk,
                                                           uint32_t* __restrict__ rowtot) {
    __shared__ uint32_t lw[RS_THREADS / 64 + 1];
    uint32_t* row = counts + (size_t)blockIdx.x * nblk;
    const int per = (nblk + RS_THREADS - 1) / RS_THREADS;
    const int b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
    uint32_t sum = 0;
    for (int i = b0; i < b1; i++) sum += row[i];
    uint32_t tot;
    uint32_t run = block_exclusive_scan(sum, lw, &tot);
    for (int i = b0; i < b1; i++) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) rowtot[blockIdx.x] = tot;
}

// FIRST: keys made from (radii, depths), values the Gaussian indices. LAST: the keys are not
// written; instead each Gaussian's tile rectangle is gathered to its rank (srect).
template <bool FIRST, bool LAST>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(int N, const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                           uint32_t* __restrict__ vout, int shift,
                                                           const uint32_t* __restrict__ counts,
                                                           const uint32_t* __restrict__ rowtot, int nblk,
                                                           const int32_t* __restrict__ radii,
                                                           const float* __restrict__ depths,
                                                           const ushort4* __restrict__ rect,
                                                           ushort4* __restrict__ srect) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_THREADS / 64][256];
    __shared__ uint32_t lw[RS_THREADS / 64 + 1];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // this thread's keys and values first: their loads overlap each other and the scans
    uint32_t kr[RS_ITEMS], vr[RS_ITEMS];
    ushort4 rq[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int i = blockIdx.x * RS_TILE + r * RS_THREADS + threadIdx.x;
        kr[r] = i < N ? (FIRST ? depth_key(radii, depths, i) : kin[i]) : 0u;
        vr[r] = i < N ? (FIRST ? (uint32_t)i : vin[i]) : 0u;
    }
    if (LAST) {
#pragma unroll
        for (int r = 0; r < RS_ITEMS; r++) {
            const int i = blockIdx.x * RS_TILE + r * RS_THREADS + threadIdx.x;
            rq[r] = i < N ? rect[vr[r]] : make_ushort4(0, 0, 0, 0);
        }
    }
    // Items are ranked into an LDS copy of the workgroup's tile, ordered by digit, and written
    // out from there, so each digit's run of items goes to consecutive addresses.
    __shared__ uint32_t sk[RS_TILE], sv[RS_TILE];
    __shared__ ushort4 sq[LAST ? RS_TILE : 1];
    __shared__ uint32_t gbase[256], lbase[256];
    uint32_t tot;
    const uint32_t dbase = block_exclusive_scan(rowtot[threadIdx.x], lw, &tot);
    const uint32_t c0 = counts[(size_t)threadIdx.x * nblk + blockIdx.x];
    const uint32_t c1 = (int)blockIdx.x + 1 < nblk ? counts[(size_t)threadIdx.x * nblk + blockIdx.x + 1]
                                                     : rowtot[threadIdx.x];
    run[threadIdx.x] = dbase + c0;
    gbase[threadIdx.x] = dbase + c0;  // first output slot of this workgroup's digit-d items
    lbase[threadIdx.x] = block_exclusive_scan(c1 - c0, lw, &tot);  // and their first LDS slot
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; w++) wcnt[w][threadIdx.x] = 0u;
        __syncthreads();
        const int i = blockIdx.x * RS_TILE + r * RS_THREADS + threadIdx.x;
        const bool valid = i < N;
        const uint32_t key = kr[r];
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__builtin_popcountll(peers & lt);
        if (valid && rank == 0u) wcnt[wid][d] = (uint32_t)__builtin_popcountll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wid; w++) pos += wcnt[w][d];
            const uint32_t loc = pos - gbase[d] + lbase[d];
            sk[loc] = key;
            sv[loc] = vr[r];
            if (LAST) sq[loc] = rq[r];
        }
        __syncthreads();
        uint32_t add = 0u;
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; w++) add += wcnt[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
    const int items = min(RS_TILE, N - (int)blockIdx.x * RS_TILE);
    for (int j = threadIdx.x; j < items; j += RS_THREADS) {
        const uint32_t key = sk[j];
        const uint32_t d = (key >> shift) & 255u;
        const uint32_t pos = gbase[d] + ((uint32_t)j - lbase[d]);
        if (LAST) srect[pos] = sq[j];
        else kout[pos] = key;
        vout[pos] = sv[j];
    }
}

size_t lists_temp_bytes(int N) { return ((size_t)256 * (size_t)rs_blocks(N > 0 ? N : 1) + 256 + 64) * 4; }

// [chunks x T] tile histogram rows, then (the same buffer) the [chunks x blocks] block
// histogram and offsets
size_t lists_hist_bytes(int N, int T, int chunk_target) {
    return (size_t)std::max(hist_blocks(N, chunk_target), l1_blocks(N)) * (size_t)T * 4;  // blocks <= tiles
}

// 4x4-tile blocks of a tile grid
static int blocks_of(const Geom& g) { return ((g.tbx + BLK - 1) / BLK) * ((g.tby + BLK - 1) / BLK); }
size_t lists_block_bytes(const Geom& g) { return ((size_t)blocks_of(g) * 2 + 2) * 4; }

static size_t g_lds_limit = 0;
static size_t lds_limit() {
    if (g_lds_limit == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        g_lds_limit = 65536;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            g_lds_limit = prop.sharedMemPerBlock;
    }
    return g_lds_limit;
}

bool lists_supported(int N, int T) {
    (void)N;
    return (size_t)T * 4 <= lds_limit() && bucket_lds_bytes(T, 1) <= lds_limit();
}

// Depth ranks of all N Gaussians: four 8-bit passes ping-pong between (dkeys, dvals) and
// (dkeys_out, sorted_ids); the first makes its keys from the projection, the last writes the
// sorted ids and the rectangles in rank order (srect) and drops the keys.
void lists_rank(hipStream_t s, int N, const ProjOut& pj, ListBufs& L) {
    if (N <= 0) return;
    const int nblk = rs_blocks(N);
    uint32_t* counts = reinterpret_cast<uint32_t*>(L.temp);
    uint32_t* rowtot = counts + (size_t)256 * nblk;
    uint32_t *ka = nullptr, *va = nullptr, *kb = L.dkeys, *vb = L.dvals;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        if (pass == 0)
            hipLaunchKernelGGL(k_rs_hist<true>, dim3(nblk), dim3(RS_THREADS), 0, s, N, ka, shift, counts, nblk,
                               pj.radii, pj.depths);
        else
            hipLaunchKernelGGL(k_rs_hist<false>, dim3(nblk), dim3(RS_THREADS), 0, s, N, ka, shift, counts, nblk,
                               pj.radii, pj.depths);
        hipLaunchKernelGGL(k_rs_rowscan, dim3(256), dim3(RS_THREADS), 0, s, counts, nblk, rowtot);
        if (pass == 0)
            hipLaunchKernelGGL((k_rs_scatter<true, false>), dim3(nblk), dim3(RS_THREADS), 0, s, N, ka, va, kb, vb,
                               shift, counts, rowtot, nblk, pj.radii, pj.depths, pj.rect, L.srect);
        else if (pass < 3)
            hipLaunchKernelGGL((k_rs_scatter<false, false>), dim3(nblk), dim3(RS_THREADS), 0, s, N, ka, va, kb, vb,
                               shift, counts, rowtot, nblk, pj.radii, pj.depths, pj.rect, L.srect);
        else
            hipLaunchKernelGGL((k_rs_scatter<false, true>), dim3(nblk), dim3(RS_THREADS), 0, s, N, ka, va, kb, vb,
                               shift, counts, rowtot, nblk, pj.radii, pj.depths, pj.rect, L.srect);
        ka = kb;
        va = vb;
        kb = (pass % 2 == 0) ? L.dkeys_out : L.dkeys;
        vb = (pass % 2 == 0) ? L.sorted_ids : L.dvals;
    }
}

// A launch that fails leaves the lists unwritten; the raster must not be queued behind it.
static void check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

static void allow_lds(const void* fn, size_t lds, const char* what) {
    if (lds <= 65536) return;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void lists_count(hipStream_t s, int N, const ProjOut& pj, const ViewConsts* vc, const Geom& g, ListBufs& L,
                 const RasterWork* wq) {
    if (N > 0) {
        const int nblk = hist_blocks(N, L.chunk_target), G = (N + nblk - 1) / nblk;
        const size_t lds = (size_t)g.T * 4;
        allow_lds((const void*)k_tile_hist, lds, "tile histogram LDS");
        hipLaunchKernelGGL(k_tile_hist, dim3(nblk), dim3(HIST_THREADS), lds, s, N, G, g.T, L.srect, vc, L.hist);
        check_launch("tile histogram launch");
        hipLaunchKernelGGL(k_col_sum, dim3(cdiv(g.T, 64)), dim3(64 * COLSCAN_PARTS), 0, s, vc, nblk, g.T, L.hist,
                           L.tile_counts);
    } else {
        fill_zero(s, L.tile_counts, (size_t)g.T * 4);
    }
    hipLaunchKernelGGL(k_prefix, dim3(1), dim3(1024), 0, s, vc, L.tile_counts, L.exact_off, L.bins, L.totals,
                       (uint32_t)L.capacity, wq ? wq->queue : nullptr, wq ? wq->n_work : nullptr,
                       wq ? wq->ctr : nullptr);
}

// Level-1 bucket launch with WAVES waves (chunks) per workgroup (LDS by the launch's block
// count NBL).
template <int WAVES>
static void launch_bucket(hipStream_t s, int nblk, int N, int G, int NBL, const ViewConsts* vc, ListBufs& L) {
    const size_t lds = bucket_lds_bytes(NBL, WAVES);
    allow_lds((const void*)k_block_bucket<WAVES>, lds, "block bucket LDS");
    uint32_t* btot = L.blocks;
    uint32_t* boff = L.blocks + NBL;
    hipLaunchKernelGGL(k_block_bucket<WAVES>, dim3(8 * cdiv(cdiv(nblk, WAVES), 8)), dim3(64 * WAVES), lds, s, N, G, nblk, NBL,
                       L.sorted_ids, L.srect, vc, L.hist, btot, L.exact_off, L.tile_counts, L.totals, L.l1_ids,
                       L.l1_mask, boff, (uint32_t)L.capacity);
    check_launch("block bucket launch");
}

static std::atomic<int> g_bucket_waves{0};
void set_bucket_waves(int waves) { g_bucket_waves.store(waves == 1 || waves == 2 || waves == 4 ? waves : 0); }

void lists_bucket(hipStream_t s, int N, const ViewConsts* vc, const Geom& g, ListBufs& L) {
    if (N <= 0) return;
    const int nblk = l1_blocks(N), G = (N + nblk - 1) / nblk;
    const int NBL = blocks_of(g);
    uint32_t* btot = L.blocks;
    uint32_t* boff = L.blocks + NBL;
    // level 1: block histogram (the tile histogram's buffer is free once the tile counts are
    // summed), column offsets, bucket
    const size_t lds = (size_t)NBL * 4;
    allow_lds((const void*)k_block_hist, lds, "block histogram LDS");
    hipLaunchKernelGGL(k_block_hist, dim3(nblk), dim3(HIST_THREADS), lds, s, N, G, NBL, L.srect, vc, L.exact_off,
                       L.tile_counts, L.totals, (uint32_t)L.capacity, L.hist);
    check_launch("block histogram launch");
    hipLaunchKernelGGL(k_block_scan, dim3(cdiv(NBL, BSCAN_COLS)), dim3(BSCAN_COLS * BSCAN_PARTS), 0, s, vc, nblk, NBL,
                       L.hist, btot);
    switch (bucket_waves(NBL, lds_limit(), L.bucket_waves > 0 ? L.bucket_waves : g_bucket_waves.load())) {
        case 4: launch_bucket<4>(s, nblk, N, G, NBL, vc, L); break;
        case 2: launch_bucket<2>(s, nblk, N, G, NBL, vc, L); break;
        case 1: launch_bucket<1>(s, nblk, N, G, NBL, vc, L); break;
        default: throw std::runtime_error("block count exceeds the bucket kernel's LDS");
    }
    // level 2: tile lists
    hipLaunchKernelGGL(k_tile_compact, dim3(8 * cdiv(NBL, 8) * BLK * BLK), dim3(COMPACT_THREADS), 0, s, vc, boff,
                       L.l1_ids, L.l1_mask, L.exact_off, L.tile_counts, L.ids, (uint32_t)L.capacity, (uint32_t)N);
    check_launch("tile compaction launch");
}

// ------------------------------------------------------------------------------------
// A10/A11: rasterisation forward. One 256-thread workgroup per 16x16 tile (monolithic,
// nd_rasterize_forward_kernel metal:499-618) or per (tile, 512-entry chunk)
// (rasterize_forward_chunked_kernel metal:2960-3076 + merge metal:3081-3134).
constexpr int RB = 256;

// Pixel of a thread inside its 16x16 tile: each 64-lane wave covers one 8x8 quadrant, so
// the lanes of a wave see nearby pixels (coherent early-outs, fewer divergent passes).
__device__ __forceinline__ int tile_lx() { return (int)((threadIdx.x & 7) | ((threadIdx.x >> 3) & 8)); }
__device__ __forceinline__ int tile_ly() { return (int)(((threadIdx.x >> 3) & 7) | ((threadIdx.x >> 4) & 8)); }

// ms_expf_core (ms_math.h) of four arguments in lockstep, step by step, so that the four
// chains interleave: alpha = min(0.999, opacity exp(-sigma)) for the sigma in range
// [0, 5.55) (exp(0) for the others, whose `ok` is cleared). The same operations in the same
// order as ms_expf_core: bit-identical results.
__device__ __forceinline__ void ms_expf_core4(const float sg[4], bool ok[4], const float4 a[4], float al[4]) {
    float x[4], n[4], r[4], p[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const bool in_range = sg[u] >= 0.f && sg[u] < 5.55f;
        ok[u] = ok[u] & in_range;
        x[u] = in_range ? -sg[u] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) n[u] = rintf(x[u] * 1.44269502162933349609375f);
#pragma unroll
    for (int u = 0; u < 4; u++) r[u] = fmaf(-n[u], 0.693147182464599609375f, x[u]);
#pragma unroll
    for (int u = 0; u < 4; u++) r[u] = fmaf(-n[u], -1.904654323148236e-09f, r[u]);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(1.98412701e-04f, r[u], 1.38888892e-03f);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(p[u], r[u], 8.33333377e-03f);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(p[u], r[u], 4.16666679e-02f);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(p[u], r[u], 1.66666672e-01f);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(p[u], r[u], 0.5f);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(p[u], r[u], 1.0f);
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = fmaf(p[u], r[u], 1.0f);
#pragma unroll
    for (int u = 0; u < 4; u++)
        al[u] = fminf(0.999f, a[u].z * (p[u] * ms_bits_to_float((uint32_t)((int)n[u] + 127) << 23)));
}

// Quadrants of the entry's 16x16 tile (bit q: the 8x8 quadrant of wave q, x half q & 1, y
// half q >> 1) holding a pixel centre that can take the entry. A pixel takes it only where
// sigma < 5.55 and o exp(-sigma) >= 1/255, i.e. sigma <= ln(255 o) (the early-outs of
// metal:580-588, 1081-1090); sigma is a convex quadratic, so its minimum over a quadrant's
// box of pixel centres is 0 when the Gaussian's centre lies in the box and is otherwise
// reached on one of the box's edges, each a one-dimensional quadratic minimised in closed
// form. A quadrant is left out only when that minimum exceeds the threshold by a margin
// (1e-4 of the largest term of sigma over the box, plus 1e-4) that covers the rounding of
// both this test and the per-pixel evaluation: every pixel of the quadrant would reject the
// entry, so leaving it out changes no result. Conics that are not positive definite or not
// finite keep every quadrant (the per-pixel test decides).
__device__ __forceinline__ uint32_t quad_mask(const float4 A, const float4 B, float ox, float oy) {
    const float cx = A.x, cy = A.y, o = A.z, ca = A.w, cb = B.x, cc = B.y;
    if (o < (1.f / 255.f) * (1.f - 1e-5f)) return 0u;  // alpha <= o below 1/255 everywhere
    const float det = ca * cc - cb * cb;
    if (!(ca > 0.f && cc > 0.f && det > 0.f && ms_finite(cx) && ms_finite(cy) && ms_finite(det) && o <= 1.f))
        return 15u;
    const float t = fminf(5.55f, logf(255.f * o));
    const float ia = 1.f / ca, ic = 1.f / cc;
    auto sig = [&](float dx, float dy) { return 0.5f * (ca * dx * dx + cc * dy * dy) + cb * dx * dy; };
    uint32_t m = 0u;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float x0 = ox + 8.f * (float)(q & 1), y0 = oy + 8.f * (float)(q >> 1);
        const float dlx = cx - (x0 + 7.f), dhx = cx - x0, dly = cy - (y0 + 7.f), dhy = cy - y0;
        float mn = 0.f;
        if (!(dlx <= 0.f && dhx >= 0.f && dly <= 0.f && dhy >= 0.f)) {
            const float ya = fminf(fmaxf(-cb * dlx * ic, dly), dhy), yb = fminf(fmaxf(-cb * dhx * ic, dly), dhy);
            const float xa = fminf(fmaxf(-cb * dly * ia, dlx), dhx), xb = fminf(fmaxf(-cb * dhy * ia, dlx), dhx);
            mn = fminf(fminf(sig(dlx, ya), sig(dhx, yb)), fminf(sig(xa, dly), sig(xb, dhy)));
        }
        const float Dx = fmaxf(fabsf(dlx), fabsf(dhx)), Dy = fmaxf(fabsf(dly), fabsf(dhy));
        const float big = 0.5f * (ca * Dx * Dx + cc * Dy * Dy) + fabsf(cb) * Dx * Dy;
        if (!(mn >= t + 1e-4f * (big + 1.f))) m |= 1u << q;
    }
    return m;
}

// Indices (batch slots, increasing) of the staged entries whose mask holds this wave's
// quadrant and that pass `extra`; returns their count. Wave-local: the caller synchronises
// before reading the list.
template <int NB, class Extra>
__device__ __forceinline__ int wave_entry_list(const uint8_t* sM, uint8_t* list, int n, Extra extra) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int cnt = 0;
#pragma unroll
    for (int j0 = 0; j0 < NB; j0 += 64) {
        const int e = j0 + lane;
        const bool take = e < n && ((sM[e] >> wid) & 1u) && extra(e);
        const uint64_t bal = __ballot(take);
        if (take) list[cnt + __builtin_popcountll(bal & below)] = (uint8_t)e;
        cnt += __builtin_popcountll(bal);
    }
    return cnt;
}

__device__ __forceinline__ void composite_range(int lo, int hi, float px, float py, bool inside, const EntryView& ev,
                                                float& T, float C[3], int& last, int& taken) {
    __shared__ float4 sA[RB];
    __shared__ float4 sB[RB];
    __shared__ float sCx[RB];
    __shared__ uint8_t sM[RB];
    __shared__ __align__(4) uint8_t sL[4][RB + 4];  // + 4: a step reads its four indices as one word
    const int wid = (int)(threadIdx.x >> 6);
    const float ox = px - (float)tile_lx(), oy = py - (float)tile_ly();  // the tile's corner pixel
    T = 1.f;
    C[0] = C[1] = C[2] = 0.f;
    last = lo - 1;
    taken = 0;
    bool done = !inside;
    for (int b0 = lo; b0 < hi; b0 += RB) {
        if (__syncthreads_and(done)) break;
        const int idx = b0 + threadIdx.x;
        if (idx < hi) {  // colours stored clamped (max(c + 1/2, 0)): per entry, not per pixel
            const uint32_t g = ev.ids[idx];
            const float4 a = ev.rec.A[g];
            const float4 b = ev.rec.B[g];
            sA[threadIdx.x] = a;
            sB[threadIdx.x] = make_float4(b.x, b.y, fmaxf(b.z + 0.5f, 0.f), fmaxf(b.w + 0.5f, 0.f));
            sCx[threadIdx.x] = fmaxf(ev.rec.C[g].x + 0.5f, 0.f);
            sM[threadIdx.x] = (uint8_t)quad_mask(a, b, ox, oy);
        }
        __syncthreads();
        const int n = min(RB, hi - b0);
        // only the entries some pixel of this wave's quadrant can take (the others would be
        // rejected by every lane): same accepted entries in the same order
        const int cnt = wave_entry_list<RB>(sM, sL[wid], n, [](int) { return true; });
        __syncthreads();
        // Four entries per step, without branches inside a step: their indices come in one
        // LDS word and all their record reads are issued before the first is used, so a
        // step waits on LDS twice instead of several times per entry. Every lane evaluates
        // every entry; the per-pixel tests only select whether its result is taken, so the
        // arithmetic of the taken entries is the same as entry by entry.
        for (int k0 = 0; k0 < cnt; k0 += 4) {
            if (__ballot(!done) == 0ull) break;  // wave-uniform: every pixel of the wave is done
            const uint32_t e4 = *reinterpret_cast<const uint32_t*>(&sL[wid][k0]);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const bool have = k0 + u < cnt;  // the last step may hold fewer than four
                const int e = have ? (int)((e4 >> (8 * u)) & 255u) : 0;
                const float4 a = sA[e];
                const float4 b = sB[e];
                const float cz = sCx[e];
                const float dx = a.x - px, dy = a.y - py;
                const float sigma = fmaf(0.5f, fmaf(a.w, dx * dx, b.y * dy * dy), b.x * dx * dy);
                const bool in_range = ms_finite(sigma) && sigma >= 0.f && sigma < 5.55f;
                const float alpha = fminf(0.999f, a.z * ms_expf_core(in_range ? -sigma : 0.f));
                // bitwise: no short-circuit branches around the arithmetic above
                const bool take = have & !done & in_range & (alpha >= 1.f / 255.f);
                const float nT = T * (1.f - alpha);
                const bool sat = take & (nT <= 1e-4f);
                const bool acc = take & !sat;
                // an entry not taken adds colour x 0 (colours are finite): C stays exactly as
                // it is, with no branch around the colour reads
                const float vis = acc ? alpha * T : 0.f;
                C[0] = fmaf(b.z, vis, C[0]);
                C[1] = fmaf(b.w, vis, C[1]);
                C[2] = fmaf(cz, vis, C[2]);
                T = acc ? nT : T;
                last = acc ? b0 + e : (sat ? b0 + e - 1 : last);
                taken += acc ? 1 : 0;
                done = done | sat;
            }
        }
    }
}


// One workgroup: every valid (view, tile) item is binned by its list length (8-entry bins,
// longest first, all lists of 2040 entries or more in the first bin) with LDS counters, and
// placed bin by bin. The order inside a bin is arbitrary; it only balances the persistent
// raster kernels, whose outputs do not depend on it.
__global__ __launch_bounds__(WQ_THREADS) void k_work_queue(const RasterView* __restrict__ views, int B, int Tmax,
                                                          uint32_t* __restrict__ queue, uint32_t* __restrict__ n_work,
                                                          uint32_t* __restrict__ ctr) {
    __shared__ uint32_t cnt[WQ_BINS];
    __shared__ uint32_t lw[17];
    const int tid = (int)threadIdx.x, n = B * Tmax;
    if (tid < WQ_BINS) cnt[tid] = 0u;
    __syncthreads();
    auto bin_of = [&](int k, uint32_t& code) -> int {  // -1: not an item
        const int v = k / Tmax, t = k % Tmax;
        const RasterView& rv = views[v];
        if (rv.chunked || t >= rv.vc->T) return -1;
        const int2 r = rv.bins[t];
        const uint32_t len = (uint32_t)max(0, r.y - r.x);
        code = ((uint32_t)v << WORK_TILE_BITS) | (uint32_t)t;
        return wq_bin(len);
    };
    constexpr int U = 8;  // items per thread in flight
    for (int k0 = 0; k0 < n; k0 += WQ_THREADS * U) {
        int b[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + u * WQ_THREADS + tid;
            uint32_t code;
            b[u] = k < n ? bin_of(k, code) : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (b[u] >= 0) atomicAdd(&cnt[b[u]], 1u);
    }
    __syncthreads();
    uint32_t total = 0u;
    const uint32_t c = tid < WQ_BINS ? cnt[tid] : 0u;
    const uint32_t start = block_exclusive_scan(c, lw, &total);
    if (tid < WQ_BINS) cnt[tid] = start;
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += WQ_THREADS * U) {
        int b[U];
        uint32_t code[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + u * WQ_THREADS + tid;
            b[u] = k < n ? bin_of(k, code[u]) : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (b[u] >= 0) queue[atomicAdd(&cnt[b[u]], 1u)] = code[u];
    }
    if (tid == 0) {
        *n_work = total;
        ctr[0] = 0u;
        ctr[1] = 0u;
    }
}

__global__ __launch_bounds__(RB) void k_raster_fwd_persist(const RasterView* __restrict__ views,
                                                           const uint32_t* __restrict__ work,
                                                           const uint32_t* __restrict__ n_work,
                                                           uint32_t* __restrict__ ctr) {
    const uint32_t total = *n_work;
    for (;;) {
        const uint32_t it = next_item(ctr);
        if (it >= total) break;
        const uint32_t code = work[it];
        const RasterView& rv = views[code >> WORK_TILE_BITS];
        const ViewConsts& vc = *rv.vc;
        const int t = (int)(code & ((1u << WORK_TILE_BITS) - 1u));
        const int W = vc.W, H = vc.H, tbx = vc.tbx;
        const int tx = t % tbx, ty = t / tbx;
        const int j = tx * TILE + tile_lx(), i = ty * TILE + tile_ly();
        const bool inside = i < H && j < W;
        const int2 r = rv.bins[t];
        float T, C[3];
        int last, taken;
        composite_range(r.x, r.y, (float)j, (float)i, inside, rv.pk, T, C, last, taken);
        if (inside) {
            const size_t pid = (size_t)i * W + j;
            rv.final_T[pid] = T;
            rv.final_idx[pid] = last;
            if (rv.n_contrib) rv.n_contrib[pid] = taken;
            rv.img[3 * pid + 0] = fminf(fmaxf(fmaf(vc.bg[0], T, C[0]), 0.f), 1.f);
            rv.img[3 * pid + 1] = fminf(fmaxf(fmaf(vc.bg[1], T, C[1]), 0.f), 1.f);
            rv.img[3 * pid + 2] = fminf(fmaxf(fmaf(vc.bg[2], T, C[2]), 0.f), 1.f);
        }
    }
}

__global__ __launch_bounds__(RB) void k_raster_fwd_chunk(const ViewConsts* __restrict__ vc,
                                                         const int2* __restrict__ bins, EntryView pk,
                                                         float* __restrict__ chunk_T, float* __restrict__ chunk_C,
                                                         int32_t* __restrict__ chunk_idx) {
    const int W = vc->W, H = vc->H, tbx = vc->tbx;
    const int t = blockIdx.x, k = blockIdx.y;
    if (t >= vc->T) return;
    const int tx = t % tbx, ty = t / tbx;
    const int j = tx * TILE + tile_lx(), i = ty * TILE + tile_ly();
    const bool inside = i < H && j < W;
    const int2 r = bins[t];
    const int cs = r.x + k * CHUNK, ce = min(r.x + (k + 1) * CHUNK, r.y);
    const size_t P = (size_t)W * H;
    const size_t o = (size_t)k * P + (size_t)i * W + j;
    if (cs >= ce) {
        if (inside) {
            chunk_T[o] = 1.f;
            chunk_C[3 * o] = 0.f; chunk_C[3 * o + 1] = 0.f; chunk_C[3 * o + 2] = 0.f;
            chunk_idx[o] = -1;
        }
        return;
    }
    float T, C[3];
    int last, taken;
    composite_range(cs, ce, (float)j, (float)i, inside, pk, T, C, last, taken);
    if (inside) {
        chunk_T[o] = T;
        chunk_C[3 * o] = C[0]; chunk_C[3 * o + 1] = C[1]; chunk_C[3 * o + 2] = C[2];
        chunk_idx[o] = last;
    }
}

__global__ void k_raster_merge(const ViewConsts* __restrict__ vc, int K, const float* __restrict__ chunk_T,
                               const float* __restrict__ chunk_C, int32_t* __restrict__ chunk_idx,
                               float* __restrict__ img, float* __restrict__ final_T, int32_t* __restrict__ final_idx) {
    const size_t P = (size_t)vc->W * vc->H;
    const size_t pid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P) return;
    const float bg0 = vc->bg[0], bg1 = vc->bg[1], bg2 = vc->bg[2];
    float Tr = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f;
    int last = -1, cutoff = K;
    for (int k = 0; k < K; k++) {
        const size_t o = (size_t)k * P + pid;
        const int cf = chunk_idx[o];
        if (cf < 0 && k > 0) break;
        const float cT = chunk_T[o];
        C0 = fmaf(chunk_C[3 * o], Tr, C0);
        C1 = fmaf(chunk_C[3 * o + 1], Tr, C1);
        C2 = fmaf(chunk_C[3 * o + 2], Tr, C2);
        Tr *= cT;
        if (cf >= 0) last = cf;
        if (Tr <= 1e-4f) { cutoff = k + 1; break; }
    }
    for (int k = cutoff; k < K; k++) chunk_idx[(size_t)k * P + pid] = -1;
    final_T[pid] = Tr;
    final_idx[pid] = last;
    img[3 * pid + 0] = fminf(fmaxf(fmaf(bg0, Tr, C0), 0.f), 1.f);
    img[3 * pid + 1] = fminf(fmaxf(fmaf(bg1, Tr, C1), 0.f), 1.f);
    img[3 * pid + 2] = fminf(fmaxf(fmaf(bg2, Tr, C2), 0.f), 1.f);
}

void raster_work_build(hipStream_t s, const RasterWork& w) {
    if (w.B * w.Tmax <= 0) return;
    hipLaunchKernelGGL(k_work_queue, dim3(1), dim3(WQ_THREADS), 0, s, w.views, w.B, w.Tmax, w.queue, w.n_work, w.ctr);
}

void raster_forward(hipStream_t s, const ListBufs& L, const ViewConsts* vc, const Geom& g, RasterBufs& R,
                    const RasterWork* work) {
    if (R.K <= 1) {
        raster_forward_views(s, *work);
        return;
    }
    hipLaunchKernelGGL(k_raster_fwd_chunk, dim3(g.T, R.K), dim3(RB), 0, s, vc, L.bins, L.view, R.chunk_T, R.chunk_C,
                       R.chunk_idx);
    hipLaunchKernelGGL(k_raster_merge, dim3(cdiv((size_t)g.W * g.H, 256)), dim3(256), 0, s, vc, (int)R.K, R.chunk_T,
                       R.chunk_C, R.chunk_idx, R.img, R.final_T, R.final_idx);
}

// ------------------------------------------------------------------------------------
// A12-A14: SSIM + L1 loss forward and backward (metal:3388-3450, 3563-3681, 3764-3850),
// with the optional training mask (metal:3602-3653, 3815-3834).
// Intermediate buffers are planar ([q][H][W]) so every store of a 64-lane wave is one
// contiguous 256-byte run.
constexpr int SB = 16;
// The vertical backward takes 32 columns per workgroup: each of its halo rows is then one
// full 128-byte line per plane (16 columns fetch half-used lines). Every pixel's arithmetic
// is that of 16x16 tiles.
constexpr int FW = 32;

// Horizontal pass, vertical pass, SSIM and L1 per pixel, derivative fields and their
// horizontal transpose, for one 16x16 tile: the horizontal statistics (metal:3388-3450) of
// the tile's 26x26 halo are computed in LDS from the image and the target (15 bytes per
// pixel read, 36x26 with the halo) instead of being written out and read back (60 bytes per
// pixel each way, and a misaligned 26-column halo row spans two or three 128-byte lines);
// every value is summed in the same order as a separate pass would, so the results are the
// same bit for bit.
__global__ __launch_bounds__(256) void k_ssim_fused(const ViewConsts* __restrict__ vc, const float* __restrict__ img,
                                                    float* __restrict__ dbuf, float* __restrict__ partials,
                                                    uint32_t* __restrict__ nanc) {
    const int W = vc->W, H = vc->H;
    const int gx = (W + SB - 1) / SB;
    if ((int)blockIdx.x >= gx || (int)blockIdx.y * SB >= H) return;
    const uint8_t* __restrict__ gt = vc->gt;
    const uint8_t* __restrict__ mask = vc->mask;
    const float ssim_w = vc->ssim_w;
    constexpr int TD = SB + 10;  // 26
    constexpr int TI = SB + 20;  // 36: input columns of the horizontal pass over the halo
    // one channel of the target and render at a time (the channels are taken in turn), so
    // the workgroup's LDS is 29 KB and five workgroups fit on a CU
    __shared__ float sg[TD][TI];
    __shared__ float sr[TD][TI];
    __shared__ float shp[TD][TD][5];
    __shared__ float sf[3][SB][TD];
    __shared__ float s_ssim[SB][SB];
    __shared__ float s_l1[SB][SB];
    __shared__ float red[256];
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int bx = blockIdx.x * SB - 5, by = blockIdx.y * SB - 5;
    const size_t P = (size_t)W * H;
    s_ssim[ly][lx] = 0.f;
    s_l1[ly][lx] = 0.f;
    // the target (x = gt / 255) and render over the halo, rows by .. by + 25, columns
    // bx - 5 .. bx + 30 (0 outside), all three channels read once into registers (each
    // thread's halo pixels k = tid + 256 q), so the channel passes below read no memory
    constexpr int HALO_Q = (TD * TI + 255) / 256;
    float hg[HALO_Q][3], hr[HALO_Q][3];
#pragma unroll
    for (int q = 0; q < HALO_Q; q++) {
        const int k = (int)threadIdx.x + 256 * q;
        const int sy = k / TI, sx = k % TI;
        const int gy = by + sy, gxx = bx - 5 + sx;
        const bool in = k < TD * TI && gxx >= 0 && gxx < W && gy >= 0 && gy < H;
        const size_t pid = in ? (size_t)gy * W + gxx : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            hg[q][c] = in ? (float)gt[3 * pid + c] * (1.0f / 255.0f) : 0.f;
            hr[q][c] = in ? img[3 * pid + c] : 0.f;
        }
    }
    for (int c = 0; c < 3; c++) {
        // channel c into LDS; the previous channel's readers passed the barrier that ends
        // its iteration
#pragma unroll
        for (int q = 0; q < HALO_Q; q++) {
            const int k = (int)threadIdx.x + 256 * q;
            if (k < TD * TI) {
                sg[k / TI][k % TI] = hg[q][c];
                sr[k / TI][k % TI] = hr[q][c];
            }
        }
        __syncthreads();
        // horizontal 11-tap statistics at the halo's pixels; zero outside the image (the
        // vertical pass's padding)
        for (int k = threadIdx.x; k < TD * TD; k += 256) {
            const int sy = k / TD, sx = k % TD;
            const int gy = by + sy, gxx = bx + sx;
            float mx = 0.f, my = 0.f, sxq = 0.f, syq = 0.f, cxy = 0.f;
            if (gxx >= 0 && gxx < W && gy >= 0 && gy < H) {
#pragma unroll
                for (int d = 0; d < 11; d++) {
                    const float w = c_gauss1d[d];
                    const float gv = sg[sy][sx + d], rv = sr[sy][sx + d];
                    const float wg = w * gv, wr = w * rv;
                    mx = fmaf(w, gv, mx);
                    my = fmaf(w, rv, my);
                    sxq = fmaf(wg, gv, sxq);
                    syq = fmaf(wr, rv, syq);
                    cxy = fmaf(wg, rv, cxy);
                }
            }
            shp[sy][sx][0] = mx;
            shp[sy][sx][1] = my;
            shp[sy][sx][2] = sxq;
            shp[sy][sx][3] = syq;
            shp[sy][sx][4] = cxy;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < SB * TD; k += 256) {
            const int dy = k / TD, dx = k % TD;
            float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 11; kk++) {
                const float w = c_gauss1d[kk];
#pragma unroll
                for (int q = 0; q < 5; q++) m[q] = fmaf(w, shp[dy + kk][dx][q], m[q]);
            }
            if (!(ms_finite(m[0]) && ms_finite(m[1]) && ms_finite(m[2]) && ms_finite(m[3]) && ms_finite(m[4]))) {
                sf[0][dy][dx] = 0.f;
                sf[1][dy][dx] = 0.f;
                sf[2][dy][dx] = 0.f;
                atomicAdd(&nanc[4], 1u);
                continue;
            }
            const float mu_x = m[0], mu_y = m[1];
            const float sxx = fmaxf(m[2] - mu_x * mu_x, 0.f), syy = fmaxf(m[3] - mu_y * mu_y, 0.f);
            const float sxy = m[4] - mu_x * mu_y;
            const float A = 2.0f * mu_x * mu_y + 0.0001f, Bv = 2.0f * sxy + 0.0009f;
            const float Cd = mu_x * mu_x + mu_y * mu_y + 0.0001f, D = sxx + syy + 0.0009f;
            const float iCD = 1.0f / (Cd * D);
            const float dmu = 2.0f * Bv * (mu_x * Cd - A * mu_y) / (Cd * Cd * D);
            const float dsyq = -A * Bv * iCD / D, dsxy = 2.0f * A * iCD;
            // training mask: a masked centre pixel owns no loss, so its derivative fields
            // and its loss terms are multiplied by zero (by one otherwise: exact)
            const int gpx = bx + dx, gpy = by + dy + 5;
            const bool inimg = gpx >= 0 && gpx < W && gpy >= 0 && gpy < H;
            const float mval = (mask && inimg && mask[(size_t)gpy * W + gpx] < 128) ? 0.f : 1.f;
            sf[0][dy][dx] = mval * (dmu - 2.0f * mu_y * dsyq - mu_x * dsxy);
            sf[1][dy][dx] = mval * (2.0f * dsyq);
            sf[2][dy][dx] = mval * dsxy;
            if (dx >= 5 && dx < 5 + SB && inimg) {
                s_ssim[dy][dx - 5] += mval * (A * Bv) / (Cd * D);
                s_l1[dy][dx - 5] += mval * fabsf(sg[dy + 5][dx + 5] - sr[dy + 5][dx + 5]);
            }
        }
        __syncthreads();
        const int px = blockIdx.x * SB + lx, py = blockIdx.y * SB + ly;
        if (px < W && py < H) {
            float h1 = 0.f, h2 = 0.f, h3 = 0.f;
#pragma unroll
            for (int d = 0; d < 11; d++) {
                const float w = c_gauss1d[10 - d];
                h1 = fmaf(w, sf[0][ly][lx + d], h1);
                h2 = fmaf(w, sf[1][ly][lx + d], h2);
                h3 = fmaf(w, sf[2][ly][lx + d], h3);
            }
            const size_t pid = (size_t)py * W + px;
            dbuf[(c * 3 + 0) * P + pid] = h1;
            dbuf[(c * 3 + 1) * P + pid] = h2;
            dbuf[(c * 3 + 2) * P + pid] = h3;
        }
        __syncthreads();
    }
    // pixel loss with the reference's per-thread bookkeeping (see ssim_loss_counted in the
    // CPU restatement): in-image pixels whose sweep thread lies outside the image add only w.
    const int px = blockIdx.x * SB + lx, py = blockIdx.y * SB + ly;
    float pl = 0.f;
    if (px < W && py < H) {
        const int ii = ly * 26 + lx + 5;
        const int tr = ii < 256 ? ii : ii - 256;
        const bool counted = blockIdx.x * SB + (tr % 16) < W && blockIdx.y * SB + (tr / 16) < H;
        pl = counted ? ssim_w * (1.0f - s_ssim[ly][lx] / 3.0f) + (1.0f - ssim_w) * s_l1[ly][lx] / 3.0f : ssim_w;
    }
    red[threadIdx.x] = pl;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.y * gx + blockIdx.x] = red[0];
}

// Workgroup (0, 0) also sums the fused pass's per-workgroup loss partials (written by the
// previous launch) into the loss, in index order, after its own pixels.
__device__ __forceinline__ void loss_sum(int n, const float* __restrict__ partials, double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) acc += (double)partials[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ __launch_bounds__(256) void k_ssim_vbwd(const ViewConsts* __restrict__ vc, const float* __restrict__ img,
                                                   const float* __restrict__ dbuf, float* __restrict__ v_rend,
                                                   uint32_t* __restrict__ nanc, const float* __restrict__ partials,
                                                   double* __restrict__ loss_out) {
    constexpr int TH = SB + 10;
    __shared__ float sh[9][TH][FW];
    const int W = vc->W, H = vc->H;
    if ((int)blockIdx.x * FW >= W || (int)blockIdx.y * SB >= H) return;
    const uint8_t* __restrict__ gt = vc->gt;
    const uint8_t* __restrict__ mask = vc->mask;
    const float ssim_w = vc->ssim_w, inv_n = vc->inv_n;
    const int bx = blockIdx.x * FW, by = blockIdx.y * SB - 5;
    const size_t P = (size_t)W * H;
    for (int k = threadIdx.x; k < TH * FW; k += 256) {
        const int sy = k / FW, sx = k % FW;
        const int gy = by + sy, gx = bx + sx;
        const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
        const size_t pid = (size_t)gy * W + gx;
#pragma unroll
        for (int q = 0; q < 9; q++) sh[q][sy][sx] = in ? dbuf[q * P + pid] : 0.f;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < SB * FW; k += 256) {
        const int ly = k / FW, lx = k % FW;
        const int px = bx + lx, py = blockIdx.y * SB + ly;
        if (px >= W || py >= H) continue;
        const size_t pid = (size_t)py * W + px;
        // a masked pixel drops its own L1 gradient; SSIM gradients of its neighbours'
        // windows still flow through it
        const float m_center = (mask && mask[pid] < 128) ? 0.f : 1.f;
        for (int c = 0; c < 3; c++) {
            float f1 = 0.f, f2 = 0.f, f3 = 0.f;
#pragma unroll
            for (int d = 0; d < 11; d++) {
                const float w = c_gauss1d[10 - d];
                f1 = fmaf(w, sh[c * 3 + 0][ly + d][lx], f1);
                f2 = fmaf(w, sh[c * 3 + 1][ly + d][lx], f2);
                f3 = fmaf(w, sh[c * 3 + 2][ly + d][lx], f3);
            }
            const float rv = img[3 * pid + c];
            const float gv = (float)gt[3 * pid + c] * (1.0f / 255.0f);
            const float vs = f1 + rv * f2 + gv * f3;
            const float vl = m_center * ((gv > rv) ? -1.0f : ((gv < rv) ? 1.0f : 0.0f));
            float v = inv_n * (-ssim_w * vs + (1.0f - ssim_w) * vl);
            if (!ms_finite(v)) {
                atomicAdd(&nanc[6], 1u);
                v = 0.f;
            }
            v_rend[3 * pid + c] = v;
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0) loss_sum(((W + SB - 1) / SB) * ((H + SB - 1) / SB), partials, loss_out);
}

void ssim_loss(hipStream_t s, const ViewConsts* vc, const Geom& gm, const float* img, float* dbuf,
               float* v_rend, float* loss_partials, double* loss_out, uint32_t* nan_counts) {
    dim3 g((gm.W + SB - 1) / SB, (gm.H + SB - 1) / SB), gw((gm.W + FW - 1) / FW, (gm.H + SB - 1) / SB);
    hipLaunchKernelGGL(k_ssim_fused, g, dim3(256), 0, s, vc, img, dbuf, loss_partials, nan_counts);
    hipLaunchKernelGGL(k_ssim_vbwd, gw, dim3(256), 0, s, vc, img, dbuf, v_rend, nan_counts, loss_partials, loss_out);
}

// ------------------------------------------------------------------------------------
// A15/A16: rasterisation backward (rasterize_backward_kernel metal:931-1135; chunked
// metal:3139-3366). Entries are walked back to front in LDS batches; each wave reduces
// its 64 lanes' 9 gradient terms of four entries at a time (wave_sum4_rows), the 4 wave
// sums meet in LDS and one thread per entry issues the global atomics.
constexpr int BB = 192;  // backward batch: most tiles' walks end within one (128 was about 1 % slower, 64 6 %)

__device__ __forceinline__ void backward_range(int lo, int hi, int bin_final, bool inside, float px, float py,
                                               float T, float buf0, float buf1, float buf2, float T_final,
                                               const float bg[3], float vo0, float vo1, float vo2,
                                               const EntryView& ev, GradBufs& G, int tile_max) {
    __shared__ float4 sA[BB];
    __shared__ float4 sB[BB];
    __shared__ float2 sC[BB];
    __shared__ uint32_t sF[BB];
    __shared__ uint8_t sM[BB];
    __shared__ __align__(4) uint8_t sL[4][BB + 4];  // + 4: a step reads its four indices as one word
    __shared__ uint8_t sIn[4][BB];   // sIn[w][slot]: entry `slot` is on wave w's list
    __shared__ float red[4][BB][9];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const float ox = px - (float)tile_lx(), oy = py - (float)tile_ly();  // the tile's corner pixel
    // the colour accumulated behind the current entry plus the background term T_final bg
    // enters the alpha gradient only as its dot product with the pixel's colour gradient:
    // S = sum_c (buf_c + T_final bg_c) vo_c, advanced by fac * (colour . vo) per entry taken
    float S = fmaf(fmaf(T_final, bg[2], buf2), vo2, fmaf(fmaf(T_final, bg[1], buf1), vo1, fmaf(T_final, bg[0], buf0) * vo0));
    const int top = min(hi, tile_max + 1);  // entries past every pixel's last contributor are skipped
    // the last contributor of any pixel of this wave: entries behind it give the wave zeros
    int wmax = inside ? bin_final : INT_MIN;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off));
    const int tag = wave_sum4_tag();
    for (int bhi = top; bhi > lo; bhi -= BB) {
        const int blo = max(lo, bhi - BB);
        const int n = bhi - blo;
        __syncthreads();
        if ((int)threadIdx.x < n) {  // colours stored clamped; which of them pass the clamp in sF
            const int idx = bhi - 1 - threadIdx.x;  // slot 0 = furthest back
            const uint32_t g = ev.ids[idx];
            const float4 a = ev.rec.A[g];
            const float4 b = ev.rec.B[g];
            const float c0 = b.z + 0.5f, c1 = b.w + 0.5f, c2 = ev.rec.C[g].x + 0.5f;
            sA[threadIdx.x] = a;
            sB[threadIdx.x] = make_float4(b.x, b.y, fmaxf(c0, 0.f), fmaxf(c1, 0.f));
            sC[threadIdx.x] = make_float2(fmaxf(c2, 0.f), __builtin_bit_cast(float, g));
            sF[threadIdx.x] = (c0 >= 0.f ? 1u : 0u) | (c1 >= 0.f ? 2u : 0u) | (c2 >= 0.f ? 4u : 0u);
            sM[threadIdx.x] = (uint8_t)quad_mask(a, b, ox, oy);
        }
        __syncthreads();
        // this wave's entries, back to front: its quadrant reachable (every lane rejects the
        // others) and not behind the wave's last contributor
        const int cnt = wave_entry_list<BB>(sM, sL[wid], n, [&](int e) { return bhi - 1 - e <= wmax; });
#pragma unroll
        for (int j0 = 0; j0 < BB; j0 += 64) sIn[wid][j0 + lane] = 0;
        __syncthreads();
        for (int k = lane; k < cnt; k += 64) sIn[wid][sL[wid][k]] = 1;
        // four entries at a time: their 4 x 9 terms are reduced together (wave_sum4_rows).
        // Their indices come in one LDS word and all their record reads are issued before
        // the first is used; every lane evaluates every entry and the per-pixel tests only
        // select whether its terms and transmittance update are taken (the same arithmetic
        // as entry by entry for the entries taken, zeros for the others).
        for (int t4 = 0; t4 < cnt; t4 += 4) {
            float g[4][9];
            const uint32_t e4 = *reinterpret_cast<const uint32_t*>(&sL[wid][t4]);
            // the four entries side by side, phase by phase, so that their independent chains
            // (offsets, sigma, the exp polynomial, the reciprocal) interleave in the issue
            float4 a[4], b[4];
            float r2[4], dx[4], dy[4], sg[4], al[4];
            bool ok[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const bool have = t4 + u < cnt;  // the last step may hold fewer than four
                const int tt = have ? (int)((e4 >> (8 * u)) & 255u) : 0;
                a[u] = sA[tt];
                b[u] = sB[tt];
                r2[u] = sC[tt].x;
                ok[u] = have & inside & (bhi - 1 - tt <= bin_final);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                dx[u] = a[u].x - px;
                dy[u] = a[u].y - py;
                sg[u] = fmaf(0.5f, fmaf(a[u].w, dx[u] * dx[u], b[u].y * dy[u] * dy[u]), b[u].x * dx[u] * dy[u]);
            }
            ms_expf_core4(sg, ok, a, al);
            float au[4], ra[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                // bitwise: no short-circuit branches around the arithmetic above
                const bool use = ok[u] & (al[u] >= 1.f / 255.f) & (al[u] < 0.999f);
                // one select: an entry not taken has alpha 0, so ra = 1 exactly, T and the
                // colour sums stay as they are and every term below is a product with 0 (all
                // factors are finite), which is what selecting each of them would give
                au[u] = use ? al[u] : 0.f;
                ra[u] = recip_nr(1.f - au[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float Tn = T * ra[u];
                const float fac = au[u] * Tn;
                // d loss / d alpha = sum_c (colour_c Tn - ra (buf_c + T_final bg_c)) vo_c
                const float rv = fmaf(r2[u], vo2, fmaf(b[u].w, vo1, b[u].z * vo0));
                const float v_alpha = fmaf(Tn, rv, -ra[u] * S);
                S = fmaf(fac, rv, S);
                T = Tn;
                // the entry's own constants (conic, opacity) multiply the wave sums once per
                // entry after the reduction: per pixel only v_sigma and its products with the
                // offsets are summed
                const float vs = -au[u] * v_alpha;  // d loss / d sigma
                const float vdx = vs * dx[u], vdy = vs * dy[u];
                g[u][0] = fac * vo0;
                g[u][1] = fac * vo1;
                g[u][2] = fac * vo2;
                g[u][3] = vdx * dx[u];
                g[u][4] = vdx * dy[u];
                g[u][5] = vdy * dy[u];
                g[u][6] = vdx;
                g[u][7] = vdy;
                g[u][8] = vs;
            }
            // lane 15 of each row holds the sum of list entry t4 + tag
            const bool wr = (lane & 15) == 15 && t4 + tag < cnt;
            const int slot = wr ? sL[wid][t4 + tag] : 0;
            // the nine reductions are independent: all of them first (their DPP and permlane
            // steps interleave), then one masked store sequence
            float v[9];
#pragma unroll
            for (int k = 0; k < 9; k++) v[k] = wave_sum4_rows(g[0][k], g[1][k], g[2][k], g[3][k]);
            if (wr) {
#pragma unroll
                for (int k = 0; k < 9; k++) red[wid][slot][k] = v[k];
            }
        }
        __syncthreads();
        // per entry: the four waves' sums, then the entry's constants: sums of v_sigma dx^2,
        // dx dy, dy^2, dx, dy and v_sigma -> conic, mean and opacity-logit gradients
        // (d sigma / d mean = conic (dx, dy); the logit's derivative of alpha is
        // alpha (1 - opacity)); the results overwrite the entry's row of wave 0's sums
        if ((int)threadIdx.x < n) {
            const int tt = threadIdx.x;
            float s[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                float v = 0.f;
#pragma unroll
                for (int w = 0; w < 4; w++) v += sIn[w][tt] ? red[w][tt][k] : 0.f;
                s[k] = v;
            }
            const float4 a = sA[tt];
            const float4 b = sB[tt];
            const uint32_t f = sF[tt];
            float* o = red[0][tt];
            o[G2D_RGB + 0] = (f & 1u) ? s[0] : 0.f;  // colours clamped in the forward get none
            o[G2D_RGB + 1] = (f & 2u) ? s[1] : 0.f;
            o[G2D_RGB + 2] = (f & 4u) ? s[2] : 0.f;
            o[G2D_CONIC + 0] = 0.5f * s[3];
            o[G2D_CONIC + 1] = 0.5f * s[4];
            o[G2D_CONIC + 2] = 0.5f * s[5];
            o[G2D_XY + 0] = fmaf(a.w, s[6], b.x * s[7]);
            o[G2D_XY + 1] = fmaf(b.x, s[6], b.y * s[7]);
            o[G2D_OPAC] = -(1.f - a.z) * s[8];
        }
        __syncthreads();
        // one lane per (entry, field): consecutive lanes add the fields of one record, so a
        // wave's atomics touch about seven records' lines instead of 64 scattered ones per
        // field; zero sums are skipped
        for (int q = threadIdx.x; q < n * G2D; q += RB) {
            const int tt = q / G2D, k = q - tt * G2D;
            const float v = red[0][tt][k];
            if (v != 0.f) gadd(&G.g[G2D * (size_t)__builtin_bit_cast(uint32_t, sC[tt].y) + k], v);
        }
    }
}

__device__ __forceinline__ int block_max_int(int v) {
    __shared__ int sm;
    if (threadIdx.x == 0) sm = INT_MIN;
    __syncthreads();
    atomicMax(&sm, v);
    __syncthreads();
    const int r = sm;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(RB) void k_raster_bwd_persist(const RasterView* __restrict__ views,
                                                           const uint32_t* __restrict__ work,
                                                           const uint32_t* __restrict__ n_work,
                                                           uint32_t* __restrict__ ctr) {
    const uint32_t total = *n_work;
    for (;;) {
        const uint32_t it = next_item(ctr + 1);
        if (it >= total) break;
        const uint32_t code = work[it];
        const RasterView& rv = views[code >> WORK_TILE_BITS];
        const ViewConsts& vc = *rv.vc;
        const int t = (int)(code & ((1u << WORK_TILE_BITS) - 1u));
        const int W = vc.W, H = vc.H, tbx = vc.tbx;
        const int tx = t % tbx, ty = t / tbx;
        const int j = tx * TILE + tile_lx(), i = ty * TILE + tile_ly();
        const bool inside = i < H && j < W;
        const int2 r = rv.bins[t];
        if (r.y <= r.x) continue;  // uniform over the workgroup
        const size_t pid = inside ? (size_t)i * W + j : 0;
        const float Tf = inside ? rv.final_T[pid] : 1.f;
        const int bf = inside ? rv.final_idx[pid] : -1;
        const float vo0 = inside ? rv.v_rend[3 * pid] : 0.f, vo1 = inside ? rv.v_rend[3 * pid + 1] : 0.f,
                    vo2 = inside ? rv.v_rend[3 * pid + 2] : 0.f;
        const int tmax = block_max_int(inside ? bf : INT_MIN);
        const float bg[3] = {vc.bg[0], vc.bg[1], vc.bg[2]};
        GradBufs G = rv.G;
        backward_range(r.x, r.y, bf, inside, (float)j, (float)i, Tf, 0.f, 0.f, 0.f, Tf, bg, vo0, vo1, vo2, rv.pk, G,
                       tmax);
    }
}

__global__ void k_chunk_prefix_suffix(const ViewConsts* __restrict__ vc, int K, const float* __restrict__ chunk_T,
                                      const float* __restrict__ chunk_C, float* __restrict__ prefix_T,
                                      float* __restrict__ after_C) {
    const size_t P = (size_t)vc->W * vc->H;
    const size_t pid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P) return;
    float pT = 1.f;
    for (int k = 0; k < K; k++) {
        const size_t o = (size_t)k * P + pid;
        prefix_T[o] = pT;
        pT *= chunk_T[o];
    }
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int k = K - 1; k >= 0; k--) {
        const size_t o = (size_t)k * P + pid;
        after_C[3 * o] = a0; after_C[3 * o + 1] = a1; after_C[3 * o + 2] = a2;
        const float pk = prefix_T[o];
        a0 += pk * chunk_C[3 * o];
        a1 += pk * chunk_C[3 * o + 1];
        a2 += pk * chunk_C[3 * o + 2];
    }
}

__global__ __launch_bounds__(RB) void k_raster_bwd_chunk(const ViewConsts* __restrict__ vc,
                                                         const int2* __restrict__ bins, EntryView pk,
                                                         const float* __restrict__ final_T,
                                                         const int32_t* __restrict__ chunk_idx,
                                                         const float* __restrict__ prefix_T,
                                                         const float* __restrict__ chunk_T,
                                                         const float* __restrict__ after_C,
                                                         const float* __restrict__ v_rend, GradBufs G) {
    const int W = vc->W, H = vc->H, tbx = vc->tbx;
    const int t = blockIdx.x, k = blockIdx.y;
    if (t >= vc->T) return;
    const int tx = t % tbx, ty = t / tbx;
    const int j = tx * TILE + tile_lx(), i = ty * TILE + tile_ly();
    const bool inside = i < H && j < W;
    const int2 r = bins[t];
    const int cs = r.x + k * CHUNK, ce = min(r.x + (k + 1) * CHUNK, r.y);
    if (cs >= ce) return;
    const size_t P = (size_t)W * H;
    const size_t pid = inside ? (size_t)i * W + j : 0;
    const size_t o = (size_t)k * P + pid;
    const int bf = inside ? chunk_idx[o] : -1;
    const bool act = inside && bf >= cs;
    const int tmax = block_max_int(act ? bf : INT_MIN);
    if (tmax < cs) return;  // uniform: no pixel of this tile has a contributor in this chunk
    const float Ts = act ? prefix_T[o] * chunk_T[o] : 1.f;
    const float b0 = act ? after_C[3 * o] : 0.f, b1 = act ? after_C[3 * o + 1] : 0.f, b2 = act ? after_C[3 * o + 2] : 0.f;
    const float Tf = inside ? final_T[pid] : 1.f;
    const float vo0 = inside ? v_rend[3 * pid] : 0.f, vo1 = inside ? v_rend[3 * pid + 1] : 0.f,
                vo2 = inside ? v_rend[3 * pid + 2] : 0.f;
    const float bg[3] = {vc->bg[0], vc->bg[1], vc->bg[2]};
    backward_range(cs, ce, act ? bf : -1, act, (float)j, (float)i, Ts, b0, b1, b2, Tf, bg, vo0, vo1, vo2, pk, G, tmax);
}

void raster_backward(hipStream_t s, const ListBufs& L, const RasterBufs& R, const ViewConsts* vc, const Geom& g,
                     const float* v_rend, GradBufs& G, const RasterWork* work) {
    if (R.K <= 1) {
        raster_backward_views(s, *work);
        return;
    }
    hipLaunchKernelGGL(k_chunk_prefix_suffix, dim3(cdiv((size_t)g.W * g.H, 256)), dim3(256), 0, s, vc, (int)R.K,
                       R.chunk_T, R.chunk_C, R.prefix_T, R.after_C);
    hipLaunchKernelGGL(k_raster_bwd_chunk, dim3(g.T, R.K), dim3(RB), 0, s, vc, L.bins, L.view, R.final_T, R.chunk_idx,
                       R.prefix_T, R.chunk_T, R.after_C, v_rend, G);
}

// ---- four pixels per lane: one 64-lane wave per 16x16 tile --------------------------
// Lane l covers the 2x2 pixel block at (2 (l % 8), 2 (l / 8)) of its tile. Every pixel runs
// the same operation sequence as in the one-pixel-per-thread kernels (identical images); in
// the backward each lane first sums its four pixels' terms, so the wave reduction per entry
// is shared by 256 pixels instead of 64, and the one wave issues the entry's atomics itself.
constexpr int RB4 = 64;

__device__ __forceinline__ void px4_coords(int t, int tbx, float px[4], float py[4], int ix[4], int iy[4]) {
    const int lane = threadIdx.x & 63;
    const int x0 = (t % tbx) * TILE + 2 * (lane & 7), y0 = (t / tbx) * TILE + 2 * (lane >> 3);
    for (int q = 0; q < 4; q++) {
        ix[q] = x0 + (q & 1);
        iy[q] = y0 + (q >> 1);
        px[q] = (float)ix[q];
        py[q] = (float)iy[q];
    }
}

__global__ __launch_bounds__(RB4) void k_raster_fwd_persist4(const RasterView* __restrict__ views,
                                                             const uint32_t* __restrict__ work,
                                                             const uint32_t* __restrict__ n_work,
                                                             uint32_t* __restrict__ ctr) {
    __shared__ float4 sA[RB4];
    __shared__ float4 sB[RB4];
    __shared__ float sCx[RB4];
    const uint32_t total = *n_work;
    const int lane = threadIdx.x;
    for (;;) {
        const uint32_t it = next_item(ctr);
        if (it >= total) break;
        const uint32_t code = work[it];
        const RasterView& rv = views[code >> WORK_TILE_BITS];
        const ViewConsts& vc = *rv.vc;
        const int t = (int)(code & ((1u << WORK_TILE_BITS) - 1u));
        const int W = vc.W, H = vc.H;
        float px[4], py[4];
        int ix[4], iy[4];
        px4_coords(t, vc.tbx, px, py, ix, iy);
        const int2 r = rv.bins[t];
        float T[4], C[4][3];
        int last[4];
        bool done[4];
        for (int q = 0; q < 4; q++) {
            T[q] = 1.f;
            C[q][0] = C[q][1] = C[q][2] = 0.f;
            last[q] = r.x - 1;
            done[q] = !(ix[q] < W && iy[q] < H);
        }
        for (int b0 = r.x; b0 < r.y; b0 += RB4) {
            if (__syncthreads_and(done[0] && done[1] && done[2] && done[3])) break;
            const int idx = b0 + lane;
            if (idx < r.y) {
                const uint32_t g = rv.pk.ids[idx];
                sA[lane] = rv.pk.rec.A[g];
                sB[lane] = rv.pk.rec.B[g];
                sCx[lane] = rv.pk.rec.C[g].x;
            }
            __syncthreads();
            const int n = min(RB4, r.y - b0);
            for (int e = 0; e < n; e++) {
                if (done[0] && done[1] && done[2] && done[3]) break;
                const float4 a = sA[e];
                const float4 b = sB[e];
                const float cz = sCx[e];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (done[q]) continue;
                    const float dx = a.x - px[q], dy = a.y - py[q];
                    const float sigma = fmaf(0.5f, fmaf(a.w, dx * dx, b.y * dy * dy), b.x * dx * dy);
                    if (!ms_finite(sigma) || sigma < 0.f || sigma >= 5.55f) continue;
                    const float alpha = fminf(0.999f, a.z * ms_expf_core(-sigma));
                    if (alpha < 1.f / 255.f) continue;
                    const float nT = T[q] * (1.f - alpha);
                    if (nT <= 1e-4f) {
                        last[q] = b0 + e - 1;
                        done[q] = true;
                        continue;
                    }
                    const float vis = alpha * T[q];
                    C[q][0] = fmaf(fmaxf(b.z + 0.5f, 0.f), vis, C[q][0]);
                    C[q][1] = fmaf(fmaxf(b.w + 0.5f, 0.f), vis, C[q][1]);
                    C[q][2] = fmaf(fmaxf(cz + 0.5f, 0.f), vis, C[q][2]);
                    T[q] = nT;
                    last[q] = b0 + e;
                }
            }
            __syncthreads();
        }
        for (int q = 0; q < 4; q++) {
            if (!(ix[q] < W && iy[q] < H)) continue;
            const size_t pid = (size_t)iy[q] * W + ix[q];
            rv.final_T[pid] = T[q];
            rv.final_idx[pid] = last[q];
            rv.img[3 * pid + 0] = fminf(fmaxf(fmaf(vc.bg[0], T[q], C[q][0]), 0.f), 1.f);
            rv.img[3 * pid + 1] = fminf(fmaxf(fmaf(vc.bg[1], T[q], C[q][1]), 0.f), 1.f);
            rv.img[3 * pid + 2] = fminf(fmaxf(fmaf(vc.bg[2], T[q], C[q][2]), 0.f), 1.f);
        }
    }
}

__global__ __launch_bounds__(RB4) void k_raster_bwd_persist4(const RasterView* __restrict__ views,
                                                             const uint32_t* __restrict__ work,
                                                             const uint32_t* __restrict__ n_work,
                                                             uint32_t* __restrict__ ctr) {
    __shared__ float4 sA[RB4];
    __shared__ float4 sB[RB4];
    __shared__ float2 sC[RB4];
    const uint32_t total = *n_work;
    const int lane = threadIdx.x;
    for (;;) {
        const uint32_t it = next_item(ctr + 1);
        if (it >= total) break;
        const uint32_t code = work[it];
        const RasterView& rv = views[code >> WORK_TILE_BITS];
        const ViewConsts& vc = *rv.vc;
        const int t = (int)(code & ((1u << WORK_TILE_BITS) - 1u));
        const int2 r = rv.bins[t];
        if (r.y <= r.x) continue;  // uniform over the wave
        const int W = vc.W, H = vc.H;
        float px[4], py[4];
        int ix[4], iy[4];
        px4_coords(t, vc.tbx, px, py, ix, iy);
        float T[4], S[4], vo[4][3];
        int bf[4];
        bool inside[4];
        int mx = INT_MIN;
        for (int q = 0; q < 4; q++) {
            inside[q] = ix[q] < W && iy[q] < H;
            const size_t pid = inside[q] ? (size_t)iy[q] * W + ix[q] : 0;
            T[q] = inside[q] ? rv.final_T[pid] : 1.f;
            bf[q] = inside[q] ? rv.final_idx[pid] : -1;
            for (int c = 0; c < 3; c++) vo[q][c] = inside[q] ? rv.v_rend[3 * pid + c] : 0.f;
            // sum_c (colour behind the entry + T_final bg_c) vo_c, as in backward_range
            S[q] = fmaf(T[q] * vc.bg[2], vo[q][2], fmaf(T[q] * vc.bg[1], vo[q][1], T[q] * vc.bg[0] * vo[q][0]));
            if (inside[q]) mx = max(mx, bf[q]);
        }
        const int tmax = block_max_int(mx);
        const int top = min(r.y, tmax + 1);
        for (int bhi = top; bhi > r.x; bhi -= RB4) {
            const int blo = max(r.x, bhi - RB4);
            const int n = bhi - blo;
            __syncthreads();
            if (lane < n) {
                const int idx = bhi - 1 - lane;  // slot 0 = furthest back
                const uint32_t g = rv.pk.ids[idx];
                sA[lane] = rv.pk.rec.A[g];
                sB[lane] = rv.pk.rec.B[g];
                sC[lane] = make_float2(rv.pk.rec.C[g].x, __builtin_bit_cast(float, g));
            }
            __syncthreads();
            for (int tt = 0; tt < n; tt++) {
                const int e = bhi - 1 - tt;
                const float4 a = sA[tt];
                const float4 b = sB[tt];
                const float cz = sC[tt].x;
                const float r0 = fmaxf(b.z + 0.5f, 0.f), r1 = fmaxf(b.w + 0.5f, 0.f), r2 = fmaxf(cz + 0.5f, 0.f);
                float g[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                bool any = false;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    bool valid = inside[q] && e <= bf[q];
                    float alpha = 0.f, dx = 0.f, dy = 0.f;
                    if (valid) {
                        dx = a.x - px[q];
                        dy = a.y - py[q];
                        const float sigma = fmaf(0.5f, fmaf(a.w, dx * dx, b.y * dy * dy), b.x * dx * dy);
                        if (sigma < 0.f || sigma >= 5.55f) {
                            valid = false;
                        } else {
                            alpha = fminf(0.999f, a.z * ms_expf_core(-sigma));
                            if (alpha < 1.f / 255.f) valid = false;
                        }
                    }
                    if (!(valid && alpha < 0.999f)) continue;
                    any = true;
                    const float ra = recip_nr(1.f - alpha);
                    T[q] *= ra;
                    const float fac = alpha * T[q];
                    const float rv = fmaf(r2, vo[q][2], fmaf(r1, vo[q][1], r0 * vo[q][0]));
                    const float v_alpha = fmaf(T[q], rv, -ra * S[q]);
                    S[q] = fmaf(fac, rv, S[q]);
                    const float vs = -alpha * v_alpha;
                    const float vdx = vs * dx, vdy = vs * dy;
                    g[0] += fac * vo[q][0];
                    g[1] += fac * vo[q][1];
                    g[2] += fac * vo[q][2];
                    g[3] += vdx * dx;
                    g[4] += vdx * dy;
                    g[5] += vdy * dy;
                    g[6] += vdx;
                    g[7] += vdy;
                    g[8] += vs;
                }
                if (__ballot(any) == 0ull) continue;
#pragma unroll
                for (int k = 0; k < 9; k++) g[k] = wave_sum_to_lane63(g[k]);
                if (lane == 63) {
                    bool nz = false;
#pragma unroll
                    for (int k = 0; k < 9; k++) nz |= (g[k] != 0.f);
                    if (nz) {
                        const uint32_t gid = __builtin_bit_cast(uint32_t, sC[tt].y);
                        float* o = rv.G.g + G2D * (size_t)gid;
                        if (b.z + 0.5f >= 0.f) gadd(o + G2D_RGB + 0, g[0]);
                        if (b.w + 0.5f >= 0.f) gadd(o + G2D_RGB + 1, g[1]);
                        if (cz + 0.5f >= 0.f) gadd(o + G2D_RGB + 2, g[2]);
                        gadd(o + G2D_CONIC + 0, 0.5f * g[3]);
                        gadd(o + G2D_CONIC + 1, 0.5f * g[4]);
                        gadd(o + G2D_CONIC + 2, 0.5f * g[5]);
                        gadd(o + G2D_XY + 0, fmaf(a.w, g[6], b.x * g[7]));
                        gadd(o + G2D_XY + 1, fmaf(b.x, g[6], b.y * g[7]));
                        gadd(o + G2D_OPAC, -(1.f - a.z) * g[8]);
                    }
                }
            }
        }
    }
}

void raster_backward_views(hipStream_t s, const RasterWork& w) {
    if (w.ppl == 4)
        hipLaunchKernelGGL(k_raster_bwd_persist4, dim3(w.grid), dim3(RB4), 0, s, w.views, w.queue, w.n_work, w.ctr);
    else
        hipLaunchKernelGGL(k_raster_bwd_persist, dim3(w.grid), dim3(RB), 0, s, w.views, w.queue, w.n_work, w.ctr);
}

void raster_forward_views(hipStream_t s, const RasterWork& w) {
    if (w.ppl == 4)
        hipLaunchKernelGGL(k_raster_fwd_persist4, dim3(w.grid), dim3(RB4), 0, s, w.views, w.queue, w.n_work, w.ctr);
    else
        hipLaunchKernelGGL(k_raster_fwd_persist, dim3(w.grid), dim3(RB), 0, s, w.views, w.queue, w.n_work, w.ctr);
}

// ------------------------------------------------------------------------------------
// A17 (SH part): Adam on features_rest, one thread per coefficient. The reference fuses
// this into the projection backward (metal:1987-2041); here each coefficient's Adam
// update is an independent coalesced element, the basis value being recomputed from the
// Gaussian's (pre-update) mean. Visible Gaussians and bases <= degrees_to_use only.
__device__ __forceinline__ void adam_el(float& p, float& m, float& v, float g, float step, float b1, float b2,
                                        float bc2s, float eps) {
    const float mn = fmaf(b1, m, (1.0f - b1) * g);
    const float vn = fmaf(b2, v, (1.0f - b2) * g * g);
    p -= step * mn / (sqrtf(vn) / bc2s + eps);
    m = mn;
    v = vn;
}

__device__ __forceinline__ float sh_basis_one(int b, float x, float y, float z) {
    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    switch (b) {
        case 0: return -SH_C1 * y;
        case 1: return SH_C1 * z;
        case 2: return -SH_C1 * x;
        case 3: return SH_C2_0 * xy;
        case 4: return SH_C2_1 * yz;
        case 5: return SH_C2_2 * (2.f * zz - xx - yy);
        case 6: return SH_C2_3 * xz;
        case 7: return SH_C2_4 * (xx - yy);
        case 8: return SH_C3_0 * y * (3.f * xx - yy);
        case 9: return SH_C3_1 * xy * z;
        case 10: return SH_C3_2 * y * (4.f * zz - xx - yy);
        case 11: return SH_C3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy);
        case 12: return SH_C3_4 * x * (4.f * zz - xx - yy);
        case 13: return SH_C3_5 * z * (xx - yy);
        default: return SH_C3_6 * x * (xx - 3.f * yy);
    }
}

// SH basis values of degrees 1..3 in the order of the fused SH backward (metal:1993-2033).
__device__ __forceinline__ void sh_basis_all(float x, float y, float z, float* b) {
    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    b[0] = -SH_C1 * y;
    b[1] = SH_C1 * z;
    b[2] = -SH_C1 * x;
    b[3] = SH_C2_0 * xy;
    b[4] = SH_C2_1 * yz;
    b[5] = SH_C2_2 * (2.f * zz - xx - yy);
    b[6] = SH_C2_3 * xz;
    b[7] = SH_C2_4 * (xx - yy);
    b[8] = SH_C3_0 * y * (3.f * xx - yy);
    b[9] = SH_C3_1 * xy * z;
    b[10] = SH_C3_2 * y * (4.f * zz - xx - yy);
    b[11] = SH_C3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy);
    b[12] = SH_C3_4 * x * (4.f * zz - xx - yy);
    b[13] = SH_C3_5 * z * (xx - yy);
    b[14] = SH_C3_6 * x * (xx - 3.f * yy);
}

// Each workgroup updates 1024 consecutive coefficients: the Gaussians they belong to get
// their basis values and colour gradient staged in LDS once, then every thread updates 4
// coefficients with 16-byte loads and stores of parameter, exp_avg and exp_avg_sq.
constexpr int SH_EPB = 1024;

// Moment storage of the SH groups: fp32, or bf16 (rounded to nearest even on store).
template <bool BF> struct Mom;
template <> struct Mom<false> {
    using T = float;
    static __device__ __forceinline__ void ld4(const float* p, size_t e, float o[4]) {
        const float4 a = reinterpret_cast<const float4*>(p)[e >> 2];
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    }
    static __device__ __forceinline__ void st4(float* p, size_t e, const float o[4]) {
        reinterpret_cast<float4*>(p)[e >> 2] = make_float4(o[0], o[1], o[2], o[3]);
    }
    static __device__ __forceinline__ float ld(const float* p, size_t e) { return p[e]; }
    static __device__ __forceinline__ void st(float* p, size_t e, float x) { p[e] = x; }
};
template <> struct Mom<true> {
    using T = uint16_t;
    static __device__ __forceinline__ void ld4(const uint16_t* p, size_t e, float o[4]) {
        const uint2 a = reinterpret_cast<const uint2*>(p)[e >> 2];
        o[0] = ms_bits_to_float(a.x << 16); o[1] = ms_bits_to_float(a.x & 0xFFFF0000u);
        o[2] = ms_bits_to_float(a.y << 16); o[3] = ms_bits_to_float(a.y & 0xFFFF0000u);
    }
    static __device__ __forceinline__ void st4(uint16_t* p, size_t e, const float o[4]) {
        uint2 a;
        a.x = (uint32_t)ms_bf16_bits(o[0]) | ((uint32_t)ms_bf16_bits(o[1]) << 16);
        a.y = (uint32_t)ms_bf16_bits(o[2]) | ((uint32_t)ms_bf16_bits(o[3]) << 16);
        reinterpret_cast<uint2*>(p)[e >> 2] = a;
    }
    static __device__ __forceinline__ float ld(const uint16_t* p, size_t e) { return ms_bf16_float(p[e]); }
    static __device__ __forceinline__ void st(uint16_t* p, size_t e, float x) { p[e] = ms_bf16_bits(x); }
};

template <int NB3, bool BF>
__global__ __launch_bounds__(256) void k_sh_rest_adam(int N, int nuse3, const float* __restrict__ means, float cpx,
                                                      float cpy, float cpz, const int32_t* __restrict__ radii,
                                                      const float* __restrict__ g2d, float* __restrict__ f_rest,
                                                      typename Mom<BF>::T* __restrict__ m4,
                                                      typename Mom<BF>::T* __restrict__ v4,
                                                      uint16_t* __restrict__ shadow, float step, float bc2s, float b1,
                                                      float b2, float eps) {
    constexpr int MAXG = SH_EPB / NB3 + 2;
    __shared__ float s_b[MAXG][16];  // 15 basis values + visibility flag
    __shared__ float s_vc[MAXG][3];
    const size_t total = (size_t)N * NB3;
    const size_t e0 = (size_t)blockIdx.x * SH_EPB;
    const int g0 = (int)(e0 / NB3);
    const size_t e_last = min(total, e0 + SH_EPB) - 1;
    const int ng = (int)(e_last / NB3) - g0 + 1;
    for (int k = threadIdx.x; k < ng; k += 256) {
        const int g = g0 + k;
        const float* vr = g2d + G2D * (size_t)g + G2D_RGB;
        const float vc0 = vr[0], vc1 = vr[1], vc2 = vr[2];
        // visible Gaussians with a finite colour gradient only (metal:1894, 1975-1978)
        const bool ok = radii[g] > 0 && ms_finite(vc0) && ms_finite(vc1) && ms_finite(vc2);
        s_b[k][15] = ok ? 1.f : 0.f;
        s_vc[k][0] = vc0; s_vc[k][1] = vc1; s_vc[k][2] = vc2;
        if (ok) {
            const float d0 = means[3 * (size_t)g] - cpx, d1 = means[3 * (size_t)g + 1] - cpy, d2 = means[3 * (size_t)g + 2] - cpz;
            const float inv = 1.f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
            sh_basis_all(d0 * inv, d1 * inv, d2 * inv, s_b[k]);
        }
    }
    __syncthreads();
    const size_t e = e0 + 4 * (size_t)threadIdx.x;
    if (e >= total) return;
    float p[4], m[4], v[4];
    const bool vec = e + 4 <= total;
    if (vec) {
        const float4 a = reinterpret_cast<const float4*>(f_rest)[e >> 2];
        p[0] = a.x; p[1] = a.y; p[2] = a.z; p[3] = a.w;
        Mom<BF>::ld4(m4, e, m);
        Mom<BF>::ld4(v4, e, v);
    } else {
        for (int q = 0; q < 4; q++) {
            const bool in = e + q < total;
            p[q] = in ? f_rest[e + q] : 0.f;
            m[q] = in ? Mom<BF>::ld(m4, e + q) : 0.f;
            v[q] = in ? Mom<BF>::ld(v4, e + q) : 0.f;
        }
    }
    bool touched = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const size_t el = e + q;
        if (el >= total) break;
        const int g = (int)(el / NB3);
        const int r = (int)(el - (size_t)g * NB3);
        const int k = g - g0;
        if (s_b[k][15] == 0.f || r >= nuse3) continue;
        const int b = r / 3, c = r - 3 * b;
        const float g_el = s_b[k][b] * s_vc[k][c];
        adam_el(p[q], m[q], v[q], g_el, step, b1, b2, bc2s, eps);
        touched = true;
    }
    if (!touched) return;  // nothing of these four elements was updated
    if (vec) {
        reinterpret_cast<float4*>(f_rest)[e >> 2] = make_float4(p[0], p[1], p[2], p[3]);
        Mom<BF>::st4(m4, e, m);
        Mom<BF>::st4(v4, e, v);
        if (BF) {
            uint2 h;
            h.x = (uint32_t)ms_bf16_bits(p[0]) | ((uint32_t)ms_bf16_bits(p[1]) << 16);
            h.y = (uint32_t)ms_bf16_bits(p[2]) | ((uint32_t)ms_bf16_bits(p[3]) << 16);
            reinterpret_cast<uint2*>(shadow)[e >> 2] = h;
        }
    } else {
        for (int q = 0; q < 4; q++)
            if (e + q < total) {
                f_rest[e + q] = p[q];
                Mom<BF>::st(m4, e + q, m[q]);
                Mom<BF>::st(v4, e + q, v[q]);
                if (BF) shadow[e + q] = ms_bf16_bits(p[q]);
            }
    }
}

void sh_rest_adam(hipStream_t s, const Params& P, const MsplatCamera& cam, const int32_t* radii, const float* g2d,
                  const MsplatStepParams& hp) {
    const int du = hp.degrees_to_use < P.degree ? hp.degrees_to_use : P.degree;
    if (du < 1 || P.N <= 0) return;
    const int nuse3 = (du >= 3 ? 15 : du == 2 ? 8 : 3) * 3;
    const int nb3 = (P.degree == 1 ? 3 : P.degree == 2 ? 8 : 15) * 3;
    const size_t total = (size_t)P.N * nb3;
#define LAUNCH_SH(NB, BF, MT)                                                                                      \
    hipLaunchKernelGGL((k_sh_rest_adam<NB, BF>), dim3(cdiv(total, SH_EPB)), dim3(256), 0, s, P.N, nuse3, P.means,  \
                       cam.cam_pos[0], cam.cam_pos[1], cam.cam_pos[2], radii, g2d, P.f_rest, (MT*)P.m[4],          \
                       (MT*)P.v[4], P.f_rest_h, hp.step_size[4], hp.bc2_sqrt[4], hp.beta1, hp.beta2, hp.eps)
    if (P.sh_bf16) {
        if (nb3 == 9) LAUNCH_SH(9, true, uint16_t);
        else if (nb3 == 24) LAUNCH_SH(24, true, uint16_t);
        else LAUNCH_SH(45, true, uint16_t);
    } else {
        if (nb3 == 9) LAUNCH_SH(9, false, float);
        else if (nb3 == 24) LAUNCH_SH(24, false, float);
        else LAUNCH_SH(45, false, float);
    }
#undef LAUNCH_SH
}

// ------------------------------------------------------------------------------------
// A17 (projection part) + A18 + A19 + A20/A21, one thread per Gaussian: projection VJP
// (metal:1860-1958), DC Adam (metal:1980-1985), regulariser gradients (metal:3987-4026),
// Adam on means / log-scales / quaternions / opacity (metal:1680-1713) and the SGLD
// perturbation (metal:3871-3967) of the updated parameters.
struct M3 {
    float m[3][3];
};
__device__ __forceinline__ M3 mm(const M3& a, const M3& b) {
    M3 o;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) o.m[r][c] = a.m[r][0] * b.m[0][c] + a.m[r][1] * b.m[1][c] + a.m[r][2] * b.m[2][c];
    return o;
}
__device__ __forceinline__ M3 tr(const M3& a) {
    M3 o;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) o.m[r][c] = a.m[c][r];
    return o;
}

__device__ void proj_vjp(const float p[3], const float ls[3], const float q[4], const MsplatCamera& cam,
                         const float conic[3], float vxy0, float vxy1, const float vcn[3], float v_mean[3],
                         float v_scale[3], float v_quat[4]) {
    const float* pm = cam.projmat;
    const float* vm = cam.viewmat;
    const float W = (float)cam.width, H = (float)cam.height;
    // project_pix_vjp (metal:1252-1270), perspective divide held fixed
    const float hw = pm[12] * p[0] + pm[13] * p[1] + pm[14] * p[2] + pm[15];
    const float rw = 1.f / (hw + 1e-6f);
    const float vp0 = 0.5f * W * vxy0 * rw, vp1 = 0.5f * H * vxy1 * rw;
    v_mean[0] = pm[0] * vp0 + pm[4] * vp1;
    v_mean[1] = pm[1] * vp0 + pm[5] * vp1;
    v_mean[2] = pm[2] * vp0 + pm[6] * vp1;
    // cov2d_to_conic_vjp (metal:1289-1300)
    const float X00 = conic[0], X01 = conic[1], X11 = conic[2];
    const float G00 = vcn[0], G01 = vcn[1], G11 = vcn[2];
    const float P00 = -(X00 * G00 + X01 * G01), P01 = -(X00 * G01 + X01 * G11);
    const float P10 = -(X01 * G00 + X11 * G01), P11 = -(X01 * G01 + X11 * G11);
    const float Q00 = P00 * X00 + P01 * X01, Q01 = P00 * X01 + P01 * X11;
    const float Q10 = P10 * X00 + P11 * X01, Q11 = P10 * X01 + P11 * X11;
    const float vc2_0 = Q00, vc2_1 = Q01 + Q10, vc2_2 = Q11;
    // cov3d from scale / rotation
    float s[3];
#pragma unroll
    for (int k = 0; k < 3; k++) s[k] = ms_expf_core(fminf(fmaxf(ls[k], -15.f), 15.f));
    float R[3][3];
    quat_rot(q, R);
    M3 Mm;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) Mm.m[r][c] = R[r][c] * s[c];
    M3 V = mm(Mm, tr(Mm));
    // project_cov3d_ewa_vjp (metal:1375-1443)
    const float fx = cam.fx, fy = cam.fy;
    const float tan_fovx = 0.5f * W / fx, tan_fovy = 0.5f * H / fy;
    const float pvx0 = vm[0] * p[0] + vm[1] * p[1] + vm[2] * p[2] + vm[3];
    const float pvy0 = vm[4] * p[0] + vm[5] * p[1] + vm[6] * p[2] + vm[7];
    const float pvz = vm[8] * p[0] + vm[9] * p[1] + vm[10] * p[2] + vm[11];
    const float lim_x = 1.3f * tan_fovx, lim_y = 1.3f * tan_fovy;
    const float pvx = pvz * fminf(lim_x, fmaxf(-lim_x, pvx0 / pvz));
    const float pvy = pvz * fminf(lim_y, fmaxf(-lim_y, pvy0 / pvz));
    const float rz = 1.f / pvz, rz2 = rz * rz;
    M3 Wm, J, vc;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) Wm.m[r][c] = vm[r * 4 + c];
    J.m[0][0] = fx * rz; J.m[0][1] = 0.f; J.m[0][2] = -fx * pvx * rz2;
    J.m[1][0] = 0.f; J.m[1][1] = fy * rz; J.m[1][2] = -fy * pvy * rz2;
    J.m[2][0] = 0.f; J.m[2][1] = 0.f; J.m[2][2] = 0.f;
    vc.m[0][0] = vc2_0; vc.m[0][1] = 0.5f * vc2_1; vc.m[0][2] = 0.f;
    vc.m[1][0] = 0.5f * vc2_1; vc.m[1][1] = vc2_2; vc.m[1][2] = 0.f;
    vc.m[2][0] = 0.f; vc.m[2][1] = 0.f; vc.m[2][2] = 0.f;
    const M3 Tm = mm(J, Wm);
    const M3 vV = mm(mm(tr(Tm), vc), Tm);
    const M3 vT1 = mm(mm(vc, Tm), V);   // V symmetric: V^T == V
    const M3 vT2 = mm(mm(tr(vc), Tm), V);
    M3 vT;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) vT.m[r][c] = vT1.m[r][c] + vT2.m[r][c];
    const float vcov3[6] = {vV.m[0][0], vV.m[1][0] + vV.m[0][1], vV.m[2][0] + vV.m[0][2],
                            vV.m[1][1], vV.m[2][1] + vV.m[1][2], vV.m[2][2]};
    const M3 vJ = mm(vT, tr(Wm));
    const float fx_rz2 = fx * rz2, fy_rz2 = fy * rz2, rz3 = rz2 * rz;
    const float vt0 = -fx_rz2 * vJ.m[0][2], vt1 = -fy_rz2 * vJ.m[1][2];
    const float vt2 = -fx_rz2 * vJ.m[0][0] + 2.f * fx * pvx * rz3 * vJ.m[0][2] - fy_rz2 * vJ.m[1][1] +
                      2.f * fy * pvy * rz3 * vJ.m[1][2];
#pragma unroll
    for (int k = 0; k < 3; k++) v_mean[k] += vt0 * Wm.m[0][k] + vt1 * Wm.m[1][k] + vt2 * Wm.m[2][k];
    // scale_rot_to_cov3d_vjp (metal:1536-1569) + quat_to_rotmat_vjp (metal:1445-1490)
    M3 vVs;
    vVs.m[0][0] = 2.f * vcov3[0]; vVs.m[0][1] = 2.f * (0.5f * vcov3[1]); vVs.m[0][2] = 2.f * (0.5f * vcov3[2]);
    vVs.m[1][0] = 2.f * (0.5f * vcov3[1]); vVs.m[1][1] = 2.f * vcov3[3]; vVs.m[1][2] = 2.f * (0.5f * vcov3[4]);
    vVs.m[2][0] = 2.f * (0.5f * vcov3[2]); vVs.m[2][1] = 2.f * (0.5f * vcov3[4]); vVs.m[2][2] = 2.f * vcov3[5];
    const M3 vM = mm(vVs, Mm);
#pragma unroll
    for (int k = 0; k < 3; k++) v_scale[k] = R[0][k] * vM.m[0][k] + R[1][k] * vM.m[1][k] + R[2][k] * vM.m[2][k];
    float g[3][3];  // g[c][r] = column-major v_R
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) g[c][r] = vM.m[r][c] * s[c];
    const float n2 = q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const float sn = 1.f / sqrtf(fmaxf(n2, 1e-8f));
    const float w = q[0] * sn, x = q[1] * sn, y = q[2] * sn, z = q[3] * sn;
    v_quat[0] = 2.f * (x * (g[1][2] - g[2][1]) + y * (g[2][0] - g[0][2]) + z * (g[0][1] - g[1][0]));
    v_quat[1] = 2.f * (-2.f * x * (g[1][1] + g[2][2]) + y * (g[0][1] + g[1][0]) + z * (g[0][2] + g[2][0]) +
                       w * (g[1][2] - g[2][1]));
    v_quat[2] = 2.f * (x * (g[0][1] + g[1][0]) - 2.f * y * (g[0][0] + g[2][2]) + z * (g[1][2] + g[2][1]) +
                       w * (g[2][0] - g[0][2]));
    v_quat[3] = 2.f * (x * (g[0][2] + g[2][0]) + y * (g[1][2] + g[2][1]) - 2.f * z * (g[0][0] + g[1][1]) +
                       w * (g[0][1] - g[1][0]));
#pragma unroll
    for (int k = 0; k < 3; k++) v_scale[k] *= s[k];
}

__device__ __forceinline__ uint32_t pcg_hash32(uint32_t seed, uint32_t idx) {
    uint32_t x = seed * 747796405u + idx * 2891336453u;
    x = ((x >> ((x >> 28) + 4)) ^ x) * 277803737u;
    return (x >> 22) ^ x;
}
__device__ __forceinline__ float pcg_unit(uint32_t h) { return ((float)(h >> 9) + 0.5f) * (1.0f / 8388608.0f); }

// regularizer gradients (mcmc_reg_grad_kernel, metal:3987-4026); vdc = sum of the DC exp_avg_sq
__device__ __forceinline__ void reg_grads(int N, float opac_i, const float ls[3], float vdc, const MsplatStepParams& hp,
                                          float& vo, float vs[3]) {
    const float duty = (vdc == 0.0f) ? 1.0f : hp.reg_invis_scale;
    const float invN = duty / (float)N;
    const float sig = 1.0f / (1.0f + ms_expf(-opac_i));
    vo += hp.op_reg * sig * (1.0f - sig) * invN;
    const float inv3N = invN * (1.0f / 3.0f);
#pragma unroll
    for (int k = 0; k < 3; k++) vs[k] += hp.sc_reg * ms_expf_core(fminf(fmaxf(ls[k], -15.f), 15.f)) * inv3N;
}

// Adam on means / log-scales / quaternions / opacity with the non-finite gradient gate, then
// SGLD on the updated parameters, then write-back of row i.
// Adam moments of groups 0, 1, 2, 5 and the opacity logit of one Gaussian, loaded at the
// start of an update kernel so that their latency overlaps the gradient computation.
struct TailState {
    float m0[3], v0[3], m1[3], v1[3], m2[4], v2[4], m5, v5, o;
};

__device__ __forceinline__ void tail_load(int i, const float* __restrict__ opac, const float* __restrict__ m0,
                                          const float* __restrict__ v0, const float* __restrict__ m1,
                                          const float* __restrict__ v1, const float* __restrict__ m2,
                                          const float* __restrict__ v2, const float* __restrict__ m5,
                                          const float* __restrict__ v5, TailState& t) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const size_t e = 3 * (size_t)i + k;
        t.m0[k] = m0[e]; t.v0[k] = v0[e];
        t.m1[k] = m1[e]; t.v1[k] = v1[e];
    }
    const float4 a = reinterpret_cast<const float4*>(m2)[i], b = reinterpret_cast<const float4*>(v2)[i];
    t.m2[0] = a.x; t.m2[1] = a.y; t.m2[2] = a.z; t.m2[3] = a.w;
    t.v2[0] = b.x; t.v2[1] = b.y; t.v2[2] = b.z; t.v2[3] = b.w;
    t.m5 = m5[i];
    t.v5 = v5[i];
    t.o = opac[i];
}

__device__ __forceinline__ void adam_sgld_tail(int i, float p[3], float ls[3], float q[4], const float vm[3],
                                               const float vs[3], const float vq[4], float vo,
                                               float* __restrict__ means, float* __restrict__ scales,
                                               float* __restrict__ quats, float* __restrict__ opac,
                                               float* __restrict__ m0, float* __restrict__ v0, float* __restrict__ m1,
                                               float* __restrict__ v1, float* __restrict__ m2, float* __restrict__ v2,
                                               float* __restrict__ m5, float* __restrict__ v5, TailState& t,
                                               const MsplatStepParams& hp, uint32_t* __restrict__ nanc) {
    const float b1 = hp.beta1, b2 = hp.beta2, eps = hp.eps;
    // fused_adam (metal:1680-1713) with the non-finite gradient gate, groups 0, 1, 2, 5
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const size_t e = 3 * (size_t)i + k;
        if (ms_finite(vm[k])) {
            adam_el(p[k], t.m0[k], t.v0[k], vm[k], hp.step_size[0], b1, b2, hp.bc2_sqrt[0], eps);
            m0[e] = t.m0[k]; v0[e] = t.v0[k];
        } else atomicAdd(&nanc[0], 1u);
        if (ms_finite(vs[k])) {
            adam_el(ls[k], t.m1[k], t.v1[k], vs[k], hp.step_size[1], b1, b2, hp.bc2_sqrt[1], eps);
            m1[e] = t.m1[k]; v1[e] = t.v1[k];
        } else atomicAdd(&nanc[1], 1u);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t e = 4 * (size_t)i + k;
        if (ms_finite(vq[k])) {
            adam_el(q[k], t.m2[k], t.v2[k], vq[k], hp.step_size[2], b1, b2, hp.bc2_sqrt[2], eps);
            m2[e] = t.m2[k]; v2[e] = t.v2[k];
        } else atomicAdd(&nanc[2], 1u);
    }
    float o = t.o;
    if (ms_finite(vo)) {
        adam_el(o, t.m5, t.v5, vo, hp.step_size[5], b1, b2, hp.bc2_sqrt[5], eps);
        m5[i] = t.m5; v5[i] = t.v5;
    } else atomicAdd(&nanc[5], 1u);
    // SGLD (sgld_noise_gen metal:3871-3890, sgld_noise metal:3892-3967) on the updated state
    if (hp.sgld_enable) {
        const uint32_t base = (uint32_t)i * 4u;
        const float u0 = pcg_unit(pcg_hash32(hp.sgld_seed, base)), u1 = pcg_unit(pcg_hash32(hp.sgld_seed, base + 1));
        const float u2 = pcg_unit(pcg_hash32(hp.sgld_seed, base + 2)), u3 = pcg_unit(pcg_hash32(hp.sgld_seed, base + 3));
        const float r01 = sqrtf(-2.0f * ms_logf(u0)), r23 = sqrtf(-2.0f * ms_logf(u2));
        float s01, c01, s23, c23;
        ms_sincosf(6.283185307179586f * u1, &s01, &c01);
        ms_sincosf(6.283185307179586f * u3, &s23, &c23);
        const float nz0 = r01 * c01, nz1 = r01 * s01, nz2 = r23 * c23;
        const float op_sig = 1.0f / (1.0f + ms_expf(-o));
        const float weight = 1.0f / (1.0f + ms_expf(-100.0f * ((1.0f - op_sig) - 0.995f)));
        const float ns = hp.noise_lr * hp.xyz_lr * weight;
        const float n0 = nz0 * ns, n1 = nz1 * ns, n2 = nz2 * ns;
        float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
        const float ql = sqrtf(fmaxf(qw * qw + qx * qx + qy * qy + qz * qz, 1e-8f));
        qw /= ql; qx /= ql; qy /= ql; qz /= ql;
        const float sx = ms_expf_core(fminf(fmaxf(ls[0], -15.f), 15.f));
        const float sy = ms_expf_core(fminf(fmaxf(ls[1], -15.f), 15.f));
        const float sz = ms_expf_core(fminf(fmaxf(ls[2], -15.f), 15.f));
        float rt0 = (1 - 2 * (qy * qy + qz * qz)) * n0 + 2 * (qx * qy + qw * qz) * n1 + 2 * (qx * qz - qw * qy) * n2;
        float rt1 = 2 * (qx * qy - qw * qz) * n0 + (1 - 2 * (qx * qx + qz * qz)) * n1 + 2 * (qy * qz + qw * qx) * n2;
        float rt2 = 2 * (qx * qz + qw * qy) * n0 + 2 * (qy * qz - qw * qx) * n1 + (1 - 2 * (qx * qx + qy * qy)) * n2;
        rt0 *= sx * sx; rt1 *= sy * sy; rt2 *= sz * sz;
        float d0 = (1 - 2 * (qy * qy + qz * qz)) * rt0 + 2 * (qx * qy - qw * qz) * rt1 + 2 * (qx * qz + qw * qy) * rt2;
        float d1 = 2 * (qx * qy + qw * qz) * rt0 + (1 - 2 * (qx * qx + qz * qz)) * rt1 + 2 * (qy * qz - qw * qx) * rt2;
        float d2 = 2 * (qx * qz - qw * qy) * rt0 + 2 * (qy * qz + qw * qx) * rt1 + (1 - 2 * (qx * qx + qy * qy)) * rt2;
        if (ms_finite(d0) && ms_finite(d1) && ms_finite(d2)) {
            const float len2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (ms_finite(len2)) {
                if (len2 > hp.max_disp * hp.max_disp) {
                    const float inv = hp.max_disp / sqrtf(len2);
                    d0 *= inv; d1 *= inv; d2 *= inv;
                }
                p[0] += d0; p[1] += d1; p[2] += d2;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) { means[3 * (size_t)i + k] = p[k]; scales[3 * (size_t)i + k] = ls[k]; }
    reinterpret_cast<float4*>(quats)[i] = make_float4(q[0], q[1], q[2], q[3]);
    opac[i] = o;
}

template <bool BF>
__global__ __launch_bounds__(256) void k_gaussian_update(
    int N, float* __restrict__ means, float* __restrict__ scales, float* __restrict__ quats, float* __restrict__ f_dc,
    float* __restrict__ opac, float* __restrict__ m0, float* __restrict__ v0, float* __restrict__ m1,
    float* __restrict__ v1, float* __restrict__ m2, float* __restrict__ v2, typename Mom<BF>::T* __restrict__ m3,
    typename Mom<BF>::T* __restrict__ v3, uint16_t* __restrict__ dc_shadow, float* __restrict__ m5,
    float* __restrict__ v5, MsplatCamera cam,
    const int32_t* __restrict__ radii, const float4* __restrict__ rA, const float4* __restrict__ rB, GradBufs G,
    MsplatStepParams hp, uint32_t* __restrict__ nanc, float* __restrict__ dbg_vm, float* __restrict__ dbg_vs,
    float* __restrict__ dbg_vq, float* __restrict__ dbg_vo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float p[3] = {means[3 * (size_t)i], means[3 * (size_t)i + 1], means[3 * (size_t)i + 2]};
    float ls[3] = {scales[3 * (size_t)i], scales[3 * (size_t)i + 1], scales[3 * (size_t)i + 2]};
    const float4 q4 = reinterpret_cast<const float4*>(quats)[i];
    float q[4] = {q4.x, q4.y, q4.z, q4.w};
    TailState ts;
    tail_load(i, opac, m0, v0, m1, v1, m2, v2, m5, v5, ts);
    // Every row of the gradient and projection records is loaded before the visibility test
    // (all are sized N), so their latency overlaps instead of following the radius load.
    const int rad = radii[i];
    const float4 a = rA[i], b = rB[i];
    const float* gr = G.g + G2D * (size_t)i;
    const float2 vxy = make_float2(gr[G2D_XY], gr[G2D_XY + 1]);
    const float vcn[3] = {gr[G2D_CONIC], gr[G2D_CONIC + 1], gr[G2D_CONIC + 2]};
    const float vc[3] = {gr[G2D_RGB], gr[G2D_RGB + 1], gr[G2D_RGB + 2]};
    float dc[3], dm[3], dv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const size_t k = 3 * (size_t)i + c;
        dc[c] = f_dc[k]; dm[c] = Mom<BF>::ld(m3, k); dv[c] = Mom<BF>::ld(v3, k);
    }
    float vo = gr[G2D_OPAC];
    float vm[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    const float b1 = hp.beta1, b2 = hp.beta2, eps = hp.eps;
    if (rad > 0) {
        const float conic[3] = {a.w, b.x, b.y};
        proj_vjp(p, ls, q, cam, conic, vxy.x, vxy.y, vcn, vm, vs, vq);
        if (ms_finite(vc[0]) && ms_finite(vc[1]) && ms_finite(vc[2])) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const size_t k = 3 * (size_t)i + c;
                adam_el(dc[c], dm[c], dv[c], SH_C0 * vc[c], hp.step_size[3], b1, b2, hp.bc2_sqrt[3], eps);
                f_dc[k] = dc[c];
                Mom<BF>::st(m3, k, dm[c]);
                Mom<BF>::st(v3, k, dv[c]);
                if (BF) {
                    dc_shadow[k] = ms_bf16_bits(dc[c]);
                    dv[c] = ms_bf16_round(dv[c]);  // the value the store holds
                }
            }
        } else {
            atomicAdd(&nanc[3], 1u);
        }
    }
    if (hp.op_reg > 0.0f || hp.sc_reg > 0.0f) {
        const float vdc = dv[0] + dv[1] + dv[2];  // second moments of the colour after this step
        reg_grads(N, ts.o, ls, vdc, hp, vo, vs);
    }
    if (dbg_vm) {
#pragma unroll
        for (int k = 0; k < 3; k++) { dbg_vm[3 * (size_t)i + k] = vm[k]; dbg_vs[3 * (size_t)i + k] = vs[k]; }
#pragma unroll
        for (int k = 0; k < 4; k++) dbg_vq[4 * (size_t)i + k] = vq[k];
        dbg_vo[i] = vo;
    }
    adam_sgld_tail(i, p, ls, q, vm, vs, vq, vo, means, scales, quats, opac, m0, v0, m1, v1, m2, v2, m5, v5, ts, hp,
                   nanc);
}

void gaussian_update(hipStream_t s, const Params& P, const MsplatCamera& cam, const ProjOut& pj, const GradBufs& G,
                     const MsplatStepParams& hp, uint32_t* nan_counts, float* dbg_v_mean, float* dbg_v_scale,
                     float* dbg_v_quat, float* dbg_v_opac) {
    if (P.N <= 0) return;
    if (P.sh_bf16)
        hipLaunchKernelGGL(k_gaussian_update<true>, dim3(cdiv(P.N, 256)), dim3(256), 0, s, P.N, P.means, P.scales,
                           P.quats, P.f_dc, P.opac, P.m[0], P.v[0], P.m[1], P.v[1], P.m[2], P.v[2], (uint16_t*)P.m[3],
                           (uint16_t*)P.v[3], P.f_dc_h, P.m[5], P.v[5], cam, pj.radii, pj.rec.A, pj.rec.B, G, hp,
                           nan_counts, dbg_v_mean, dbg_v_scale, dbg_v_quat, dbg_v_opac);
    else
        hipLaunchKernelGGL(k_gaussian_update<false>, dim3(cdiv(P.N, 256)), dim3(256), 0, s, P.N, P.means, P.scales,
                           P.quats, P.f_dc, P.opac, P.m[0], P.v[0], P.m[1], P.v[1], P.m[2], P.v[2], P.m[3], P.v[3],
                           (uint16_t*)nullptr, P.m[5], P.v[5], cam, pj.radii, pj.rec.A, pj.rec.B, G, hp, nan_counts,
                           dbg_v_mean, dbg_v_scale, dbg_v_quat, dbg_v_opac);
}

__global__ void k_f32_to_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = ms_bf16_bits(src[i]);
}
__global__ void k_bf16_to_f32(const uint16_t* __restrict__ src, float* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = ms_bf16_float(src[i]);
}

void f32_to_bf16(hipStream_t s, const float* src, uint16_t* dst, size_t n) {
    if (n) hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}
void bf16_to_f32(hipStream_t s, const uint16_t* src, float* dst, size_t n) {
    if (n) hipLaunchKernelGGL(k_bf16_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

void sh_refresh_shadows(hipStream_t s, const Params& P) {
    if (!P.sh_bf16 || P.N <= 0) return;
    const int nb3 = (P.degree <= 0 ? 0 : (P.degree == 1 ? 3 : P.degree == 2 ? 8 : 15)) * 3;
    f32_to_bf16(s, P.f_dc, P.f_dc_h, (size_t)P.N * 3);
    f32_to_bf16(s, P.f_rest, P.f_rest_h, (size_t)P.N * nb3);
}


// ------------------------------------------------------------------------------------
// Multi-view batched step (this build's addition; the reference trains one view per step).
template <int NB3>
__global__ __launch_bounds__(256) void k_view_accum(int N, int nuse3, const float* __restrict__ means,
                                                    const float* __restrict__ scales, const float* __restrict__ quats,
                                                    const ViewConsts* __restrict__ vc,
                                                    const int32_t* __restrict__ radii,
                                                    const float4* __restrict__ rA, const float4* __restrict__ rB,
                                                    GradBufs G, BatchGrads A, uint32_t* __restrict__ nanc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || radii[i] <= 0) return;
    const MsplatCamera cam = vc->cam;
    const float p[3] = {means[3 * (size_t)i], means[3 * (size_t)i + 1], means[3 * (size_t)i + 2]};
    const float ls[3] = {scales[3 * (size_t)i], scales[3 * (size_t)i + 1], scales[3 * (size_t)i + 2]};
    const float4 q4 = reinterpret_cast<const float4*>(quats)[i];
    const float q[4] = {q4.x, q4.y, q4.z, q4.w};
    const float4 a = rA[i], b = rB[i];
    const float conic[3] = {a.w, b.x, b.y};
    const float* gr = G.g + G2D * (size_t)i;
    const float2 vxy = make_float2(gr[G2D_XY], gr[G2D_XY + 1]);
    const float vcn[3] = {gr[G2D_CONIC], gr[G2D_CONIC + 1], gr[G2D_CONIC + 2]};
    float vm[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    proj_vjp(p, ls, q, cam, conic, vxy.x, vxy.y, vcn, vm, vs, vq);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        A.v_mean[3 * (size_t)i + k] += vm[k];
        A.v_scale[3 * (size_t)i + k] += vs[k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) A.v_quat[4 * (size_t)i + k] += vq[k];
    A.v_opac[i] += gr[G2D_OPAC];
    const float vcol[3] = {gr[G2D_RGB], gr[G2D_RGB + 1], gr[G2D_RGB + 2]};
    if (!(ms_finite(vcol[0]) && ms_finite(vcol[1]) && ms_finite(vcol[2]))) {
        atomicAdd(&nanc[3], 1u);
        return;
    }
    A.vis[i] = 1u;
#pragma unroll
    for (int c = 0; c < 3; c++) A.v_dc[3 * (size_t)i + c] += SH_C0 * vcol[c];
    if (NB3 > 0 && nuse3 > 0) {
        const float d0 = p[0] - cam.cam_pos[0], d1 = p[1] - cam.cam_pos[1], d2 = p[2] - cam.cam_pos[2];
        const float inv = 1.f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
        float bs[16];
        sh_basis_all(d0 * inv, d1 * inv, d2 * inv, bs);
        float* row = A.v_rest + (size_t)NB3 * i;
        for (int r = 0; r < nuse3; r++) row[r] += bs[r / 3] * vcol[r % 3];
    }
}

template <bool BF, int NB3>
__global__ __launch_bounds__(256) void k_batch_update(
    int N, int nuse3, float* __restrict__ means, float* __restrict__ scales, float* __restrict__ quats,
    float* __restrict__ f_dc, float* __restrict__ f_rest, float* __restrict__ opac, float* __restrict__ m0,
    float* __restrict__ v0, float* __restrict__ m1, float* __restrict__ v1, float* __restrict__ m2,
    float* __restrict__ v2, typename Mom<BF>::T* __restrict__ m3, typename Mom<BF>::T* __restrict__ v3,
    typename Mom<BF>::T* __restrict__ m4, typename Mom<BF>::T* __restrict__ v4, uint16_t* __restrict__ dc_shadow,
    uint16_t* __restrict__ rest_shadow, float* __restrict__ m5, float* __restrict__ v5, BatchGrads A,
    const MsplatStepParams* __restrict__ hpd, uint32_t* __restrict__ nanc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const MsplatStepParams hp = *hpd;
    float p[3] = {means[3 * (size_t)i], means[3 * (size_t)i + 1], means[3 * (size_t)i + 2]};
    float ls[3] = {scales[3 * (size_t)i], scales[3 * (size_t)i + 1], scales[3 * (size_t)i + 2]};
    const float4 q4 = reinterpret_cast<const float4*>(quats)[i];
    float q[4] = {q4.x, q4.y, q4.z, q4.w};
    TailState ts;
    tail_load(i, opac, m0, v0, m1, v1, m2, v2, m5, v5, ts);
    float vm[3], vs[3], vq[4];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        vm[k] = A.v_mean[3 * (size_t)i + k];
        vs[k] = A.v_scale[3 * (size_t)i + k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) vq[k] = A.v_quat[4 * (size_t)i + k];
    float vo = A.v_opac[i];
    const float b1 = hp.beta1, b2 = hp.beta2, eps = hp.eps;
    if (A.vis[i]) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const size_t k = 3 * (size_t)i + c;
            float pp = f_dc[k], mm_ = Mom<BF>::ld(m3, k), vv = Mom<BF>::ld(v3, k);
            adam_el(pp, mm_, vv, A.v_dc[k], hp.step_size[3], b1, b2, hp.bc2_sqrt[3], eps);
            f_dc[k] = pp;
            Mom<BF>::st(m3, k, mm_);
            Mom<BF>::st(v3, k, vv);
            if (BF) dc_shadow[k] = ms_bf16_bits(pp);
        }
        for (int r = 0; r < nuse3; r++) {
            const size_t k = (size_t)NB3 * i + r;
            float pp = f_rest[k], mm_ = Mom<BF>::ld(m4, k), vv = Mom<BF>::ld(v4, k);
            adam_el(pp, mm_, vv, A.v_rest[k], hp.step_size[4], b1, b2, hp.bc2_sqrt[4], eps);
            f_rest[k] = pp;
            Mom<BF>::st(m4, k, mm_);
            Mom<BF>::st(v4, k, vv);
            if (BF) rest_shadow[k] = ms_bf16_bits(pp);
        }
    }
    if (hp.op_reg > 0.0f || hp.sc_reg > 0.0f) {
        const float vdc = Mom<BF>::ld(v3, 3 * (size_t)i) + Mom<BF>::ld(v3, 3 * (size_t)i + 1) +
                          Mom<BF>::ld(v3, 3 * (size_t)i + 2);
        reg_grads(N, ts.o, ls, vdc, hp, vo, vs);
    }
    adam_sgld_tail(i, p, ls, q, vm, vs, vq, vo, means, scales, quats, opac, m0, v0, m1, v1, m2, v2, m5, v5, ts, hp,
                   nanc);
}

static int nuse3_of(const Params& P, int degrees_to_use) {
    const int du = degrees_to_use < P.degree ? degrees_to_use : P.degree;
    return (du >= 3 ? 15 : du == 2 ? 8 : du == 1 ? 3 : 0) * 3;
}

void view_accumulate(hipStream_t s, const Params& P, const ViewConsts* vc, int degrees_to_use, const ProjOut& pj,
                     const GradBufs& G, BatchGrads& A, uint32_t* nan_counts) {
    if (P.N <= 0) return;
    const int nuse3 = nuse3_of(P, degrees_to_use);
    const dim3 grid(cdiv(P.N, 256));
#define LAUNCH_ACC(NB) \
    hipLaunchKernelGGL(k_view_accum<NB>, grid, dim3(256), 0, s, P.N, nuse3, P.means, P.scales, P.quats, vc, pj.radii, \
                       pj.rec.A, pj.rec.B, G, A, nan_counts)
    if (P.degree <= 0) LAUNCH_ACC(0);
    else if (P.degree == 1) LAUNCH_ACC(9);
    else if (P.degree == 2) LAUNCH_ACC(24);
    else LAUNCH_ACC(45);
#undef LAUNCH_ACC
}

void batch_update(hipStream_t s, const Params& P, const BatchGrads& A, const MsplatStepParams* hpd, int degrees_to_use,
                  uint32_t* nan_counts) {
    if (P.N <= 0) return;
    const int nuse3 = nuse3_of(P, degrees_to_use);
    const dim3 grid(cdiv(P.N, 256));
#define LAUNCH_UPD(BF, MT, NB)                                                                                      \
    hipLaunchKernelGGL((k_batch_update<BF, NB>), grid, dim3(256), 0, s, P.N, nuse3, P.means, P.scales, P.quats,     \
                       P.f_dc, P.f_rest, P.opac, P.m[0], P.v[0], P.m[1], P.v[1], P.m[2], P.v[2], (MT*)P.m[3],       \
                       (MT*)P.v[3], (MT*)P.m[4], (MT*)P.v[4], P.f_dc_h, P.f_rest_h, P.m[5], P.v[5], A, hpd, nan_counts)
#define LAUNCH_UPD_NB(BF, MT)                        \
    if (P.degree <= 0) LAUNCH_UPD(BF, MT, 0);        \
    else if (P.degree == 1) LAUNCH_UPD(BF, MT, 9);   \
    else if (P.degree == 2) LAUNCH_UPD(BF, MT, 24);  \
    else LAUNCH_UPD(BF, MT, 45);
    if (P.sh_bf16) {
        LAUNCH_UPD_NB(true, uint16_t)
    } else {
        LAUNCH_UPD_NB(false, float)
    }
#undef LAUNCH_UPD_NB
#undef LAUNCH_UPD
}

}  // namespace msk

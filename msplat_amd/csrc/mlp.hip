// mlp.hip — relocation-proposal network on the matrix cores (gfx950, wave64).
//
// Every wave evaluates 16 Gaussians at a time with v_mfma_f32_16x16x16_bf16. Activations are
// kept transposed ([unit][Gaussian]) so that the accumulator layout of one layer (lane l
// holds rows 4 (l / 16) + q, column l % 16 of its 16x16 tile) is exactly the B-operand
// layout of the next: the forward chain never leaves registers. The weight fragments (A
// operands) of all three layers stay in registers for the whole launch. Weight gradients
// sum over Gaussians, which is the MFMA reduction dimension: each wave keeps them in its
// accumulators across its groups and adds them to the global gradient once; only their
// operands need a transpose, through a wave-private LDS tile.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "mlp.h"
#include "ms_math.h"

namespace msk {

typedef short bf16x4 __attribute__((vector_size(8)));
typedef float f32x4 __attribute__((vector_size(16)));

static inline int cdiv_i(long long a, int b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ short bfs(float x) { return (short)ms_bf16_bits(x); }
__device__ __forceinline__ float bff(short h) { return ms_bf16_float((uint16_t)h); }

// GELU, tanh form, and its derivative.
__device__ __forceinline__ float gelu(float x) {
    const float u = 0.7978845608f * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.f + tanhf(u));
}
__device__ __forceinline__ float gelu_d(float x) {
    const float u = 0.7978845608f * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * 0.7978845608f * (1.f + 0.134145f * x * x);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------------------
// statistics and features

__global__ void k_relo_stats(int N, const int32_t* __restrict__ radii, const float* __restrict__ depths,
                             const float* __restrict__ g2d, ReloStats st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int r = radii[i];
    if (r <= 0) return;
    const float2 g = make_float2(g2d[G2D * (size_t)i + G2D_XY], g2d[G2D * (size_t)i + G2D_XY + 1]);
    const float d = depths[i];
    st.vis[i] += 1.f;
    st.gsum[i] += sqrtf(g.x * g.x + g.y * g.y);
    st.maxr[i] = fmaxf(st.maxr[i], (float)r);
    st.dsum[i] += d;
    st.dsq[i] += d * d;
}

void relo_stats_accum(hipStream_t s, int N, const int32_t* radii, const float* depths, const float* g2d,
                      ReloStats st) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_relo_stats, dim3(cdiv_i(N, 256)), dim3(256), 0, s, N, radii, depths, g2d, st);
}

__global__ void k_relo_features(int N, const float* __restrict__ scales, const float* __restrict__ quats,
                                const float* __restrict__ opac, const float* __restrict__ f_dc, ReloStats st,
                                float inv_views, uint16_t* __restrict__ X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float f[MLP_IN];
    f[0] = 1.f / (1.f + ms_expf(-opac[i]));
    for (int k = 0; k < 3; k++) f[1 + k] = scales[3 * (size_t)i + k];
    const float4 q = reinterpret_cast<const float4*>(quats)[i];
    const float qn = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    const float iq = qn > 0.f ? 1.f / qn : 0.f;
    f[4] = q.x * iq; f[5] = q.y * iq; f[6] = q.z * iq; f[7] = q.w * iq;
    const float vis = st.vis[i];
    const float iv = vis > 0.f ? 1.f / vis : 0.f;
    const float gn = st.gsum[i] * iv;
    f[8] = log10f(gn + 1e-10f) * 0.2f + 1.f;
    f[9] = log1pf(st.maxr[i]) * 0.25f;
    f[10] = vis * inv_views;
    const float dm = st.dsum[i] * iv;
    const float dv = fmaxf(st.dsq[i] * iv - dm * dm, 0.f);
    f[11] = dm * 0.2f;
    f[12] = sqrtf(dv) * 0.2f;
    for (int k = 0; k < 3; k++) f[13 + k] = f_dc[3 * (size_t)i + k];
    uint4 o[2];
    uint32_t* ow = reinterpret_cast<uint32_t*>(o);
    for (int k = 0; k < 8; k++) {
        const float a = ms_finite(f[2 * k]) ? f[2 * k] : 0.f, b = ms_finite(f[2 * k + 1]) ? f[2 * k + 1] : 0.f;
        ow[k] = (uint32_t)ms_bf16_bits(a) | ((uint32_t)ms_bf16_bits(b) << 16);
    }
    reinterpret_cast<uint4*>(X)[2 * (size_t)i] = o[0];
    reinterpret_cast<uint4*>(X)[2 * (size_t)i + 1] = o[1];
}

void relo_features(hipStream_t s, const Params& P, const ReloStats& st, int views, uint16_t* X) {
    if (P.N <= 0) return;
    hipLaunchKernelGGL(k_relo_features, dim3(cdiv_i(P.N, 256)), dim3(256), 0, s, P.N, P.scales, P.quats, P.opac,
                       P.f_dc, st, views > 0 ? 1.f / (float)views : 0.f, X);
}

__global__ void k_relo_record(int n, const uint32_t* __restrict__ pair_src, const uint32_t* __restrict__ pair_dst,
                              const uint16_t* __restrict__ X, const float* __restrict__ means,
                              const float* __restrict__ scales, const float* __restrict__ opac, ReloRecords R) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t src = pair_src[j], dst = pair_dst[j];
    const uint4* xs = reinterpret_cast<const uint4*>(X) + 2 * (size_t)src;
    uint4* xd = reinterpret_cast<uint4*>(R.X) + 2 * (size_t)j;
    xd[0] = xs[0];
    xd[1] = xs[1];
    R.dst[j] = dst;
    R.o[j] = opac[dst];
    for (int k = 0; k < 3; k++) R.mu[3 * (size_t)j + k] = means[3 * (size_t)src + k];
    const float* ls = scales + 3 * (size_t)dst;
    R.s[j] = ms_expf(fmaxf(fmaxf(ls[0], ls[1]), ls[2]));
}

void relo_record(hipStream_t s, int n, const uint32_t* pair_src, const uint32_t* pair_dst, const uint16_t* X,
                 const Params& P, ReloRecords R) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_relo_record, dim3(cdiv_i(n, 256)), dim3(256), 0, s, n, pair_src, pair_dst, X, P.means,
                       P.scales, P.opac, R);
}

__global__ void k_relo_label(int n, const ReloRecords R, const float* __restrict__ means,
                             const float* __restrict__ opac, float wl, float wo, uint16_t* __restrict__ X_out,
                             float* __restrict__ T_out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t d = R.dst[j];
    const uint4* xs = reinterpret_cast<const uint4*>(R.X) + 2 * (size_t)j;
    uint4* xd = reinterpret_cast<uint4*>(X_out) + 2 * (size_t)j;
    xd[0] = xs[0];
    xd[1] = xs[1];
    const float on = opac[d];
    float t[8];
    t[0] = (ms_finite(on) && on > R.o[j]) ? 1.f : 0.f;
    t[1] = wl;
    const float inv = R.s[j] > 0.f ? 1.f / R.s[j] : 0.f;
    bool ok = ms_finite(inv);
    for (int k = 0; k < 3; k++) {
        const float v = (means[3 * (size_t)d + k] - R.mu[3 * (size_t)j + k]) * inv;
        ok = ok && ms_finite(v);
        t[2 + k] = fminf(fmaxf(v, -3.f), 3.f);
    }
    if (!ok) t[2] = t[3] = t[4] = 0.f;
    t[5] = ok ? wo : 0.f;
    t[6] = t[7] = 0.f;
    float4* to = reinterpret_cast<float4*>(T_out) + 2 * (size_t)j;
    to[0] = make_float4(t[0], t[1], t[2], t[3]);
    to[1] = make_float4(t[4], t[5], t[6], t[7]);
}

void relo_label(hipStream_t s, int n, const ReloRecords R, const Params& P, float wl, float wo, uint16_t* X_out,
                float* T_out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_relo_label, dim3(cdiv_i(n, 256)), dim3(256), 0, s, n, R, P.means, P.opac, wl, wo, X_out,
                       T_out);
}

__global__ void k_relo_weights(int n, float* __restrict__ T, float wl, float wo) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    T[8 * (size_t)j + 1] = wl;
    if (T[8 * (size_t)j + 5] != 0.f) T[8 * (size_t)j + 5] = wo;
}

void relo_weights(hipStream_t s, int n, float* T, float wl, float wo) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_relo_weights, dim3(cdiv_i(n, 256)), dim3(256), 0, s, n, T, wl, wo);
}

// ------------------------------------------------------------------------------------
// packing and Adam

__global__ void k_mlp_pack(const float* __restrict__ th, MlpPacked* __restrict__ w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < MLP_HID * MLP_IN) w->W1[k] = ms_bf16_bits(th[MLP_OFF_W1 + k]);
    if (k < MLP_HID * MLP_HID) w->W2[k] = ms_bf16_bits(th[MLP_OFF_W2 + k]);
    if (k < MLP_OUT * MLP_HID) w->W3[k] = ms_bf16_bits(th[MLP_OFF_W3 + k]);
    if (k < MLP_HID) {
        w->b1[k] = th[MLP_OFF_B1 + k];
        w->b2[k] = th[MLP_OFF_B2 + k];
    }
    if (k < MLP_OUT) w->b3[k] = th[MLP_OFF_B3 + k];
}

void mlp_pack(hipStream_t s, const float* theta, MlpPacked* w) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(cdiv_i(MLP_HID * MLP_HID, 256)), dim3(256), 0, s, theta, w);
}

__global__ void k_mlp_adam(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float step, float b1, float b2, float eps, float bc2s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    const float gk = g[k];
    if (!ms_finite(gk)) return;
    const float mn = fmaf(b1, m[k], (1.f - b1) * gk);
    const float vn = fmaf(b2, v[k], (1.f - b2) * gk * gk);
    th[k] -= step * mn / (sqrtf(vn) / bc2s + eps);
    m[k] = mn;
    v[k] = vn;
}

void mlp_adam(hipStream_t s, float* theta, float* m, float* v, const float* grad, float lr, float beta1, float beta2,
              float eps, float bc1, float bc2_sqrt, MlpPacked* w) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv_i(MLP_NPARAM, 256)), dim3(256), 0, s, theta, m, v, grad, lr / bc1, beta1,
                       beta2, eps, bc2_sqrt);
    mlp_pack(s, theta, w);
}

// ------------------------------------------------------------------------------------
// forward / backward. Lane l: c = l % 16 (a column of the 16x16 tiles: a Gaussian, or a
// row of an A operand), kb = l / 16 (which 4 of the 16 reduction indices it holds).

// Register-resident operands: W1 and W3 fragments (W2 and the biases are read from the
// workgroup's LDS copy of the packed block per use).
struct Frags {
    bf16x4 a1[4];       // W1 rows 16m + c, inputs 4 kb + q
    bf16x4 a3[4];       // W3 row c (< 4), inputs 16r + 4 kb + q
};

// The packed block plus W2 transposed, staged in LDS by the whole workgroup.
struct alignas(16) MlpLds {
    MlpPacked w;
    uint16_t W2T[MLP_HID * MLP_HID];
};

__device__ __forceinline__ void stage_weights(const MlpPacked* __restrict__ w, MlpLds& L, bool transposed) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(w);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&L.w);
    for (int k = threadIdx.x; k < (int)(sizeof(MlpPacked) / 4); k += blockDim.x) dst[k] = src[k];
    if (transposed)
        for (int k = threadIdx.x; k < MLP_HID * MLP_HID; k += blockDim.x)
            L.W2T[(k % MLP_HID) * MLP_HID + k / MLP_HID] = w->W2[k];
    __syncthreads();
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {
    const uint2 raw = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(raw.x & 0xFFFFu);
    r[1] = (short)(raw.x >> 16);
    r[2] = (short)(raw.y & 0xFFFFu);
    r[3] = (short)(raw.y >> 16);
    return r;
}
__device__ __forceinline__ f32x4 ldf4(const float* p) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    f32x4 r = {v.x, v.y, v.z, v.w};
    return r;
}

__device__ __forceinline__ void load_frags(const MlpLds& L, int c, int kb, Frags& F) {
    for (int m = 0; m < 4; m++) F.a1[m] = ld4(&L.w.W1[(16 * m + c) * MLP_IN + 4 * kb]);
    const bf16x4 z = {0, 0, 0, 0};
    for (int r = 0; r < 4; r++) F.a3[r] = c < MLP_OUT ? ld4(&L.w.W3[c * MLP_HID + 16 * r + 4 * kb]) : z;
}

// Forward of the group whose column c is Gaussian g (zero input when g >= N).
struct Acts {
    f32x4 z1[4], z2[4], z3;
    bf16x4 h1[4], h2[4];
};

__device__ __forceinline__ void forward_group(const uint16_t* __restrict__ X, int N, int g, int c, int kb,
                                              const Frags& F, const MlpLds& L, Acts& A) {
    bf16x4 xb = {0, 0, 0, 0};
    if (g < N) xb = ld4(X + (size_t)g * MLP_IN + 4 * kb);
    for (int m = 0; m < 4; m++) {
        A.z1[m] = mfma16(F.a1[m], xb, ldf4(&L.w.b1[16 * m + 4 * kb]));
        for (int q = 0; q < 4; q++) A.h1[m][q] = bfs(gelu(A.z1[m][q]));
    }
    for (int m = 0; m < 4; m++) {
        f32x4 acc = ldf4(&L.w.b2[16 * m + 4 * kb]);
        for (int r = 0; r < 4; r++) acc = mfma16(ld4(&L.w.W2[(16 * m + c) * MLP_HID + 16 * r + 4 * kb]), A.h1[r], acc);
        A.z2[m] = acc;
        for (int q = 0; q < 4; q++) A.h2[m][q] = bfs(gelu(acc[q]));
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (kb == 0) acc = ldf4(L.w.b3);
    for (int r = 0; r < 4; r++) acc = mfma16(F.a3[r], A.h2[r], acc);
    A.z3 = acc;
}

constexpr int MLP_THREADS = 256;

__global__ __launch_bounds__(MLP_THREADS) void k_mlp_fwd(int N, const uint16_t* __restrict__ X,
                                                         const MlpPacked* __restrict__ w, float* __restrict__ out) {
    __shared__ MlpLds L;
    stage_weights(w, L, false);
    const int lane = threadIdx.x & 63, c = lane & 15, kb = lane >> 4;
    Frags F;
    load_frags(L, c, kb, F);
    const int wave = blockIdx.x * (MLP_THREADS / 64) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (MLP_THREADS / 64);
    const int groups = (N + 15) / 16;
    for (int gr = wave; gr < groups; gr += nwaves) {
        const int g = gr * 16 + c;
        Acts A;
        forward_group(X, N, g, c, kb, F, L, A);
        if (kb == 0 && g < N)
            *reinterpret_cast<float4*>(out + (size_t)g * MLP_OUT) = make_float4(A.z3[0], A.z3[1], A.z3[2], A.z3[3]);
    }
}

void mlp_forward(hipStream_t s, int N, const uint16_t* X, const MlpPacked* w, float* out) {
    if (N <= 0) return;
    const int groups = (N + 15) / 16;
    const int blocks = std::min(cdiv_i(groups, MLP_THREADS / 64), 2048);
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks), dim3(MLP_THREADS), 0, s, N, X, w, out);
}

// Wave-private transposition tiles (bf16): dz3 [16][16], then h2, dz2, h1, dz1 [64][16].
constexpr int TP_WAVE = 16 * 16 + 4 * 64 * 16;

// Writes a tile held in accumulator layout (rows 16 m + 4 kb + q, column c) into T [rows][16].
__device__ __forceinline__ void put_tile(uint16_t* T, int m, int c, int kb, bf16x4 v) {
    for (int q = 0; q < 4; q++) T[(16 * m + 4 * kb + q) * 16 + c] = (uint16_t)v[q];
}
// Reads row 16 m + c, columns 4 kb .. 4 kb + 3 of T (the A / B layout with the Gaussians
// as the reduction index).
__device__ __forceinline__ bf16x4 get_rowfrag(const uint16_t* T, int m, int c, int kb) {
    const uint2 raw = *reinterpret_cast<const uint2*>(T + (16 * m + c) * 16 + 4 * kb);
    bf16x4 r;
    r[0] = (short)(raw.x & 0xFFFFu);
    r[1] = (short)(raw.x >> 16);
    r[2] = (short)(raw.y & 0xFFFFu);
    r[3] = (short)(raw.y >> 16);
    return r;
}

// Sum over the 16 lanes of a row group (lanes sharing kb).
__device__ __forceinline__ float sum16(float x) {
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    return x;
}

__global__ __launch_bounds__(MLP_THREADS) void k_mlp_bwd(int N, const uint16_t* __restrict__ X,
                                                         const float* __restrict__ tgt,
                                                         const MlpPacked* __restrict__ w, float* __restrict__ grad,
                                                         double* __restrict__ loss2) {
    __shared__ MlpLds L;
    __shared__ alignas(16) uint16_t tp[MLP_THREADS / 64][TP_WAVE];
    stage_weights(w, L, true);
    const int lane = threadIdx.x & 63, c = lane & 15, kb = lane >> 4;
    uint16_t* T3 = tp[threadIdx.x >> 6];
    uint16_t* TH2 = T3 + 256;
    uint16_t* TZ2 = TH2 + 1024;
    uint16_t* TH1 = TZ2 + 1024;
    uint16_t* TZ1 = TH1 + 1024;
    Frags F;
    load_frags(L, c, kb, F);
    // W3 transposed for the input-gradient chain (W2 transposed is read from LDS)
    bf16x4 a3t[4];
    for (int m = 0; m < 4; m++)
        for (int q = 0; q < 4; q++) {
            const int o = 4 * kb + q;
            a3t[m][q] = o < MLP_OUT ? (short)L.w.W3[o * MLP_HID + 16 * m + c] : (short)0;
        }
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gw1[4], gw2[4][4], gw3[4];
    f32x4 gb1[4], gb2[4], gb3 = zero;
    for (int m = 0; m < 4; m++) {
        gw1[m] = zero; gw3[m] = zero; gb1[m] = zero; gb2[m] = zero;
        for (int n = 0; n < 4; n++) gw2[m][n] = zero;
    }
    double l_logit = 0.0, l_off = 0.0;
    const int wave = blockIdx.x * (MLP_THREADS / 64) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (MLP_THREADS / 64);
    const int groups = (N + 15) / 16;
    for (int gr = wave; gr < groups; gr += nwaves) {
        const int g = gr * 16 + c;
        Acts A;
        forward_group(X, N, g, c, kb, F, L, A);
        // loss gradient at the outputs (lanes kb == 0 hold Gaussian c's four outputs)
        f32x4 dz3 = zero;
        if (kb == 0 && g < N) {
            const float4 t0 = *reinterpret_cast<const float4*>(tgt + (size_t)g * 8);
            const float4 t1 = *reinterpret_cast<const float4*>(tgt + (size_t)g * 8 + 4);
            const float z = A.z3[0];
            const float sg = 1.f / (1.f + ms_expf(-z));
            dz3[0] = t0.y * (sg - t0.x);
            l_logit += (double)t0.y * (double)(fmaxf(z, 0.f) - t0.x * z + log1pf(ms_expf(-fabsf(z))));
            const float tt[3] = {t0.z, t0.w, t1.x};
            for (int k = 0; k < 3; k++) {
                const float d = A.z3[1 + k] - tt[k];
                dz3[1 + k] = 2.f * t1.y * d;
                l_off += (double)t1.y * (double)(d * d);
            }
        }
        bf16x4 dz3b;
        for (int q = 0; q < 4; q++) dz3b[q] = bfs(dz3[q]);
        // input-gradient chain
        bf16x4 dz2b[4];
        f32x4 dz2[4];
        for (int m = 0; m < 4; m++) {
            const f32x4 dh2 = mfma16(a3t[m], dz3b, zero);
            for (int q = 0; q < 4; q++) {
                dz2[m][q] = dh2[q] * gelu_d(A.z2[m][q]);
                dz2b[m][q] = bfs(dz2[m][q]);
            }
        }
        bf16x4 dz1b[4];
        f32x4 dz1[4];
        for (int m = 0; m < 4; m++) {
            f32x4 acc = zero;
            for (int r = 0; r < 4; r++) acc = mfma16(ld4(&L.W2T[(16 * m + c) * MLP_HID + 16 * r + 4 * kb]), dz2b[r], acc);
            for (int q = 0; q < 4; q++) {
                dz1[m][q] = acc[q] * gelu_d(A.z1[m][q]);
                dz1b[m][q] = bfs(dz1[m][q]);
            }
        }
        // bias gradients (fp32)
        for (int q = 0; q < 4; q++) gb3[q] += dz3[q];
        for (int m = 0; m < 4; m++) {
            gb1[m] += dz1[m];
            gb2[m] += dz2[m];
        }
        // weight gradients: operands transposed through LDS
        put_tile(T3, 0, c, kb, dz3b);
        for (int m = 0; m < 4; m++) {
            put_tile(TH2, m, c, kb, A.h2[m]);
            put_tile(TZ2, m, c, kb, dz2b[m]);
            put_tile(TH1, m, c, kb, A.h1[m]);
            put_tile(TZ1, m, c, kb, dz1b[m]);
        }
        wave_sync_lds();
        const bf16x4 a_dz3 = get_rowfrag(T3, 0, c, kb);
        for (int n = 0; n < 4; n++) gw3[n] = mfma16(a_dz3, get_rowfrag(TH2, n, c, kb), gw3[n]);
        for (int m = 0; m < 4; m++) {
            const bf16x4 a_dz2 = get_rowfrag(TZ2, m, c, kb);
            for (int n = 0; n < 4; n++) gw2[m][n] = mfma16(a_dz2, get_rowfrag(TH1, n, c, kb), gw2[m][n]);
        }
        bf16x4 bx = {0, 0, 0, 0};
        for (int q = 0; q < 4; q++) {
            const int gg = gr * 16 + 4 * kb + q;
            bx[q] = gg < N ? (short)X[(size_t)gg * MLP_IN + c] : (short)0;
        }
        for (int m = 0; m < 4; m++) gw1[m] = mfma16(get_rowfrag(TZ1, m, c, kb), bx, gw1[m]);
        wave_sync_lds();
    }
    // one global add per weight per wave
    for (int q = 0; q < 4; q++) {
        const int row = 4 * kb + q;  // accumulator row
        for (int m = 0; m < 4; m++) {
            atomicAdd(&grad[MLP_OFF_W1 + (16 * m + row) * MLP_IN + c], gw1[m][q]);
            for (int n = 0; n < 4; n++) atomicAdd(&grad[MLP_OFF_W2 + (16 * m + row) * MLP_HID + 16 * n + c], gw2[m][n][q]);
        }
        if (row < MLP_OUT)
            for (int n = 0; n < 4; n++) atomicAdd(&grad[MLP_OFF_W3 + row * MLP_HID + 16 * n + c], gw3[n][q]);
    }
    for (int q = 0; q < 4; q++) {
        const int row = 4 * kb + q;
        for (int m = 0; m < 4; m++) {
            const float s1 = sum16(gb1[m][q]), s2 = sum16(gb2[m][q]);
            if (c == 0) {
                atomicAdd(&grad[MLP_OFF_B1 + 16 * m + row], s1);
                atomicAdd(&grad[MLP_OFF_B2 + 16 * m + row], s2);
            }
        }
        const float s3 = sum16(gb3[q]);
        if (c == 0 && kb == 0) atomicAdd(&grad[MLP_OFF_B3 + q], s3);
    }
    // losses: lanes 0..15 hold them
    float ll = (float)l_logit, lo = (float)l_off;
    ll = sum16(ll);
    lo = sum16(lo);
    if (lane == 0) {
        atomicAdd(&loss2[0], (double)ll);
        atomicAdd(&loss2[1], (double)lo);
    }
}

void mlp_backward(hipStream_t s, int N, const uint16_t* X, const float* targets, const MlpPacked* w, float* grad,
                  double* loss2) {
    if (N <= 0) return;
    const int groups = (N + 15) / 16;
    const int blocks = std::min(cdiv_i(groups, 4 * (MLP_THREADS / 64)), 256);
    hipLaunchKernelGGL(k_mlp_bwd, dim3(blocks), dim3(MLP_THREADS), 0, s, N, X, targets, w, grad, loss2);
}

}  // namespace msk

// oracle_capi.cpp — extern "C" entry points of the CPU restatement (TEST INFRASTRUCTURE).
// Loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// only; the product never links it.
#include <cstdio>
#include "oracle.hpp"
#include "oracle_mcmc.hpp"
#include "oracle_mlp.hpp"

using namespace oracle;

#define ORACLE_STRUCTS(SUF, FT)                                                                  \
    typedef struct {                                                                             \
        int32_t N, degree;                                                                       \
        FT *means, *scales, *quats, *f_dc, *f_rest, *opac;                                       \
        FT* m[6];                                                                                \
        FT* v[6];                                                                                \
        int32_t sh_bf16;                                                                         \
    } OracleState##SUF;                                                                          \
    typedef struct {                                                                             \
        FT *xys, *depths, *conics, *colors, *aabb;                                               \
        int32_t *radii, *tiles_hit;                                                              \
        int32_t *tile_counts, *tile_bins, *ids;                                                  \
        int64_t ids_cap, num_isect, num_padded;                                                  \
        int32_t overflow;                                                                        \
        uint32_t k_max;                                                                          \
        FT *img, *final_T;                                                                       \
        int32_t* final_idx;                                                                      \
        FT* v_rendered;                                                                          \
        double loss_sum;                                                                         \
        FT *v_xy, *v_conic, *v_rgb, *v_opac, *v_mean, *v_scale, *v_quat;                         \
        uint32_t nan_adam[6];                                                                    \
    } OracleOut##SUF;

extern "C" {
ORACLE_STRUCTS(F32, float)
ORACLE_STRUCTS(F64, double)
}

template <class T, class SS> static State<T> make_state(SS* s) {
    State<T> S;
    S.N = s->N;
    S.degree = s->degree;
    S.means = s->means; S.scales = s->scales; S.quats = s->quats;
    S.f_dc = s->f_dc; S.f_rest = s->f_rest; S.opac = s->opac;
    for (int g = 0; g < 6; g++) { S.m[g] = s->m[g]; S.v[g] = s->v[g]; }
    S.sh_bf16 = s->sh_bf16;
    return S;
}

template <class T> static void cp(T* dst, const std::vector<T>& src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(T));
}

template <class T, class OO> static void export_out(const StepOut<T>& O, OO* o, bool train) {
    cp(o->xys, O.proj.xys); cp(o->depths, O.proj.depths); cp(o->conics, O.proj.conics);
    cp(o->colors, O.proj.colors); cp(o->aabb, O.proj.aabb);
    cp(o->radii, O.proj.radii); cp(o->tiles_hit, O.proj.tiles_hit);
    if (o->tile_counts) for (size_t t = 0; t < O.lists.counts.size(); t++) o->tile_counts[t] = (int32_t)O.lists.counts[t];
    cp(o->tile_bins, O.lists.bins);
    o->num_isect = O.lists.num_isect;
    o->num_padded = O.lists.num_padded;
    o->overflow = O.lists.overflow ? 1 : 0;
    if (o->ids && (int64_t)O.lists.ids.size() <= o->ids_cap) cp(o->ids, O.lists.ids);
    o->k_max = O.raster.K;
    cp(o->img, O.raster.img); cp(o->final_T, O.raster.final_T); cp(o->final_idx, O.raster.final_idx);
    if (!train) return;
    cp(o->v_rendered, O.v_rend);
    o->loss_sum = O.loss_sum;
    cp(o->v_xy, O.rg.v_xy); cp(o->v_conic, O.rg.v_conic); cp(o->v_rgb, O.rg.v_rgb);
    cp(o->v_opac, O.v_opac_total); cp(o->v_mean, O.v_mean); cp(o->v_scale, O.v_scale); cp(o->v_quat, O.v_quat);
    for (int k = 0; k < 6; k++) o->nan_adam[k] = O.nan_adam[k];
}

extern "C" {

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// One training step; mask (uint8 [H W], pixels below 128 own no loss) may be null.
int oracle_train_step_f32(OracleStateF32* s, const MsplatCamera* cam, const uint8_t* gt, const MsplatStepParams* hp,
                          OracleOutF32* o, const uint8_t* mask) {
    State<float> S = make_state<float>(s);
    StepOut<float> O;
    train_step<float>(S, *cam, gt, *hp, O, mask);
    export_out(O, o, true);
    return O.lists.overflow ? 1 : 0;
}

int oracle_train_step_f64(OracleStateF64* s, const MsplatCamera* cam, const uint8_t* gt, const MsplatStepParams* hp,
                          OracleOutF64* o, const uint8_t* mask) {
    State<double> S = make_state<double>(s);
    StepOut<double> O;
    train_step<double>(S, *cam, gt, *hp, O, mask);
    export_out(O, o, true);
    return O.lists.overflow ? 1 : 0;
}

int oracle_render_f32(OracleStateF32* s, const MsplatCamera* cam, int degrees_to_use, const float* bg, OracleOutF32* o) {
    State<float> S = make_state<float>(s);
    StepOut<float> O;
    render<float>(S, *cam, degrees_to_use, bg, O);
    export_out(O, o, false);
    return O.lists.overflow ? 1 : 0;
}

int oracle_render_f64(OracleStateF64* s, const MsplatCamera* cam, int degrees_to_use, const float* bg, OracleOutF64* o) {
    State<double> S = make_state<double>(s);
    StepOut<double> O;
    render<double>(S, *cam, degrees_to_use, bg, O);
    export_out(O, o, false);
    return O.lists.overflow ? 1 : 0;
}

// SSIM + L1 loss on a given image (used to pin the loss against an independent SSIM); the
// masked variants take a uint8 [H W] training mask.
double oracle_ssim_loss_f64(int W, int H, const double* rend, const uint8_t* gt, double w, double inv_n, double* v_rend) {
    std::vector<double> v;
    double l = ssim_loss<double>(W, H, rend, gt, w, inv_n, v);
    if (v_rend) std::memcpy(v_rend, v.data(), v.size() * sizeof(double));
    return l;
}
double oracle_ssim_loss_f32(int W, int H, const float* rend, const uint8_t* gt, float w, float inv_n, float* v_rend) {
    std::vector<float> v;
    double l = ssim_loss<float>(W, H, rend, gt, w, inv_n, v);
    if (v_rend) std::memcpy(v_rend, v.data(), v.size() * sizeof(float));
    return l;
}
double oracle_ssim_loss_masked_f64(int W, int H, const double* rend, const uint8_t* gt, const uint8_t* mask, double w,
                                   double inv_n, double* v_rend) {
    std::vector<double> v;
    double l = ssim_loss<double>(W, H, rend, gt, w, inv_n, v, mask);
    if (v_rend) std::memcpy(v_rend, v.data(), v.size() * sizeof(double));
    return l;
}
double oracle_ssim_loss_masked_f32(int W, int H, const float* rend, const uint8_t* gt, const uint8_t* mask, float w,
                                   float inv_n, float* v_rend) {
    std::vector<float> v;
    double l = ssim_loss<float>(W, H, rend, gt, w, inv_n, v, mask);
    if (v_rend) std::memcpy(v_rend, v.data(), v.size() * sizeof(float));
    return l;
}

void oracle_sgld_noise_f32(int N, uint32_t seed, float* out) {
    for (int i = 0; i < N; i++) sgld_noise3<float>(seed, (uint32_t)i, out + 3 * (size_t)i);
}
uint32_t oracle_pcg_hash32(uint32_t seed, uint32_t idx) { return pcg_hash32(seed, idx); }

// One Gaussian's SGLD move (sgld_apply): mean[3] updated in place.
void oracle_sgld_apply_f64(double* mean, const double* ls, const double* q, double opac_logit, const double* n,
                           double noise_lr, double xyz_lr, double max_disp) {
    sgld_apply<double>(mean, ls, q, opac_logit, n, noise_lr, xyz_lr, max_disp);
}
void oracle_sgld_apply_f32(float* mean, const float* ls, const float* q, float opac_logit, const float* n,
                           float noise_lr, float xyz_lr, float max_disp) {
    sgld_apply<float>(mean, ls, q, opac_logit, n, noise_lr, xyz_lr, max_disp);
}

// Regulariser gradients of N rows (reg_grad): v_opac[N] and v_scale[N 3] are added to.
void oracle_reg_grad_f64(int N, const double* opac, const double* ls, const double* dc_sq, double op_reg,
                         double sc_reg, double reg_invis_scale, double* v_opac, double* v_scale) {
    for (int i = 0; i < N; i++)
        reg_grad<double>(N, opac[i], ls + 3 * (size_t)i, dc_sq + 3 * (size_t)i, op_reg, sc_reg, reg_invis_scale,
                         v_opac[i], v_scale + 3 * (size_t)i);
}

float oracle_ms_expf(float x) { return ms_expf(x); }
float oracle_ms_logf(float x) { return ms_logf(x); }
void oracle_ms_sincosf(float a, float* s, float* c) { ms_sincosf(a, s, c); }

uint32_t oracle_k_max(int64_t capacity, int num_tiles) { return k_max_rule(capacity, num_tiles); }

// Eq. 9 (model:311-332) in the relocation's arithmetic (see oracle_mcmc.hpp).
void oracle_eq9(float o_old, const float* s_old, int n, float* o_new, float* s_new) {
    relocation_eq9(o_old, s_old, n, o_new, s_new);
}

// Batched step (FP32 or FP64 state); outputs per view: image and raster gradients when the
// pointers are set; masks (per view uint8 [H W] or null; the array itself may be null);
// returns the mean of the per-view losses.
#define ORACLE_BATCH(SUF, NAME, FT)                                                                         \
    double oracle_train_step_batch_##NAME(OracleState##SUF* s, const MsplatCamera* cams,                    \
                                         const uint8_t* const* gts, const MsplatStepParams* hps, int B,     \
                                         FT* const* imgs, FT* const* v_xy, const uint8_t* const* masks) {   \
        State<FT> S = make_state<FT>(s);                                                                    \
        BatchOut<FT> O;                                                                                     \
        train_step_batch<FT>(S, cams, gts, hps, B, O, masks);                                               \
        for (int v = 0; v < B; v++) {                                                                       \
            if (imgs && imgs[v]) cp(imgs[v], O.views[v].raster.img);                                        \
            if (v_xy && v_xy[v]) cp(v_xy[v], O.views[v].rg.v_xy);                                          \
        }                                                                                                   \
        return O.loss_mean;                                                                                 \
    }
ORACLE_BATCH(F32, f32, float)
ORACLE_BATCH(F64, f64, double)

int oracle_mcmc_step_learned(OracleStateF32* s, int32_t cap_max, float cull_radius, uint32_t reloc_count_in,
                             const float* mlp_out, int32_t* out_info) {
    State<float> S = make_state<float>(s);
    McmcInfo info;
    mcmc_relocate_and_grow(S, cap_max, cull_radius, reloc_count_in, info, mlp_out);
    out_info[0] = info.num_active;
    out_info[1] = info.dead;
    out_info[2] = info.grown;
    out_info[3] = info.nonfinite;
    out_info[4] = info.culled;
    out_info[5] = (int32_t)info.reloc_count_out;
    return 0;
}

int oracle_mcmc_step(OracleStateF32* s, int32_t cap_max, float cull_radius, uint32_t reloc_count_in, int32_t* out_info) {
    State<float> S = make_state<float>(s);
    McmcInfo info;
    mcmc_relocate_and_grow(S, cap_max, cull_radius, reloc_count_in, info);
    out_info[0] = info.num_active;
    out_info[1] = info.dead;
    out_info[2] = info.grown;
    out_info[3] = info.nonfinite;
    out_info[4] = info.culled;
    out_info[5] = (int32_t)info.reloc_count_out;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Entry points for the pins in tests/test_oracle.py (FP64 unless noted).
extern "C" {

// Rasterise n 2-D Gaussians given directly in screen space (xy, conic, raw rgb, opacity
// logit, depth) through the same list build, compositing and (optionally) backward as a
// training step. K = 0 selects the reference's K rule.
int oracle_raster2d_f64(int n, const double* xy, const double* conic, const double* rgb, const double* opac_logit,
                        const double* depth, int W, int H, const float* bg, int K, double* img, double* final_T,
                        int32_t* final_idx, const double* v_out, double* v_xy, double* v_conic, double* v_rgb,
                        double* v_opac) {
    using T = double;
    Proj<T> pj;
    pj.resize(n);
    const int tbx = (W + 15) / 16, tby = (H + 15) / 16;
    for (int i = 0; i < n; i++) {
        pj.xys[2 * i] = xy[2 * i];
        pj.xys[2 * i + 1] = xy[2 * i + 1];
        for (int c = 0; c < 3; c++) { pj.conics[3 * i + c] = conic[3 * i + c]; pj.colors[3 * i + c] = rgb[3 * i + c]; }
        pj.depths[i] = depth[i];
        pj.opac_sig[i] = sigmoid<T>(opac_logit[i]);
        // covariance extents from the conic's inverse (det of the conic = 1/det cov)
        T det = conic[3 * i] * conic[3 * i + 2] - conic[3 * i + 1] * conic[3 * i + 1];
        T sxx = conic[3 * i + 2] / det, syy = conic[3 * i] / det;
        pj.aabb[2 * i] = std::ceil(3.0 * std::sqrt(sxx));
        pj.aabb[2 * i + 1] = std::ceil(3.0 * std::sqrt(syy));
        int x0, y0, x1, y1;
        tile_bbox<T>(pj.xys[2 * i], pj.xys[2 * i + 1], pj.aabb[2 * i], pj.aabb[2 * i + 1], tbx, tby, x0, y0, x1, y1);
        int area = (x1 - x0) * (y1 - y0);
        pj.radii[i] = area > 0 ? 1 : 0;
        pj.tiles_hit[i] = area > 0 ? area : 0;
    }
    Lists<T> L;
    build_lists<T>(n, pj, W, H, L);
    uint32_t KK = K > 0 ? (uint32_t)K : k_max_rule(L.capacity, L.num_tiles);
    Raster<T> R;
    raster_forward<T>(L, W, H, bg, KK, R);
    std::memcpy(img, R.img.data(), R.img.size() * sizeof(T));
    std::memcpy(final_T, R.final_T.data(), R.final_T.size() * sizeof(T));
    std::memcpy(final_idx, R.final_idx.data(), R.final_idx.size() * sizeof(int32_t));
    if (v_out) {
        RasterGrads<T> G;
        raster_backward<T>(L, R, bg, v_out, n, G);
        std::memcpy(v_xy, G.v_xy.data(), G.v_xy.size() * sizeof(T));
        std::memcpy(v_conic, G.v_conic.data(), G.v_conic.size() * sizeof(T));
        std::memcpy(v_rgb, G.v_rgb.data(), G.v_rgb.size() * sizeof(T));
        std::memcpy(v_opac, G.v_opac.data(), G.v_opac.size() * sizeof(T));
    }
    return (int)KK;
}

// Projection of n Gaussians (FP64): pixel centres, conics, depths; and the VJP of the
// projection for given v_xy / v_conic.
void oracle_project_f64(int n, const double* means, const double* scales, const double* quats, const MsplatCamera* cam,
                        double* xys, double* conics, double* depths, int32_t* radii) {
    using T = double;
    std::vector<T> dc(3 * (size_t)n, 0.0), rest(1, 0.0), op(n, 0.0);
    Proj<T> pj;
    project_forward<T>(n, means, scales, quats, dc.data(), rest.data(), op.data(), 0, 0, *cam, pj);
    std::memcpy(xys, pj.xys.data(), pj.xys.size() * sizeof(T));
    std::memcpy(conics, pj.conics.data(), pj.conics.size() * sizeof(T));
    std::memcpy(depths, pj.depths.data(), pj.depths.size() * sizeof(T));
    std::memcpy(radii, pj.radii.data(), pj.radii.size() * sizeof(int32_t));
}

void oracle_project_vjp_f64(int n, const double* means, const double* scales, const double* quats,
                            const MsplatCamera* cam, const double* conics, const double* v_xy, const double* v_conic,
                            double* v_mean, double* v_scale, double* v_quat) {
    for (int i = 0; i < n; i++)
        proj_backward_one<double>(means + 3 * i, scales + 3 * i, quats + 4 * i, *cam, conics + 3 * i, v_xy + 2 * i,
                                  v_conic + 3 * i, v_mean + 3 * i, v_scale + 3 * i, v_quat + 4 * i);
}

// SH colour (raw, before +0.5) for a unit view direction.
void oracle_sh_color_f64(int degree, const double* dir, const double* dc, const double* rest, double* out) {
    sh_to_color<double>(degree, dir, dc, rest, out);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Evaluation metrics: psnr / l1_loss (model:34-49) and ssim_eval (ssim:61-123), the 1-D
// window as createSSIMWindow computes it (ssim:7-24, host expf), replicate padding, the
// blur accumulated with fused multiply-add (clang's default contraction of
// `sum += in * k` on the reference's arm64 targets). FP64 sums in index order.
// Pins: SSIM(x, x) = 1 exactly; PSNR of a constant offset in closed form; scipy's
// gaussian_filter(mode="nearest") SSIM in FP64 (tests/test_oracle.py).
extern "C" void oracle_ssim_window(float out[11]) {
    float sum = 0.f;
    for (int i = 0; i < 11; i++) {
        const float x = (float)(i - 5);
        out[i] = std::exp(-(x * x) / (2.0f * 1.5f * 1.5f));
        sum += out[i];
    }
    for (int i = 0; i < 11; i++) out[i] /= sum;
}

static void blur_replicate(const float* in, float* out, float* tmp, int H, int W, const float* k) {
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float s = 0.f;
            for (int t = 0; t < 11; t++) {
                const int sx = std::min(std::max(x + t - 5, 0), W - 1);
                s = std::fma(in[(size_t)y * W + sx], k[t], s);
            }
            tmp[(size_t)y * W + x] = s;
        }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float s = 0.f;
            for (int t = 0; t < 11; t++) {
                const int sy = std::min(std::max(y + t - 5, 0), H - 1);
                s = std::fma(tmp[(size_t)sy * W + x], k[t], s);
            }
            out[(size_t)y * W + x] = s;
        }
}

// out3 = {sum (r-g)^2, sum |r-g|, sum SSIM} over H*W*3 values; g = gt_u8 / 255.
extern "C" void oracle_eval_sums(int H, int W, const float* img, const uint8_t* gt, double* out3) {
    const size_t P = (size_t)H * W;
    float k[11];
    oracle_ssim_window(k);
    std::vector<float> r(P), g(P), rr(P), gg(P), rg(P), m1(P), m2(P), s11(P), s22(P), s12(P), tmp(P);
    double sq = 0, ab = 0, ss = 0;
    for (size_t i = 0; i < P * 3; i++) {
        const float a = img[i], b = (float)gt[i] * (1.0f / 255.0f);
        const double d = (double)(a - b);
        sq += d * d;
        ab += (double)std::fabs(a - b);
    }
    const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
    for (int c = 0; c < 3; c++) {
        for (size_t i = 0; i < P; i++) {
            r[i] = img[i * 3 + c];
            g[i] = (float)gt[i * 3 + c] * (1.0f / 255.0f);
            rr[i] = r[i] * r[i];
            gg[i] = g[i] * g[i];
            rg[i] = r[i] * g[i];
        }
        blur_replicate(r.data(), m1.data(), tmp.data(), H, W, k);
        blur_replicate(g.data(), m2.data(), tmp.data(), H, W, k);
        blur_replicate(rr.data(), s11.data(), tmp.data(), H, W, k);
        blur_replicate(gg.data(), s22.data(), tmp.data(), H, W, k);
        blur_replicate(rg.data(), s12.data(), tmp.data(), H, W, k);
        for (size_t i = 0; i < P; i++) {
            const float m1sq = m1[i] * m1[i], m2sq = m2[i] * m2[i], m12 = m1[i] * m2[i];
            const float sig1 = s11[i] - m1sq, sig2 = s22[i] - m2sq, sig12 = s12[i] - m12;
            const float num = (2.0f * m12 + C1) * (2.0f * sig12 + C2);
            const float den = (m1sq + m2sq + C1) * (sig1 + sig2 + C2);
            ss += (double)(num / den);
        }
    }
    out3[0] = sq;
    out3[1] = ab;
    out3[2] = ss;
}

// ---- learned relocation proposal -------------------------------------------------------
extern "C" {
// out [N 4]; exact != 0: double arithmetic without bf16 rounding.
void oracle_mlp_forward(const float* theta, const uint16_t* X, int N, int exact, double* out) {
    for (int g = 0; g < N; g++) {
        float x[oracle::MI];
        for (int f = 0; f < oracle::MI; f++) x[f] = oracle::bf(X[(size_t)g * oracle::MI + f]);
        if (exact) {
            oracle::MlpActs<double> A;
            oracle::mlp_forward_one<double>(theta, x, true, A);
            for (int o = 0; o < oracle::MO; o++) out[(size_t)g * 4 + o] = A.z3[o];
        } else {
            oracle::MlpActs<float> A;
            oracle::mlp_forward_one<float>(theta, x, false, A);
            for (int o = 0; o < oracle::MO; o++) out[(size_t)g * 4 + o] = A.z3[o];
        }
    }
}
void oracle_mlp_backward(const float* theta, const uint16_t* X, const float* tgt, int N, int exact, double* grad,
                         double* loss2) {
    if (exact) {
        oracle::mlp_backward<double>(theta, X, tgt, N, true, grad, loss2);
    } else {
        std::vector<float> g(oracle::NPAR);
        oracle::mlp_backward<float>(theta, X, tgt, N, false, g.data(), loss2);
        for (int k = 0; k < oracle::NPAR; k++) grad[k] = g[k];
    }
}
double oracle_mlp_loss_exact(const double* theta, const uint16_t* X, const float* tgt, int N) {
    return oracle::mlp_loss_exact(theta, X, tgt, N);
}
}

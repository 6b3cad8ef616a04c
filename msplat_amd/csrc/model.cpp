// model.cpp — device-resident trainable state and the host side of one training step.
// Reference: core/src/model.cpp (setupOptimizers 168-244, schedulersStep 256-261,
// mcmcAfterTrain 497-565, checkpoint 698-843, prepareCam 845-889, fullIteration 901-980).
//
// All parameters and Adam moments live in cap_max-row device buffers allocated once; the
// active prefix [0, num_active) is what the kernels see. Relocation and growth run on the
// device (mcmc.hip), so no step touches the parameters from the host.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "../../include/msplat_api_ext.h"
#include "ms_math.h"

namespace msplat_hip {

void HostGaussians::resize(int rows, int deg) {
    n = rows;
    degree = deg;
    means.assign((size_t)rows * 3, 0.f);
    scales.assign((size_t)rows * 3, 0.f);
    quats.assign((size_t)rows * 4, 0.f);
    f_dc.assign((size_t)rows * 3, 0.f);
    f_rest.assign((size_t)rows * rest_len(), 0.f);
    opac.assign((size_t)rows, 0.f);
}

int row_len(int group, int degree) {
    static const int base[6] = {3, 3, 4, 3, 0, 1};
    return group == 4 ? ((degree + 1) * (degree + 1) - 1) * 3 : base[group];
}

static std::vector<float>* group_vec(HostGaussians& g, int k) {
    std::vector<float>* v[6] = {&g.means, &g.scales, &g.quats, &g.f_dc, &g.f_rest, &g.opac};
    return v[k];
}
static const std::vector<float>* group_vec(const HostGaussians& g, int k) {
    const std::vector<float>* v[6] = {&g.means, &g.scales, &g.quats, &g.f_dc, &g.f_rest, &g.opac};
    return v[k];
}

static uint64_t mix64(uint64_t& s) {
    uint64_t x = (s += 0x9E3779B97F4A7C15ull);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

Model::Model(const ModelOptions& opt, const HostGaussians& init, float scale, const float translation[3])
    : opt_(opt), scale_(scale) {
    if (opt.cap_max <= 0) throw std::invalid_argument("cap_max must be positive");
    if (opt.sh_degree < 0 || opt.sh_degree > 3) throw std::invalid_argument("sh_degree must be in [0, 3]");
    if (opt.sh_degree_interval <= 0) throw std::invalid_argument("sh_degree_interval must be positive");
    if (init.degree != opt.sh_degree) throw std::invalid_argument("initial Gaussians have the wrong SH degree");
    for (int k = 0; k < 3; k++) translation_[k] = translation ? translation[k] : 0.f;
    cap_ = opt.cap_max;
    if (const char* e = std::getenv("MSPLAT_NAN_INJECT")) nan_inject_step_ = std::atoi(e);
    allocate(opt.sh_degree);

    // Seed selection (setupOptimizers, model.cpp:176-204): fresh seeds above 75% of the cap
    // are decimated to 50% of it (uniform sample without replacement, index order kept);
    // otherwise rows beyond the cap are dropped. The sample is this build's own
    // (splitmix64 Fisher-Yates, seed 42): libc++'s std::shuffle stream is not portable.
    HostGaussians g = init;
    int64_t rows = std::min<int64_t>(init.n, cap_);
    if (opt.decimate_seeds && init.n > cap_ * 0.75) {
        rows = (int64_t)(cap_ * 0.5);
        std::vector<int> idx(init.n);
        for (int i = 0; i < init.n; i++) idx[i] = i;
        uint64_t s = 42;
        for (int i = init.n - 1; i > 0; i--) {
            const int j = (int)(mix64(s) % (uint64_t)(i + 1));
            std::swap(idx[i], idx[j]);
        }
        idx.resize(rows);
        std::sort(idx.begin(), idx.end());
        HostGaussians d;
        d.resize((int)rows, init.degree);
        for (int k = 0; k < 6; k++) {
            const int L = row_len(k, init.degree);
            const std::vector<float>& src = *group_vec(init, k);
            std::vector<float>& dst = *group_vec(d, k);
            for (int64_t r = 0; r < rows; r++)
                std::memcpy(dst.data() + r * L, src.data() + (size_t)idx[r] * L, L * sizeof(float));
        }
        g = std::move(d);
    } else if (init.n > cap_) {
        HostGaussians d;
        d.resize((int)rows, init.degree);
        for (int k = 0; k < 6; k++) {
            const int L = row_len(k, init.degree);
            std::memcpy(group_vec(d, k)->data(), group_vec(init, k)->data(), (size_t)rows * L * sizeof(float));
        }
        g = std::move(d);
    }
    upload(g, nullptr, nullptr);
}

Model::~Model() {
    (void)hipStreamSynchronize(eng_.stream());
    for (int k = 0; k < 6; k++) {
        p_[k].release();
        m_[k].release();
        v_[k].release();
    }
    dc_h_.release();
    rest_h_.release();
    for (auto& b : acc_) b.release();
    mcmc_.release();
    prop_.release();
}

void Model::allocate(int degree) {
    degree_ = degree;
    hipStream_t s = eng_.stream();
    for (int k = 0; k < 6; k++) {
        const size_t rows = (size_t)cap_ * std::max(row_len(k, degree), 1);
        p_[k].release();
        m_[k].release();
        v_[k].release();
        p_[k].ensure(rows * 4);
        m_[k].ensure(rows * moment_elem_bytes(k));
        v_[k].ensure(rows * moment_elem_bytes(k));
        msk::fill_zero(s, p_[k].p, rows * 4);
        msk::fill_zero(s, m_[k].p, rows * moment_elem_bytes(k));
        msk::fill_zero(s, v_[k].p, rows * moment_elem_bytes(k));
    }
    dc_h_.release();
    rest_h_.release();
    if (sh_bf16_) {
        dc_h_.ensure((size_t)cap_ * 3 * 2);
        rest_h_.ensure((size_t)cap_ * std::max(row_len(4, degree), 1) * 2);
        msk::fill_zero(s, dc_h_.p, (size_t)cap_ * 3 * 2);
        msk::fill_zero(s, rest_h_.p, (size_t)cap_ * std::max(row_len(4, degree), 1) * 2);
    }
    hip_check(hipStreamSynchronize(s), "allocate");
}

void Model::refresh_shadows() {
    if (!sh_bf16_) return;
    msk::Params P;
    P.f_dc = p_[3].as<float>();
    P.f_rest = p_[4].as<float>();
    P.f_dc_h = dc_h_.as<uint16_t>();
    P.f_rest_h = rest_h_.as<uint16_t>();
    P.N = num_active_;
    P.degree = degree_;
    P.sh_bf16 = 1;
    msk::sh_refresh_shadows(eng_.stream(), P);
}

void Model::set_sh_bf16(bool on) {
    if (on == sh_bf16_) return;
    hipStream_t s = eng_.stream();
    for (int g = 3; g <= 4; g++) {
        const size_t n = (size_t)cap_ * std::max(row_len(g, degree_), 1);
        DBuf* bufs[2] = {&m_[g], &v_[g]};
        for (DBuf* b : bufs) {
            DBuf nb;
            nb.ensure(n * (on ? 2 : 4));
            if (on) msk::f32_to_bf16(s, b->as<float>(), nb.as<uint16_t>(), n);
            else msk::bf16_to_f32(s, b->as<uint16_t>(), nb.as<float>(), n);
            hip_check(hipStreamSynchronize(s), "moment conversion");
            b->release();
            *b = nb;
        }
    }
    sh_bf16_ = on;
    dc_h_.release();
    rest_h_.release();
    if (on) {
        dc_h_.ensure((size_t)cap_ * 3 * 2);
        rest_h_.ensure((size_t)cap_ * std::max(row_len(4, degree_), 1) * 2);
        msk::fill_zero(s, dc_h_.p, (size_t)cap_ * 3 * 2);
        msk::fill_zero(s, rest_h_.p, (size_t)cap_ * std::max(row_len(4, degree_), 1) * 2);
        refresh_shadows();
    }
    hip_check(hipStreamSynchronize(s), "sh storage switch");
}

void Model::upload(const HostGaussians& g, const std::vector<float>* m, const std::vector<float>* v) {
    if (g.n > cap_) throw std::runtime_error("more Gaussians than cap_max");
    if (g.degree != degree_) allocate(g.degree);
    hipStream_t s = eng_.stream();
    std::vector<uint16_t> half;
    for (int k = 0; k < 6; k++) {
        const size_t count = (size_t)g.n * row_len(k, degree_);
        if (count == 0) continue;
        hip_check(hipMemcpyAsync(p_[k].p, group_vec(g, k)->data(), count * 4, hipMemcpyHostToDevice, s), "upload");
        const std::vector<float>* src[2] = {m ? &m[k] : nullptr, v ? &v[k] : nullptr};
        DBuf* dst[2] = {&m_[k], &v_[k]};
        for (int w = 0; w < 2; w++) {
            if (!src[w]) continue;
            if (moment_elem_bytes(k) == 2) {
                half.resize(count);
                for (size_t e = 0; e < count; e++) half[e] = ms_bf16_bits((*src[w])[e]);
                hip_check(hipMemcpyAsync(dst[w]->p, half.data(), count * 2, hipMemcpyHostToDevice, s), "upload");
                hip_check(hipStreamSynchronize(s), "upload");
            } else {
                hip_check(hipMemcpyAsync(dst[w]->p, src[w]->data(), count * 4, hipMemcpyHostToDevice, s), "upload");
            }
        }
    }
    hip_check(hipStreamSynchronize(s), "upload");
    num_active_ = g.n;
    refresh_shadows();
    hip_check(hipStreamSynchronize(s), "upload");
}

MsplatDeviceState Model::state() const {
    MsplatDeviceState st;
    st.N = num_active_;
    st.degree = degree_;
    st.means = p_[0].as<float>();
    st.scales = p_[1].as<float>();
    st.quats = p_[2].as<float>();
    st.f_dc = p_[3].as<float>();
    st.f_rest = p_[4].as<float>();
    st.opac = p_[5].as<float>();
    for (int k = 0; k < 6; k++) {
        st.m[k] = m_[k].as<float>();
        st.v[k] = v_[k].as<float>();
    }
    st.sh_bf16 = sh_bf16_ ? 1 : 0;
    st.f_dc_bf16 = sh_bf16_ ? dc_h_.as<uint16_t>() : nullptr;
    st.f_rest_bf16 = sh_bf16_ ? rest_h_.as<uint16_t>() : nullptr;
    return st;
}

int Model::degrees_to_use(int step) const {
    return std::min(step / opt_.sh_degree_interval, opt_.sh_degree);
}

static MsplatCamera prepare(const TrainCamera& cam) {
    MsplatCamera c;
    msplat_camera_prepare(cam.c2w, cam.fx, cam.fy, cam.cx, cam.cy, cam.width, cam.height, &c);
    return c;
}

MsplatStepParams Model::step_params(const TrainCamera& cam, int step, float ssim_weight, int adam_step) const {
    MsplatStepParams hp;
    std::memset(&hp, 0, sizeof(hp));
    // bias corrections as the reference computes them: std::pow(float, int) promotes to
    // double, and the difference is rounded to float (model.cpp:914-915)
    const float bc1 = (float)(1.0 - std::pow((double)os_.beta1, (double)adam_step));
    const float bc2 = (float)(1.0 - std::pow((double)os_.beta2, (double)adam_step));
    for (int g = 0; g < 6; g++) {
        hp.step_size[g] = os_.adam_lr[g] / bc1;
        hp.bc2_sqrt[g] = std::sqrt(bc2);
    }
    hp.beta1 = os_.beta1;
    hp.beta2 = os_.beta2;
    hp.eps = os_.eps;
    hp.ssim_weight = ssim_weight;
    // masked training normalises by the pixels that own loss (reference model.cpp:929-934)
    hp.loss_inv_n = (cam.mask && cam.unmasked > 0) ? 1.0f / (float)(cam.unmasked * 3)
                                                   : 1.0f / (float)(cam.height * cam.width * 3);
    hp.op_reg = step >= 500 ? opt_.opacity_reg : 0.0f;
    hp.sc_reg = step >= 500 ? opt_.scale_reg : 0.0f;
    hp.reg_invis_scale = 1.0f;
    for (int c = 0; c < 3; c++) hp.background[c] = opt_.bg[c];
    hp.noise_lr = opt_.noise_lr;
    hp.xyz_lr = os_.adam_lr[0];
    hp.max_disp = opt_.cull_radius > 0.0f ? opt_.cull_radius * 0.0167f : 0.05f;
    hp.sgld_seed = (uint32_t)step;
    hp.sgld_enable = 1;
    hp.degree = degree_;
    hp.degrees_to_use = degrees_to_use(step);
    return hp;
}

void Model::full_iteration(const TrainCamera& cam, int step, float ssim_weight) {
    if (!cam.gt) throw std::runtime_error("camera has no ground-truth image");
    const MsplatCamera c = prepare(cam);
    os_.adam_step_count++;
    const MsplatStepParams hp = step_params(cam, step, ssim_weight, os_.adam_step_count);
    const int rc = eng_.train_step(state(), c, cam.gt, hp, false, (cam.mask && cam.unmasked > 0) ? cam.mask : nullptr);
    if (rc != 0) throw std::runtime_error("training step rejected its inputs (rc=" + std::to_string(rc) + ")");
    if (prop_.enabled) proposal_accumulate(eng_);
}

void Model::full_iteration_batch(const TrainCamera* const* cams, int B, int step, float ssim_weight) {
    if (B <= 0 || B > Engine::kLossSlots) throw std::invalid_argument("batch size must be in [1, 64]");
    for (int v = 0; v < B; v++)
        if (!cams[v] || !cams[v]->gt) throw std::runtime_error("camera has no ground-truth image");
    const size_t rows = (size_t)cap_;
    const size_t sizes[7] = {rows * 12, rows * 12, rows * 16, rows * 4, rows * 12,
                             rows * std::max(row_len(4, degree_), 1) * 4, rows * 4};
    for (int k = 0; k < 7; k++) acc_[k].ensure(sizes[k]);
    msk::BatchGrads A{acc_[0].as<float>(), acc_[1].as<float>(), acc_[2].as<float>(), acc_[3].as<float>(),
                      acc_[4].as<float>(), acc_[5].as<float>(), acc_[6].as<uint32_t>()};
    os_.adam_step_count++;
    std::vector<BatchView> views((size_t)B);
    MsplatStepParams hp{};
    for (int v = 0; v < B; v++) {
        const TrainCamera& cam = *cams[v];
        views[v].cam = prepare(cam);
        views[v].gt = cam.gt;
        const bool masked = cam.mask && cam.unmasked > 0;
        views[v].mask = masked ? cam.mask : nullptr;
        views[v].inv_n = 1.0f / (float)((masked ? cam.unmasked : (int64_t)cam.height * cam.width) * 3 * B);
        hp = step_params(cam, step, ssim_weight, os_.adam_step_count);
    }
    // large-batch learning-rate scaling of the Adam steps (1 = the single-view rates)
    for (int g = 0; g < 6; g++) hp.step_size[g] *= batch_lr_scale_;
    if (!batch_) {
        batch_.reset(new BatchRunner(eng_.stream()));
        batch_->set_persist_grid(persist_grid_);
        batch_->set_raster_ppl(raster_ppl_);
        batch_->set_autotune(batch_autotune_);
    }
    batch_->set_graph(batch_graph_);
    batch_->set_sequential(batch_sequential_);
    batch_->run(state(), views.data(), B, hp, A, eng_.nan_counts());
    batch_n_ = B;
    if (prop_.enabled)
        for (int v = 0; v < B; v++) proposal_accumulate(*batch_->slot(v));
}

std::vector<double> Model::batch_losses() {
    std::vector<double> out;
    for (int v = 0; v < batch_n_; v++) out.push_back(batch_->loss(v));
    return out;
}

void Model::schedulers_step(int step) {
    const float t = std::min(std::max((float)step / (float)opt_.position_lr_max_steps, 0.f), 1.f);
    os_.adam_lr[0] = std::exp(std::log(os_.means_lr_init) * (1.f - t) + std::log(os_.means_lr_final) * t);
}

bool Model::is_mcmc_step(int step) const {
    return mcmc_on_ && step >= 500 && step % 100 == 0 && step < (int)((float)opt_.max_steps * 0.833f);
}

msk::McmcInfo Model::mcmc_after_train(int step) {
    msk::McmcInfo info;
    info.num_active = num_active_;
    if (!is_mcmc_step(step)) return info;
    msk::Params P;
    const MsplatDeviceState st = state();
    P.means = st.means; P.scales = st.scales; P.quats = st.quats; P.f_dc = st.f_dc; P.f_rest = st.f_rest;
    P.opac = st.opac;
    for (int g = 0; g < 6; g++) { P.m[g] = st.m[g]; P.v[g] = st.v[g]; }
    P.N = num_active_;
    P.degree = degree_;
    P.sh_bf16 = st.sh_bf16;
    P.f_dc_h = st.f_dc_bf16;
    P.f_rest_h = st.f_rest_bf16;
    msk::ReloHooks hooks;
    if (prop_.enabled) proposal_before_event(P, hooks);
    info = msk::mcmc_relocate_and_grow(eng_.stream(), P, cap_, opt_.cull_radius, reloc_count_, mcmc_,
                                       prop_.enabled ? &hooks : nullptr);
    if (prop_.enabled) proposal_after_event(P, hooks);
    num_active_ = info.num_active;
    if (step == nan_inject_step_) nan_inject(step);
    return info;
}

void Model::nan_inject(int step) {
    if (num_active_ <= 0) return;
    // 100 draws with replacement among the active rows (the reference draws among its alive
    // list with mt19937(1234); this build's own stream: splitmix64 from 1234)
    static const float kNaN3[3] = {NAN, NAN, NAN};
    uint64_t s = 1234;
    hipStream_t st = eng_.stream();
    for (int k = 0; k < 100; k++) {
        const int64_t i = (int64_t)(mix64(s) % (uint64_t)num_active_);
        hip_check(hipMemcpyAsync(p_[3].as<float>() + 3 * i, kNaN3, sizeof(kNaN3), hipMemcpyHostToDevice, st),
                  "nan inject");
    }
    refresh_shadows();
    hip_check(hipStreamSynchronize(st), "nan inject");
    std::printf("MCMC nan-inject step=%d poisoned=100 (featuresDc)\n", step);
    std::fflush(stdout);
}

const float* Model::render(const TrainCamera& cam, int step) {
    const MsplatCamera c = prepare(cam);
    const int rc = eng_.render(state(), c, degrees_to_use(step), opt_.bg);
    if (rc != 0) throw std::runtime_error("render rejected its inputs (rc=" + std::to_string(rc) + ")");
    return eng_.image();
}

HostGaussians Model::download() const {
    HostGaussians g;
    g.resize(num_active_, degree_);
    hipStream_t s = const_cast<Engine&>(eng_).stream();
    hip_check(hipStreamSynchronize(s), "download");
    for (int k = 0; k < 6; k++) {
        const size_t bytes = (size_t)num_active_ * row_len(k, degree_) * sizeof(float);
        if (bytes) hip_check(hipMemcpy(group_vec(g, k)->data(), p_[k].p, bytes, hipMemcpyDeviceToHost), "download");
    }
    return g;
}

int64_t Model::read_group(int g, int which, float* dst, int64_t count) const {
    const int64_t avail = (int64_t)num_active_ * row_len(g, degree_);
    if (!dst || count <= 0 || avail == 0) return avail;
    const int64_t n = std::min(count, avail);
    hip_check(hipStreamSynchronize(const_cast<Engine&>(eng_).stream()), "read");
    const DBuf& b = which == 0 ? p_[g] : (which == 1 ? m_[g] : v_[g]);
    if (which != 0 && moment_elem_bytes(g) == 2) {
        std::vector<uint16_t> h((size_t)n);
        hip_check(hipMemcpy(h.data(), b.p, (size_t)n * 2, hipMemcpyDeviceToHost), "read");
        for (int64_t e = 0; e < n; e++) dst[e] = ms_bf16_float(h[(size_t)e]);
    } else {
        hip_check(hipMemcpy(dst, b.p, (size_t)n * 4, hipMemcpyDeviceToHost), "read");
    }
    return avail;
}

void Model::download_moments(std::vector<float> m[6], std::vector<float> v[6]) const {
    for (int k = 0; k < 6; k++) {
        const size_t count = (size_t)num_active_ * row_len(k, degree_);
        m[k].assign(count, 0.f);
        v[k].assign(count, 0.f);
        if (count) {
            read_group(k, 1, m[k].data(), (int64_t)count);
            read_group(k, 2, v[k].data(), (int64_t)count);
        }
    }
}

void Model::save_checkpoint(const std::string& path, int step) {
    CheckpointData c;
    c.step = (uint32_t)step;
    c.num = (uint32_t)num_active_;
    c.sh_degree = (uint32_t)opt_.sh_degree;
    c.adam_steps = (uint32_t)os_.adam_step_count;
    std::memcpy(c.adam_lr, os_.adam_lr, sizeof(c.adam_lr));
    c.lr_init = os_.means_lr_init;
    c.lr_final = os_.means_lr_final;
    c.g = download();
    download_moments(c.m, c.v);
    write_checkpoint(path, c);
}

int Model::load_checkpoint(const std::string& path) {
    CheckpointData c = read_checkpoint(path);
    if ((int64_t)c.num > cap_) throw std::runtime_error("checkpoint holds more Gaussians than cap_max");
    // Every degree-dependent buffer (parameters, moments, batch sums, the SH schedule) is
    // sized for the trainer's configured degree, so a checkpoint of another degree is
    // rejected before any state changes (the reference ignores the header's degree,
    // model.cpp:779-783, and would read the rest coefficients with the wrong stride).
    if (c.g.degree != opt_.sh_degree || (int)c.sh_degree != c.g.degree)
        throw std::runtime_error("checkpoint SH degree " + std::to_string(c.g.degree) + " (header " +
                                 std::to_string(c.sh_degree) + ") differs from the trainer's " +
                                 std::to_string(opt_.sh_degree));
    // moments of rows beyond the checkpoint's count start from zero (model.cpp:801-812)
    hipStream_t s = eng_.stream();
    for (int k = 0; k < 6; k++) {
        const size_t rows = (size_t)cap_ * std::max(row_len(k, degree_), 1);
        msk::fill_zero(s, p_[k].p, rows * 4);
        msk::fill_zero(s, m_[k].p, rows * moment_elem_bytes(k));
        msk::fill_zero(s, v_[k].p, rows * moment_elem_bytes(k));
    }
    upload(c.g, c.m, c.v);
    std::memcpy(os_.adam_lr, c.adam_lr, sizeof(os_.adam_lr));
    os_.means_lr_init = c.lr_init;
    os_.means_lr_final = c.lr_final;
    os_.adam_step_count = (int)c.adam_steps;
    return (int)c.step;
}

void Model::save_ply(const std::string& path, int step) {
    write_ply(path, download(), step, opt_.keep_crs, scale_, translation_);
}

void Model::save_splat(const std::string& path) {
    write_splat(path, download(), opt_.keep_crs, scale_, translation_);
}

}  // namespace msplat_hip

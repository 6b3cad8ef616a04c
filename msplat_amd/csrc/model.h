// model.h — trainable Gaussian state on the device and the per-step host bookkeeping
// (reference core/include/model.hpp, core/src/model.cpp).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include <memory>

#include "batch.h"
#include "engine.h"
#include "mcmc.h"
#include "proposal.h"

namespace msplat_hip {

// One camera of a dataset: OpenGL-convention camera-to-world pose (row-major), pinhole
// intrinsics and the device-resident uint8 ground truth [H, W, 3] (reference Camera,
// core/include/input_data.hpp).
struct TrainCamera {
    float c2w[16];
    float fx = 0, fy = 0, cx = 0, cy = 0;
    int width = 0, height = 0;
    const uint8_t* gt = nullptr;
    // training mask (device uint8 [H W], 255 use / 0 ignore) and its count of used pixels;
    // the step is masked only when both are set (mask && unmasked > 0)
    const uint8_t* mask = nullptr;
    int64_t unmasked = -1;
};

// Host copy of a Gaussian set, row-major: means[n,3], scales[n,3] (log), quats[n,4] (w,x,y,z),
// f_dc[n,3], f_rest[n, (degree+1)^2-1, 3], opac[n] (logit).
struct HostGaussians {
    int n = 0, degree = 0;
    std::vector<float> means, scales, quats, f_dc, f_rest, opac;
    int rest_len() const { return ((degree + 1) * (degree + 1) - 1) * 3; }
    void resize(int rows, int deg);
};

struct ModelOptions {
    int sh_degree = 3;
    int sh_degree_interval = 1000;
    int cap_max = 1000000;
    int max_steps = 30000;
    int position_lr_max_steps = 30000;
    float noise_lr = 5e5f, opacity_reg = 0.01f, scale_reg = 0.01f, cull_radius = 3.0f;
    bool keep_crs = false;
    bool decimate_seeds = true;
    float bg[3] = {0.6130f, 0.0101f, 0.3984f};
};

// Adam state and schedule scalars (model.hpp:60-70).
struct OptimState {
    float adam_lr[6] = {0.00016f, 0.005f, 0.001f, 0.0025f, 0.000125f, 0.05f};
    float means_lr_init = 0.00016f, means_lr_final = 0.0000016f;
    int adam_step_count = 0;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

int row_len(int group, int degree);

class Model {
public:
    Model(const ModelOptions& opt, const HostGaussians& init, float scale, const float translation[3]);
    ~Model();
    Model(const Model&) = delete;
    Model& operator=(const Model&) = delete;

    // Model::fullIteration (model.cpp:901-980): one training step on `cam`.
    void full_iteration(const TrainCamera& cam, int step, float ssim_weight);
    // One multi-view batched step over B cameras (one Adam update of the summed gradients;
    // the objective is the mean of the per-view losses). Returns nothing; the per-view losses
    // are in batch_losses().
    void full_iteration_batch(const TrainCamera* const* cams, int B, int step, float ssim_weight);
    std::vector<double> batch_losses();
    // Model::schedulersStep (model.cpp:256-261).
    void schedulers_step(int step);
    // Model::mcmcAfterTrain (model.cpp:497-565), on the device.
    msk::McmcInfo mcmc_after_train(int step);
    // Model::render (model.cpp:891-899): device RGB image [H, W, 3], valid until the next call.
    const float* render(const TrainCamera& cam, int step);
    // The scalars full_iteration hands to the kernels for this step (no state change).
    MsplatStepParams step_params(const TrainCamera& cam, int step, float ssim_weight, int adam_step) const;

    void save_checkpoint(const std::string& path, int step);
    int load_checkpoint(const std::string& path);
    void save_ply(const std::string& path, int step);
    void save_splat(const std::string& path);

    HostGaussians download() const;
    void download_moments(std::vector<float> m[6], std::vector<float> v[6]) const;
    // Copies up to `count` floats of the active rows of group g (which: 0 parameter,
    // 1 exp_avg, 2 exp_avg_sq), converting bf16 storage to float; returns the row-major size.
    int64_t read_group(int g, int which, float* dst, int64_t count) const;
    // SH coefficients and their Adam moments in bfloat16 storage with fp32 masters
    // (converts the current state in place).
    void set_sh_bf16(bool on);
    bool sh_bf16() const { return sh_bf16_; }
    MsplatDeviceState state() const;
    int num_active() const { return num_active_; }
    int capacity() const { return cap_; }
    int degree() const { return degree_; }
    const ModelOptions& options() const { return opt_; }
    OptimState& optim() { return os_; }
    uint32_t relocation_count() const { return reloc_count_; }
    Engine& engine() { return eng_; }
    // Batched steps replay one recorded HIP graph per launch configuration when on.
    void set_batch_graph(bool on) { batch_graph_ = on; }
    void set_batch_sequential(bool on) { batch_sequential_ = on; }
    // Relocation / growth on their schedule (on by default; off only for benchmarks that
    // compare step throughput at a fixed Gaussian count).
    void set_mcmc_enabled(bool on) { mcmc_on_ = on; }
    // Factor on every Adam step size of batched steps (large-batch scaling; 1 by default).
    void set_batch_lr_scale(float f) { batch_lr_scale_ = f; }
    float batch_lr_scale() const { return batch_lr_scale_; }
    void set_batch_autotune(bool on) {
        batch_autotune_ = on;
        if (batch_) batch_->set_autotune(on);
    }
    bool batch_graph() const { return batch_graph_; }
    BatchRunner* batch_runner() { return batch_.get(); }
    // ---- learned relocation proposal (no counterpart in the reference) ----
    // Allocates its state and starts gathering statistics; relocation stays opacity-weighted
    // until the network has trained on the outcomes of at least one event and the judge
    // (the trainer) selects the learned mode.
    void enable_proposal(bool on, uint64_t seed);
    bool proposal_enabled() const { return prop_.enabled; }
    void set_proposal_mode(bool learned) { prop_.use_learned = learned; }
    bool proposal_ready() const { return prop_.ready(); }
    const ProposalInfo& proposal_info() const { return prop_.info; }
    void proposal_params(float* theta) const;
    void set_proposal_params(const float* theta);
    // Copies a proposal buffer to the host: "out" [N 4] network outputs of the last learned
    // event, "features" [N 16] bf16 of the last event, "stats" [5 N] statistics gathered
    // since the last event. Returns the buffer's size in bytes (copies min(size, bytes)).
    int64_t proposal_read(const std::string& name, void* host, int64_t bytes) const;
    // Model::mcmcAfterTrain's schedule (model.cpp:498-501): is `step` a relocation step?
    bool is_mcmc_step(int step) const;
    // Self-test of the non-finite gates (model.cpp:638-659): at the end of the relocation
    // pass of `step` (-1: never; the MSPLAT_NAN_INJECT environment variable sets it at
    // construction) 100 draws among the active rows get NaN DC colours. The gates keep
    // those rows invisible and the next pass culls them.
    void set_nan_inject_step(int step) { nan_inject_step_ = step; }
    // Resident workgroups of the batched persistent raster launches (0 = default).
    void set_persist_grid(int g) {
        persist_grid_ = g;
        if (batch_) batch_->set_persist_grid(g);
    }
    void set_raster_ppl(int p) {
        raster_ppl_ = p;
        if (batch_) batch_->set_raster_ppl(p);
    }
    int degrees_to_use(int step) const;

private:
    void allocate(int degree);
    size_t moment_elem_bytes(int g) const { return (sh_bf16_ && (g == 3 || g == 4)) ? 2 : 4; }
    void refresh_shadows();
    void upload(const HostGaussians& g, const std::vector<float>* m, const std::vector<float>* v);

    ModelOptions opt_;
    OptimState os_;
    int cap_ = 0, degree_ = 0, num_active_ = 0;
    uint32_t reloc_count_ = 0;
    int nan_inject_step_ = -1;
    void nan_inject(int step);
    float scale_ = 1.0f, translation_[3] = {0, 0, 0};
    DBuf p_[6], m_[6], v_[6];
    bool sh_bf16_ = false;
    DBuf dc_h_, rest_h_;  // bf16 shadows of f_dc / f_rest (sh_bf16_ only)
    DBuf acc_[7];         // batched-step gradient sums (allocated on first use)
    Engine eng_;
    std::unique_ptr<BatchRunner> batch_;
    bool batch_graph_ = false;
    bool batch_sequential_ = false;
    bool mcmc_on_ = true;
    float batch_lr_scale_ = 1.0f;
    bool batch_autotune_ = false;
    int batch_n_ = 0;
    int persist_grid_ = 0;
    int raster_ppl_ = 1;
    msk::McmcScratch mcmc_;
    Proposal prop_;
    void proposal_accumulate(const Engine& e);
    void proposal_before_event(const msk::Params& P, msk::ReloHooks& hooks);
    void proposal_after_event(const msk::Params& P, const msk::ReloHooks& hooks);
};

// On-disk formats (reference save_gaussians.cpp:10-231, model.cpp:698-843).
struct CheckpointData {
    uint32_t step = 0, num = 0, sh_degree = 0, adam_steps = 0;
    float adam_lr[6] = {0}, lr_init = 0, lr_final = 0;
    HostGaussians g;
    std::vector<float> m[6], v[6];
};
void write_checkpoint(const std::string& path, const CheckpointData& c);
CheckpointData read_checkpoint(const std::string& path);
void write_ply(const std::string& path, const HostGaussians& g, int step, bool keep_crs, float scale,
               const float translation[3]);
HostGaussians read_ply(const std::string& path, int* step, float scale, const float translation[3], bool keep_crs);
void write_splat(const std::string& path, const HostGaussians& g, bool keep_crs, float scale, const float translation[3]);

}  // namespace msplat_hip

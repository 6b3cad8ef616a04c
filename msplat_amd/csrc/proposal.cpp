// proposal.cpp — the learned relocation proposal (no counterpart in the reference, whose
// relocation samples sources by opacity only: Model::relocate, core/src/model.cpp:336-416).
//
// Between relocation events every training view adds to per-Gaussian statistics (screen
// radius, depth moments, |d loss / d mean2d|). At an event the relocations of the previous
// event are labelled with their outcome (did the copy's opacity grow, where did it move
// relative to its source, in units of its scale) and join a ring of training samples; the
// network first predicts them unseen (validation against the best constant predictor),
// then trains on the ring for a fixed number of Adam epochs. When the trainer's judge
// selects the learned mode, the relocation draw weighs sources by sigmoid(o) exp(z0) and
// places copies at the predicted offset; otherwise it stays opacity-weighted. Both modes
// record their relocations, so the plain sampler's outcomes train the network.
#include "proposal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/msplat_kernels.h"
#include "model.h"

namespace msplat_hip {

void Proposal::release() {
    DBuf* all[] = {&st[0], &st[1], &st[2], &st[3], &st[4], &theta, &m, &v, &grad, &packed, &loss2, &X, &out,
                   &pair_src, &pair_dst, &rec_X, &rec_dst, &rec_o, &rec_mu, &rec_s, &new_X, &new_T, &ring_X, &ring_T};
    for (DBuf* b : all) b->release();
    enabled = false;
}

static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void proposal_init_theta(uint64_t seed, float* th) {
    uint64_t st = seed;
    auto normal = [&]() {
        const double u1 = ((double)(splitmix(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((double)(splitmix(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int k = 0; k < msk::MLP_NPARAM; k++) th[k] = 0.f;
    const int blocks[3][3] = {{msk::MLP_OFF_W1, msk::MLP_HID * msk::MLP_IN, msk::MLP_IN},
                              {msk::MLP_OFF_W2, msk::MLP_HID * msk::MLP_HID, msk::MLP_HID},
                              {msk::MLP_OFF_W3, msk::MLP_OUT * msk::MLP_HID, msk::MLP_HID}};
    for (const auto& b : blocks) {
        const double sd = std::sqrt(2.0 / b[2]);
        for (int k = 0; k < b[1]; k++) th[b[0] + k] = (float)(normal() * sd);
    }
    // the output layer starts small: the first learned draws stay close to the plain ones
    for (int k = 0; k < msk::MLP_OUT * msk::MLP_HID; k++) th[msk::MLP_OFF_W3 + k] *= 0.1f;
}

void Model::enable_proposal(bool on, uint64_t seed) {
    if (!on) {
        hip_check(hipStreamSynchronize(eng_.stream()), "proposal");
        prop_.release();
        prop_ = Proposal();
        return;
    }
    if (prop_.enabled) return;
    hipStream_t s = eng_.stream();
    Proposal& p = prop_;
    const size_t c = (size_t)std::max(cap_, 1);
    p.cap = cap_;
    for (auto& b : p.st) b.ensure(c * 4);
    p.theta.ensure(msk::MLP_NPARAM * 4);
    p.m.ensure(msk::MLP_NPARAM * 4);
    p.v.ensure(msk::MLP_NPARAM * 4);
    p.grad.ensure(msk::MLP_NPARAM * 4);
    p.packed.ensure(sizeof(msk::MlpPacked));
    p.loss2.ensure(64);
    p.X.ensure(c * 32);
    p.out.ensure(c * 16);
    p.pair_src.ensure(c * 4);
    p.pair_dst.ensure(c * 4);
    p.rec_X.ensure(c * 32);
    p.rec_dst.ensure(c * 4);
    p.rec_o.ensure(c * 4);
    p.rec_mu.ensure(c * 12);
    p.rec_s.ensure(c * 4);
    p.new_X.ensure(c * 32);
    p.new_T.ensure(c * 32);
    p.ring_X.ensure((size_t)Proposal::kRing * 32);
    p.ring_T.ensure((size_t)Proposal::kRing * 32);
    std::vector<float> th(msk::MLP_NPARAM);
    proposal_init_theta(seed, th.data());
    hip_check(hipMemcpyAsync(p.theta.p, th.data(), th.size() * 4, hipMemcpyHostToDevice, s), "proposal");
    msk::fill_zero(s, p.m.p, msk::MLP_NPARAM * 4);
    msk::fill_zero(s, p.v.p, msk::MLP_NPARAM * 4);
    for (auto& b : p.st) msk::fill_zero(s, b.p, c * 4);
    msk::mlp_pack(s, p.theta.as<float>(), p.packed.as<msk::MlpPacked>());
    hip_check(hipStreamSynchronize(s), "proposal");
    p.enabled = true;
}

void Model::proposal_params(float* theta) const {
    if (!prop_.enabled) throw std::runtime_error("learned relocation is not enabled");
    hipStream_t s = const_cast<Engine&>(eng_).stream();
    hip_check(hipMemcpyAsync(theta, prop_.theta.p, msk::MLP_NPARAM * 4, hipMemcpyDeviceToHost, s), "proposal");
    hip_check(hipStreamSynchronize(s), "proposal");
}

void Model::set_proposal_params(const float* theta) {
    if (!prop_.enabled) throw std::runtime_error("learned relocation is not enabled");
    hipStream_t s = eng_.stream();
    hip_check(hipMemcpyAsync(prop_.theta.p, theta, msk::MLP_NPARAM * 4, hipMemcpyHostToDevice, s), "proposal");
    msk::mlp_pack(s, prop_.theta.as<float>(), prop_.packed.as<msk::MlpPacked>());
    hip_check(hipStreamSynchronize(s), "proposal");
    prop_.info.trained = std::max(prop_.info.trained, 1);  // parameters given: the network counts as trained
}

int64_t Model::proposal_read(const std::string& name, void* host, int64_t bytes) const {
    if (!prop_.enabled) throw std::runtime_error("learned relocation is not enabled");
    hipStream_t s = const_cast<Engine&>(eng_).stream();
    const size_t n = (size_t)num_active_;
    auto take = [&](const DBuf& b, size_t size, size_t at) {
        if (host && bytes > (int64_t)at) {
            const size_t c = std::min(size, (size_t)bytes - at);
            hip_check(hipMemcpyAsync((uint8_t*)host + at, b.p, c, hipMemcpyDeviceToHost, s), "proposal");
        }
    };
    int64_t total = 0;
    if (name == "out") {
        take(prop_.out, n * 16, 0);
        total = (int64_t)n * 16;
    } else if (name == "features") {
        take(prop_.X, n * 32, 0);
        total = (int64_t)n * 32;
    } else if (name == "stats") {
        for (int k = 0; k < 5; k++) take(prop_.st[k], n * 4, (size_t)k * n * 4);
        total = (int64_t)n * 20;
    } else {
        throw std::invalid_argument("unknown proposal buffer '" + name + "'");
    }
    hip_check(hipStreamSynchronize(s), "proposal");
    return total;
}

void Model::proposal_accumulate(const Engine& e) {
    msk::relo_stats_accum(eng_.stream(), num_active_, e.radii_dev(), e.depths_dev(), e.g2d_dev(), prop_.stats());
    prop_.views++;
}

void Model::proposal_before_event(const msk::Params& P, msk::ReloHooks& hooks) {
    hipStream_t s = eng_.stream();
    Proposal& p = prop_;
    const int n = p.rec_n;
    if (n > 0) {
        // outcomes of the previous event's relocations
        msk::relo_label(s, n, p.records(), P, 1.f / (float)n, 1.f, p.new_X.as<uint16_t>(), p.new_T.as<float>());
        msk::relo_weights(s, n, p.new_T.as<float>(), 1.f / (float)n, Proposal::kOffsetWeight / (float)n);
        // the network on them before it trains on them
        msk::fill_zero(s, p.grad.p, msk::MLP_NPARAM * 4);
        msk::fill_zero(s, p.loss2.p, 16);
        msk::mlp_backward(s, n, p.new_X.as<uint16_t>(), p.new_T.as<float>(), p.packed.as<msk::MlpPacked>(),
                          p.grad.as<float>(), p.loss2.as<double>());
        double l2[2] = {0, 0};
        const int nb = std::min(n, 65536);
        std::vector<float> T((size_t)nb * 8);
        hip_check(hipMemcpyAsync(l2, p.loss2.p, 16, hipMemcpyDeviceToHost, s), "proposal");
        hip_check(hipMemcpyAsync(T.data(), p.new_T.p, T.size() * 4, hipMemcpyDeviceToHost, s), "proposal");
        hip_check(hipStreamSynchronize(s), "proposal");
        double ysum = 0, ok = 0, tsum[3] = {0, 0, 0}, tsq[3] = {0, 0, 0};
        for (int j = 0; j < nb; j++) {
            ysum += T[8 * j];
            if (T[8 * j + 5] != 0.f) {
                ok += 1;
                for (int k = 0; k < 3; k++) {
                    tsum[k] += T[8 * j + 2 + k];
                    tsq[k] += (double)T[8 * j + 2 + k] * T[8 * j + 2 + k];
                }
            }
        }
        const double yb = ysum / nb;
        p.info.base_bce = (yb <= 0.0 || yb >= 1.0) ? 0.0 : -(yb * std::log(yb) + (1 - yb) * std::log(1 - yb));
        double var = 0;
        for (int k = 0; k < 3; k++)
            if (ok > 0) var += tsq[k] / ok - (tsum[k] / ok) * (tsum[k] / ok);
        p.info.base_off = var;
        p.info.val_bce = l2[0];
        p.info.val_off = l2[1] / (double)Proposal::kOffsetWeight * ((double)n / std::max(ok * n / nb, 1.0));
        // into the ring (newest overwrite oldest)
        for (int done = 0; done < n;) {
            const int take = std::min(n - done, Proposal::kRing - p.ring_head);
            msk::copy16(s, p.new_X.as<uint8_t>() + (size_t)done * 32, p.ring_X.as<uint8_t>() + (size_t)p.ring_head * 32,
                        (size_t)take * 32);
            msk::copy16(s, p.new_T.as<uint8_t>() + (size_t)done * 32, p.ring_T.as<uint8_t>() + (size_t)p.ring_head * 32,
                        (size_t)take * 32);
            p.ring_head = (p.ring_head + take) % Proposal::kRing;
            done += take;
        }
        p.ring_count = std::min<int64_t>(p.ring_count + n, Proposal::kRing);
        p.info.samples = p.ring_count;
        // train on the ring
        const int cnt = (int)p.ring_count;
        msk::relo_weights(s, cnt, p.ring_T.as<float>(), 1.f / (float)cnt, Proposal::kOffsetWeight / (float)cnt);
        for (int e = 0; e < Proposal::kEpochs; e++) {
            msk::fill_zero(s, p.grad.p, msk::MLP_NPARAM * 4);
            msk::fill_zero(s, p.loss2.p, 16);
            msk::mlp_backward(s, cnt, p.ring_X.as<uint16_t>(), p.ring_T.as<float>(), p.packed.as<msk::MlpPacked>(),
                              p.grad.as<float>(), p.loss2.as<double>());
            p.adam_t++;
            const float bc1 = 1.f - std::pow(0.9f, (float)p.adam_t);
            const float bc2s = std::sqrt(1.f - std::pow(0.999f, (float)p.adam_t));
            msk::mlp_adam(s, p.theta.as<float>(), p.m.as<float>(), p.v.as<float>(), p.grad.as<float>(), Proposal::kLr,
                          0.9f, 0.999f, 1e-8f, bc1, bc2s, p.packed.as<msk::MlpPacked>());
        }
        p.info.trained++;
    }
    // features of the rows as they enter this event
    msk::relo_features(s, P, p.stats(), p.views, p.X.as<uint16_t>());
    if (p.use_learned && p.ready()) {
        msk::mlp_forward(s, P.N, p.X.as<uint16_t>(), p.packed.as<msk::MlpPacked>(), p.out.as<float>());
        hooks.mlp_out = p.out.as<float>();
        p.info.learned_events++;
    } else {
        p.info.plain_events++;
    }
    hooks.pair_src = p.pair_src.as<uint32_t>();
    hooks.pair_dst = p.pair_dst.as<uint32_t>();
    hip_check(hipGetLastError(), "proposal");
}

void Model::proposal_after_event(const msk::Params& P, const msk::ReloHooks& hooks) {
    hipStream_t s = eng_.stream();
    Proposal& p = prop_;
    p.rec_n = hooks.pairs;
    msk::relo_record(s, p.rec_n, hooks.pair_src, hooks.pair_dst, p.X.as<uint16_t>(), P, p.records());
    for (auto& b : p.st) msk::fill_zero(s, b.p, (size_t)std::max(cap_, 1) * 4);
    p.views = 0;
    p.info.events++;
    hip_check(hipGetLastError(), "proposal");
}

}  // namespace msplat_hip

// ---------------------------------------------------------------------------------------
// kernel-level C ABI of the network (msplat_kernels.h)
extern "C" {

size_t msplat_mlp_packed_bytes(void) { return sizeof(msk::MlpPacked); }

int msplat_mlp_pack(const float* theta, void* packed) {
    if (!theta || !packed) return -1;
    msk::mlp_pack(nullptr, theta, reinterpret_cast<msk::MlpPacked*>(packed));
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

int msplat_mlp_forward(int n, const uint16_t* X, const void* packed, float* out) {
    if (n < 0 || (n > 0 && (!X || !packed || !out))) return -1;
    msk::mlp_forward(nullptr, n, X, reinterpret_cast<const msk::MlpPacked*>(packed), out);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

int msplat_mlp_backward(int n, const uint16_t* X, const float* targets, const void* packed, float* grad,
                        double* loss2) {
    if (n < 0 || (n > 0 && (!X || !targets || !packed || !grad || !loss2))) return -1;
    msk::mlp_backward(nullptr, n, X, targets, reinterpret_cast<const msk::MlpPacked*>(packed), grad, loss2);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

}  // extern "C"

// proposal.h — state of the learned relocation proposal held by a Model: per-Gaussian
// statistics since the last relocation event, the network (fp32 masters, Adam moments,
// packed bf16 operands), the relocations of the last event awaiting their outcome, and a
// ring of labelled training samples.
#pragma once
#include <cstdint>

#include "engine.h"
#include "mlp.h"

namespace msplat_hip {

struct ProposalInfo {
    int events = 0, learned_events = 0, plain_events = 0;
    int64_t samples = 0;
    // on the newest event's relocations, before the network trained on them: the
    // network's logistic loss and the best constant predictor's (the label entropy),
    // and the offset loss of the network and of the best constant offset
    double val_bce = 0, base_bce = 0, val_off = 0, base_off = 0;
    int trained = 0;   // training rounds run
};

struct Proposal {
    static constexpr int kRing = 1 << 18;   // training samples kept (newest first out)
    static constexpr int kEpochs = 40;
    static constexpr float kLr = 2e-3f;
    static constexpr float kOffsetWeight = 0.25f;

    bool enabled = false;
    bool use_learned = false;   // mode of the next relocation event (set by the judge)
    int views = 0;              // views accumulated into the statistics since the last event
    int cap = 0;
    DBuf st[5];
    DBuf theta, m, v, grad, packed, loss2;
    int adam_t = 0;
    DBuf X, out, pair_src, pair_dst;
    int rec_n = 0;
    DBuf rec_X, rec_dst, rec_o, rec_mu, rec_s;
    DBuf new_X, new_T;          // the newest labelled samples (validation before training)
    DBuf ring_X, ring_T;
    int64_t ring_count = 0;
    int ring_head = 0;
    ProposalInfo info;

    msk::ReloStats stats() const {
        return {st[0].as<float>(), st[1].as<float>(), st[2].as<float>(), st[3].as<float>(), st[4].as<float>()};
    }
    msk::ReloRecords records() const {
        return {rec_X.as<uint16_t>(), rec_dst.as<uint32_t>(), rec_o.as<float>(), rec_mu.as<float>(),
                rec_s.as<float>()};
    }
    bool ready() const { return info.trained > 0; }
    void release();
};

// He-initialised parameters (biases zero) from a splitmix64 / Box-Muller stream.
void proposal_init_theta(uint64_t seed, float* theta);

}  // namespace msplat_hip

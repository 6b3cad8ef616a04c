// oracle_mcmc.hpp — CPU restatement of the MCMC relocation / growth pass (TEST INFRASTRUCTURE).
//
// Follows Model::mcmcAfterTrain (model:497-565), relocate (model:336-416),
// addNewGaussians (model:419-495) and computeRelocation (model:311-332).
// The reference samples with libc++'s std::discrete_distribution over mt19937, whose
// stream is implementation-defined. This build defines its own sampler, shared in
// definition (not in code) by this restatement and the GPU kernels:
//   weight_k = floor(sigmoid(o_k) * 2^32)             (u64 fixed point; 0 if not finite)
//   cdf      = inclusive prefix sum of weights         (u64, exact and associative)
//   draw j   = splitmix64((seed << 32) | j)
//   target   = floor(draw * total / 2^64)              (128-bit product)
//   source   = first k with cdf_k > target             (upper bound)
// with seed = 0x9E3779B9 ^ count for relocation and 0xA24BAED4 ^ count for growth, count
// incremented before each use (model:368, model:450). All of it is integer arithmetic, so
// GPU and CPU pick bit-identical sources. Pins: chi-square test of the source histogram
// against sigmoid(o) weights (tests/test_oracle.py), the reference's own Eq. 9 known-answer
// cases (tests/test_relocation_fixture.py).
#pragma once
#include "oracle.hpp"

namespace oracle {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint64_t reloc_weight(float logit) {
    float p = 1.0f / (1.0f + ms_expf(-logit));
    if (!std::isfinite(p) || !(p > 0.0f)) return 0;
    return (uint64_t)((double)p * 4294967296.0);
}

// Learned proposal weight (this build's own design): sigmoid(o) exp(clamp(z, -8, 8)) in
// 2^-20 units.
inline uint64_t learned_weight(float logit, float z) {
    float p = 1.0f / (1.0f + ms_expf(-logit));
    if (!std::isfinite(p) || !(p > 0.0f) || !std::isfinite(z)) return 0;
    const float e = ms_expf(std::min(std::max(z, -8.0f), 8.0f));
    return (uint64_t)((double)p * (double)e * 1048576.0);
}

inline uint64_t draw_target(uint32_t seed, uint32_t j, uint64_t total) {
    uint64_t u = splitmix64(((uint64_t)seed << 32) | (uint64_t)j);
    return (uint64_t)(((unsigned __int128)u * (unsigned __int128)total) >> 64);
}

// computeRelocation (model:311-332). The opacity root is taken in double and rounded
// (the reference takes it in float); the binomial series is the reference's, in double.
inline void relocation_eq9(float o_old, const float* s_old, int N, float* o_new_out, float* s_new) {
    if (N < 1) N = 1;
    if (N > 51) N = 51;
    float o_new = (float)(1.0 - std::pow(1.0 - (double)o_old, 1.0 / (double)N));
    const double on = (double)o_new;
    double denom = 0.0, binom = 1.0;
    for (int k = 1; k <= N; k++) {
        binom = binom * (double)(N - k + 1) / (double)k;
        const double term = binom * std::pow(on, (double)k) / std::sqrt((double)k);
        denom += (k % 2 == 1) ? term : -term;
    }
    double coeff = (denom > 1e-12) ? ((double)o_old / denom) : 1.0;
    if (!std::isfinite(coeff) || coeff <= 0.0) coeff = 1.0;
    for (int i = 0; i < 3; i++) s_new[i] = (float)((double)s_old[i] * coeff);
    *o_new_out = o_new;
}

struct McmcInfo {
    int num_active = 0, dead = 0, grown = 0, nonfinite = 0, culled = 0;
    uint32_t reloc_count_out = 0;
};

template <class T> inline int row_len(const State<T>& S, int g) {
    static const int base[6] = {3, 3, 4, 3, 0, 1};
    return g == 4 ? (num_sh_bases(S.degree) - 1) * 3 : base[g];
}

inline void zero_moments(State<float>& S, int idx) {
    for (int g = 0; g < 6; g++) {
        int L = row_len(S, g);
        std::memset(S.m[g] + (size_t)idx * L, 0, L * sizeof(float));
        std::memset(S.v[g] + (size_t)idx * L, 0, L * sizeof(float));
    }
}

inline void copy_params(State<float>& S, int dst, int src) {
    const int nr = (num_sh_bases(S.degree) - 1) * 3;
    std::memcpy(S.means + 3 * (size_t)dst, S.means + 3 * (size_t)src, 12);
    std::memcpy(S.quats + 4 * (size_t)dst, S.quats + 4 * (size_t)src, 16);
    std::memcpy(S.f_dc + 3 * (size_t)dst, S.f_dc + 3 * (size_t)src, 12);
    if (nr) std::memcpy(S.f_rest + (size_t)nr * dst, S.f_rest + (size_t)nr * src, nr * sizeof(float));
    S.opac[dst] = S.opac[src];
    std::memcpy(S.scales + 3 * (size_t)dst, S.scales + 3 * (size_t)src, 12);
}

inline void apply_eq9(State<float>& S, int src, int count) {
    float op_old = 1.0f / (1.0f + ms_expf(-S.opac[src]));
    float s_old[3] = {ms_expf(S.scales[3 * (size_t)src]), ms_expf(S.scales[3 * (size_t)src + 1]),
                      ms_expf(S.scales[3 * (size_t)src + 2])};
    float op_new, s_new[3];
    relocation_eq9(op_old, s_old, count + 1, &op_new, s_new);
    op_new = std::min(std::max(op_new, 1e-6f), 1.0f - 1e-6f);
    S.opac[src] = (float)std::log((double)op_new / (1.0 - (double)op_new));
    for (int j = 0; j < 3; j++) S.scales[3 * (size_t)src + j] = (float)std::log((double)std::max(s_new[j], 1e-10f));
    zero_moments(S, src);
}

// Samples `num` sources over candidate rows `cand` (their weights given), returns the
// chosen candidate positions and per-candidate counts.
// The counter advances only when a draw happens (model:366-368, model:448-450).
inline bool sample_sources(const std::vector<uint64_t>& w, uint32_t seed_base, uint32_t& count, int num,
                           std::vector<int>& chosen, std::vector<int>& counts) {
    std::vector<uint64_t> cdf(w.size());
    uint64_t acc = 0;
    for (size_t k = 0; k < w.size(); k++) { acc += w[k]; cdf[k] = acc; }
    if (acc == 0) return false;
    const uint32_t seed = seed_base ^ (++count);
    chosen.assign(num, 0);
    counts.assign(w.size(), 0);
    for (int j = 0; j < num; j++) {
        uint64_t r = draw_target(seed, (uint32_t)j, acc);
        size_t k = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
        chosen[j] = (int)k;
        counts[k]++;
    }
    return true;
}

// mlp_out (optional, [N 4]): the learned proposal's weights and placement offsets for the
// relocation draw; growth stays opacity-weighted.
inline void mcmc_relocate_and_grow(State<float>& S, int cap_max, float cull_radius, uint32_t count, McmcInfo& info,
                                   const float* mlp_out = nullptr) {
    const int N = S.N;
    const int nr = (num_sh_bases(S.degree) - 1) * 3;
    const float r2 = cull_radius > 0.0f ? cull_radius * cull_radius : -1.0f;
    std::vector<int> dead, alive;
    for (int i = 0; i < N; i++) {
        bool fin = std::isfinite(S.opac[i]);
        for (int k = 0; k < 3 && fin; k++) fin = std::isfinite(S.means[3 * (size_t)i + k]) && std::isfinite(S.scales[3 * (size_t)i + k]);
        for (int k = 0; k < 4 && fin; k++) fin = std::isfinite(S.quats[4 * (size_t)i + k]);
        for (int k = 0; k < 3 && fin; k++) fin = std::isfinite(S.f_dc[3 * (size_t)i + k]);
        for (int k = 0; k < nr && fin; k++) fin = std::isfinite(S.f_rest[(size_t)nr * i + k]);
        if (!fin) info.nonfinite++;
        float sig = 1.0f / (1.0f + ms_expf(-S.opac[i]));
        bool is_dead = !fin || (sig <= 0.005f);
        if (!is_dead && r2 > 0.0f) {
            float x = S.means[3 * (size_t)i], y = S.means[3 * (size_t)i + 1], z = S.means[3 * (size_t)i + 2];
            if (x * x + y * y + z * z > r2) { is_dead = true; info.culled++; }
        }
        (is_dead ? dead : alive).push_back(i);
    }
    info.dead = (int)dead.size();
    if (!dead.empty() && !alive.empty()) {
        std::vector<uint64_t> w(alive.size());
        for (size_t k = 0; k < alive.size(); k++)
            w[k] = mlp_out ? learned_weight(S.opac[alive[k]], mlp_out[4 * (size_t)alive[k]])
                           : reloc_weight(S.opac[alive[k]]);
        std::vector<int> chosen, counts;
        if (sample_sources(w, 0x9E3779B9u, count, (int)dead.size(), chosen, counts)) {
            for (size_t k = 0; k < alive.size(); k++)
                if (counts[k] > 0) apply_eq9(S, alive[k], counts[k]);
            for (size_t j = 0; j < dead.size(); j++) {
                const int d = dead[j], src = alive[chosen[j]];
                copy_params(S, d, src);
                zero_moments(S, d);
                if (mlp_out) {
                    const float* ls = S.scales + 3 * (size_t)d;
                    const float sc = ms_expf(std::max(std::max(ls[0], ls[1]), ls[2]));
                    for (int k = 0; k < 3; k++) {
                        const float o = mlp_out[4 * (size_t)src + 1 + k];
                        const float oc = std::isfinite(o) ? std::min(std::max(o, -3.0f), 3.0f) : 0.0f;
                        S.means[3 * (size_t)d + k] = std::fma(sc, oc, S.means[3 * (size_t)d + k]);
                    }
                }
            }
        }
    }
    // addNewGaussians
    int target = std::min(cap_max, (int)(1.05f * (float)N));
    int num_new = target - N;
    if (num_new > 0) {
        std::vector<uint64_t> w(N);
        for (int i = 0; i < N; i++) w[i] = reloc_weight(S.opac[i]);
        std::vector<int> chosen, counts;
        if (sample_sources(w, 0xA24BAED4u, count, num_new, chosen, counts)) {
            for (int i = 0; i < N; i++)
                if (counts[i] > 0) apply_eq9(S, i, counts[i]);
            for (int j = 0; j < num_new; j++) {
                copy_params(S, N + j, chosen[j]);
                zero_moments(S, N + j);
            }
            S.N = N + num_new;
            info.grown = num_new;
        }
    }
    info.num_active = S.N;
    info.reloc_count_out = count;
}

}  // namespace oracle

// oracle_mlp.hpp — CPU restatement of the learned relocation proposal (TEST
// INFRASTRUCTURE; never linked by the product). The network and its features have no
// counterpart in the reference (its relocation samples by opacity only, core/src/model.cpp:
// 336-416); this file restates this project's own design so that the device kernels can be
// checked: a 16 -> 64 -> 64 -> 4 MLP with tanh-GELU, evaluated two ways:
//   * "device" rounding: weights, inputs and hidden activations rounded to bfloat16 where
//     the matrix-core kernels round them, fp32 accumulation, backward operands rounded to
//     bf16 as they enter the matrix products, bias gradients from the fp32 values;
//   * "exact": double precision throughout, no rounding, which the finite-difference
//     gradient check pins.
// Parameter vector layout: W1 [64 16], b1 [64], W2 [64 64], b2 [64], W3 [4 64], b3 [4].
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../msplat_amd/csrc/ms_math.h"

namespace oracle {

constexpr int MI = 16, MH = 64, MO = 4;
constexpr int OW1 = 0, OB1 = 1024, OW2 = 1088, OB2 = 5184, OW3 = 5248, OB3 = 5504, NPAR = 5508;

template <class T> inline T gelu_t(T x) {
    const T u = (T)0.7978845608 * (x + (T)0.044715 * x * x * x);
    return (T)0.5 * x * ((T)1 + std::tanh(u));
}
template <class T> inline T gelu_dt(T x) {
    const T u = (T)0.7978845608 * (x + (T)0.044715 * x * x * x);
    const T t = std::tanh(u);
    return (T)0.5 * ((T)1 + t) + (T)0.5 * x * ((T)1 - t * t) * (T)0.7978845608 * ((T)1 + (T)0.134145 * x * x);
}

inline float bf(uint16_t h) { return ms_bf16_float(h); }
inline float rb(float x) { return ms_bf16_round(x); }

// Forward of one sample. X: 16 bf16 inputs. exact: double, no rounding (weights and inputs
// as given). Outputs z3[4]; optional intermediates for the backward.
template <class T> struct MlpActs {
    T z1[MH], h1[MH], z2[MH], h2[MH], z3[MO];
};

template <class T>
inline void mlp_forward_one(const float* th, const float* x, bool exact, MlpActs<T>& A) {
    auto W = [&](int off, int k) -> T { return exact ? (T)th[off + k] : (T)rb(th[off + k]); };
    for (int u = 0; u < MH; u++) {
        T acc = (T)th[OB1 + u];
        for (int f = 0; f < MI; f++) acc += W(OW1, u * MI + f) * (T)x[f];
        A.z1[u] = acc;
        const T h = gelu_t<T>(acc);
        A.h1[u] = exact ? h : (T)rb((float)h);
    }
    for (int u = 0; u < MH; u++) {
        T acc = (T)th[OB2 + u];
        for (int v = 0; v < MH; v++) acc += W(OW2, u * MH + v) * A.h1[v];
        A.z2[u] = acc;
        const T h = gelu_t<T>(acc);
        A.h2[u] = exact ? h : (T)rb((float)h);
    }
    for (int o = 0; o < MO; o++) {
        T acc = (T)th[OB3 + o];
        for (int v = 0; v < MH; v++) acc += W(OW3, o * MH + v) * A.h2[v];
        A.z3[o] = acc;
    }
}

// Loss terms and d loss / d z3 of one sample; tgt: y, wl, t0, t1, t2, wo.
template <class T> inline void mlp_loss_one(const T* z3, const float* tgt, T dz3[MO], double& l_logit, double& l_off) {
    const T z = z3[0], y = (T)tgt[0], wl = (T)tgt[1], wo = (T)tgt[5];
    const T sg = (T)1 / ((T)1 + std::exp(-z));
    dz3[0] = wl * (sg - y);
    l_logit += (double)(wl * (std::max(z, (T)0) - y * z + std::log1p(std::exp(-std::fabs(z)))));
    for (int k = 0; k < 3; k++) {
        const T d = z3[1 + k] - (T)tgt[2 + k];
        dz3[1 + k] = (T)2 * wo * d;
        l_off += (double)(wo * d * d);
    }
}

// Gradient of the summed loss over N samples (X: [N 16] bf16, tgt: [N 8]).
template <class T>
inline void mlp_backward(const float* th, const uint16_t* X, const float* tgt, int N, bool exact, T* grad,
                         double loss2[2]) {
    auto W = [&](int off, int k) -> T { return exact ? (T)th[off + k] : (T)rb(th[off + k]); };
    auto R = [&](T x) -> T { return exact ? x : (T)rb((float)x); };
    for (int k = 0; k < NPAR; k++) grad[k] = 0;
    loss2[0] = loss2[1] = 0.0;
    for (int g = 0; g < N; g++) {
        float x[MI];
        for (int f = 0; f < MI; f++) x[f] = bf(X[(size_t)g * MI + f]);
        MlpActs<T> A;
        mlp_forward_one<T>(th, x, exact, A);
        T dz3[MO];
        mlp_loss_one<T>(A.z3, tgt + (size_t)g * 8, dz3, loss2[0], loss2[1]);
        T dz3r[MO], dz2[MH], dz2r[MH], dz1[MH], dz1r[MH];
        for (int o = 0; o < MO; o++) dz3r[o] = R(dz3[o]);
        for (int v = 0; v < MH; v++) {
            T acc = 0;
            for (int o = 0; o < MO; o++) acc += W(OW3, o * MH + v) * dz3r[o];
            dz2[v] = acc * gelu_dt<T>(A.z2[v]);
            dz2r[v] = R(dz2[v]);
        }
        for (int v = 0; v < MH; v++) {
            T acc = 0;
            for (int u = 0; u < MH; u++) acc += W(OW2, u * MH + v) * dz2r[u];
            dz1[v] = acc * gelu_dt<T>(A.z1[v]);
            dz1r[v] = R(dz1[v]);
        }
        for (int o = 0; o < MO; o++) {
            grad[OB3 + o] += dz3[o];
            for (int v = 0; v < MH; v++) grad[OW3 + o * MH + v] += dz3r[o] * A.h2[v];
        }
        for (int u = 0; u < MH; u++) {
            grad[OB2 + u] += dz2[u];
            for (int v = 0; v < MH; v++) grad[OW2 + u * MH + v] += dz2r[u] * A.h1[v];
            grad[OB1 + u] += dz1[u];
            for (int f = 0; f < MI; f++) grad[OW1 + u * MI + f] += dz1r[u] * (T)x[f];
        }
    }
}

// Summed loss only (the finite-difference check's objective), exact arithmetic.
inline double mlp_loss_exact(const double* th, const uint16_t* X, const float* tgt, int N) {
    double l0 = 0.0, l1 = 0.0;
    for (int g = 0; g < N; g++) {
        double x[MI];
        for (int f = 0; f < MI; f++) x[f] = bf(X[(size_t)g * MI + f]);
        double h1[MH], h2[MH], z3[MO];
        for (int u = 0; u < MH; u++) {
            double acc = th[OB1 + u];
            for (int f = 0; f < MI; f++) acc += th[OW1 + u * MI + f] * x[f];
            h1[u] = gelu_t<double>(acc);
        }
        for (int u = 0; u < MH; u++) {
            double acc = th[OB2 + u];
            for (int v = 0; v < MH; v++) acc += th[OW2 + u * MH + v] * h1[v];
            h2[u] = gelu_t<double>(acc);
        }
        for (int o = 0; o < MO; o++) {
            double acc = th[OB3 + o];
            for (int v = 0; v < MH; v++) acc += th[OW3 + o * MH + v] * h2[v];
            z3[o] = acc;
        }
        double dz[MO];
        mlp_loss_one<double>(z3, tgt + (size_t)g * 8, dz, l0, l1);
    }
    return l0 + l1;
}

}  // namespace oracle

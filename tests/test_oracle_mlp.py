"""Pins of the learned-proposal network's CPU restatement (oracle/oracle_mlp.hpp) before it
checks the device kernels: the exact (double) gradient against central finite differences
of the exact loss on every parameter block, closed forms of GELU / the logistic loss / a
network with zero weights, and the size of the bf16 rounding the device path adds (its error
against the exact path, stated here as the tolerance the GPU test inherits)."""
import numpy as np

import oracle_py as O


def _problem(n=12, seed=3):
    rng = np.random.default_rng(seed)
    th = O.mlp_init(seed)
    th[1024:1088] = rng.normal(0, 0.1, 64)        # non-zero biases exercise their gradients
    th[5184:5248] = rng.normal(0, 0.1, 64)
    th[5504:5508] = rng.normal(0, 0.1, 4)
    X = O.bf16_bits(rng.normal(0, 1, (n, 16)).astype(np.float32))
    tgt = np.zeros((n, 8), np.float32)
    tgt[:, 0] = rng.integers(0, 2, n)
    tgt[:, 1] = 1.0 / n
    tgt[:, 2:5] = rng.normal(0, 1, (n, 3))
    tgt[:, 5] = 0.25 / n
    tgt[n // 3, 5] = 0.0                            # a sample without an offset term
    return th, X, tgt


def test_exact_gradient_matches_finite_differences():
    th, X, tgt = _problem()
    g, _ = O.mlp_backward(th, X, tgt, exact=True)
    rng = np.random.default_rng(7)
    th64 = th.astype(np.float64)
    h = 1e-5
    for name, (a, b) in O.MLP_BLOCKS.items():
        idx = rng.choice(np.arange(a, b), size=min(12, b - a), replace=False)
        for k in idx:
            tp, tm = th64.copy(), th64.copy()
            tp[k] += h
            tm[k] -= h
            fd = (O.mlp_loss_exact(tp, X, tgt) - O.mlp_loss_exact(tm, X, tgt)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (name, int(k), fd, g[k])


def test_closed_forms():
    th = np.zeros(O.MLP_NPARAM, np.float32)
    th[5504:5508] = [0.5, -1.0, 2.0, 0.25]
    X = O.bf16_bits(np.random.default_rng(1).normal(0, 1, (5, 16)).astype(np.float32))
    # zero weights: the output is the last bias exactly, in both arithmetic modes
    for exact in (False, True):
        out = O.mlp_forward(th, X, exact=exact)
        assert np.array_equal(out, np.tile([0.5, -1.0, 2.0, 0.25], (5, 1)))
    # logistic loss at z = 0 is log 2 per unit weight; its gradient is (1/2 - y)
    th[5504] = 0.0
    tgt = np.zeros((5, 8), np.float32)
    tgt[:, 0] = [0, 1, 0, 1, 1]
    tgt[:, 1] = 1.0
    g, (ll, lo) = O.mlp_backward(th, X, tgt, exact=True)
    assert abs(ll - 5 * np.log(2.0)) < 1e-12 and lo == 0.0
    assert abs(g[5504] - (5 * 0.5 - 3)) < 1e-12
    # GELU: g(0) = 0, g'(0) = 1/2, g(x) -> x for large x (through a one-unit network)
    th = np.zeros(O.MLP_NPARAM, np.float32)
    th[0] = 1.0                     # unit 0 reads input 0
    th[1088] = 1.0                  # layer-2 unit 0 reads unit 0
    th[5248] = 1.0                  # output 0 reads layer-2 unit 0
    for x0, want in ((0.0, 0.0), (8.0, 8.0)):
        Xs = O.bf16_bits(np.array([[x0] + [0.0] * 15], np.float32))
        out = O.mlp_forward(th, Xs, exact=True)
        assert abs(out[0, 0] - want) < 1e-9


def test_device_rounding_error_is_bf16_sized():
    th, X, tgt = _problem(n=64, seed=11)
    dev = O.mlp_forward(th, X, exact=False)
    ex = O.mlp_forward(th, X, exact=True)
    scale = max(1.0, float(np.abs(ex).max()))
    err = float(np.abs(dev - ex).max()) / scale
    assert err < 2e-2, err      # bf16 weights and activations: about 2^-8 relative per rounding
    gd, _ = O.mlp_backward(th, X, tgt, exact=False)
    ge, _ = O.mlp_backward(th, X, tgt, exact=True)
    for name, (a, b) in O.MLP_BLOCKS.items():
        rel = np.linalg.norm(gd[a:b] - ge[a:b]) / max(np.linalg.norm(ge[a:b]), 1e-12)
        assert rel < 5e-2, (name, rel)

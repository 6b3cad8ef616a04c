"""The project-written exp / log / sincos (msplat_amd/csrc/ms_math.h, shared by the HIP
kernels and the oracle) against double-precision libm: at most 2 ulp of float error over
the ranges the training step uses (and beyond)."""
import ctypes as C

import numpy as np

import oracle_py as O


def _ulp_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    return np.abs(got - want) / ulp


def test_expf_ulp():
    L = O.lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-87.0, 88.0, 20000), rng.uniform(-5.55, 0.0, 20000),
                         rng.uniform(-15.0, 15.0, 20000), np.array([0.0, -0.0, 1.0, -1.0, 0.5])]).astype(np.float32)
    got = np.array([L.oracle_ms_expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64))
    assert _ulp_err(got, want).max() <= 2.0
    assert L.oracle_ms_expf(0.0) == 1.0
    assert L.oracle_ms_expf(-200.0) == 0.0
    assert np.isinf(L.oracle_ms_expf(100.0))
    assert np.isnan(L.oracle_ms_expf(float("nan")))


def test_logf_ulp():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-7, 1.0, 20000), np.exp(rng.uniform(-80, 80, 10000)),
                         (np.arange(1, 2 ** 10) + 0.5) / 2 ** 23]).astype(np.float32)
    xs = xs[xs > 0]
    got = np.array([L.oracle_ms_logf(float(x)) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64))
    # absolute error near log(1) = 0 is measured against the spacing of 1.0
    ulp = np.maximum(np.spacing(np.abs(want).astype(np.float32)).astype(np.float64), np.spacing(np.float32(1)) / 4)
    assert (np.abs(got - want) / ulp).max() <= 2.0


def test_sincos_box_muller_range():
    L = O.lib()
    a = np.linspace(0.0, 6.283185307179586, 50001).astype(np.float32)
    s = C.c_float()
    c = C.c_float()
    got_s, got_c = [], []
    for x in a:
        L.oracle_ms_sincosf(float(x), C.byref(s), C.byref(c))
        got_s.append(s.value)
        got_c.append(c.value)
    got_s = np.array(got_s)
    got_c = np.array(got_c)
    ad = a.astype(np.float64)
    assert np.abs(got_s - np.sin(ad)).max() < 3e-7
    assert np.abs(got_c - np.cos(ad)).max() < 3e-7

"""Helpers shared by the GPU parity tests: scene set-up and comparisons against the oracle."""
from __future__ import annotations

import numpy as np

from msplat_amd import scene


def small_scene(n=3000, n_hidden=1500, W=160, H=144, degree=3, kind=0, seed=11, fx=None):
    g = scene.synth_gaussians(kind, seed, n, degree)
    hidden = scene.synth_gaussians(kind, seed + 1000, n_hidden, degree)
    c2w = scene.synth_cameras(seed + 7, 4, 3.0, 0.0)
    fx = fx or float(W)
    cams = [scene.prepare_camera(c, fx, fx, W / 2.0, H / 2.0, W, H) for c in c2w]
    return g, hidden, cams


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.abs(b).max(), 1e-30) if b.size else 1.0
    return float(np.abs(a - b).max() / den) if b.size else 0.0


def assert_exact(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{name}: shape {a.shape} vs {b.shape}"
    bad = np.flatnonzero(a.reshape(-1) != b.reshape(-1))
    assert bad.size == 0, f"{name}: {bad.size} of {a.size} differ; first at {bad[:5]}: {a.reshape(-1)[bad[:5]]} vs {b.reshape(-1)[bad[:5]]}"


def assert_rel(name, a, b, tol):
    e = rel_err(a, b)
    assert e <= tol, f"{name}: relative error {e:.3e} > {tol:.1e}"
    return e

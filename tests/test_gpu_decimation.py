"""Seed decimation of fresh initial Gaussians (reference Model::setupOptimizers,
core/src/model.cpp:176-204): more than 75% of cap_max fresh seeds are subsampled to
floor(0.5 cap_max) rows, kept in index order, for growth headroom; at or below 75% every seed
is kept. The reference draws the sample with libc++'s mt19937(42) + std::shuffle, a stream
this build cannot reproduce (implementation-defined); its own sampler (splitmix64
Fisher-Yates, seed 42) is checked for the properties the reference's relies on: the count,
unchanged rows in increasing index order, zero optimiser state, determinism, and a
uniform choice (the kept fraction of every tenth of the index range within 5 binomial
standard deviations of 1/2 cap / n)."""
import numpy as np
import pytest

from msplat_amd import GaussianTrainer, TrainingConfig, scene
from test_gpu_trainer import GROUPS, _dataset, _state

pytestmark = pytest.mark.gpu


def _kept_indices(g, st, n_kept):
    """Index in the initial set of every kept row (rows are told apart by their means)."""
    where = {tuple(r): i for i, r in enumerate(g["means"].astype(np.float32).tolist())}
    means = st["means"].reshape(n_kept, 3)
    return np.array([where[tuple(r)] for r in means.astype(np.float32).tolist()], np.int64)


def _trainer(g, hidden, cap):
    ds, _ = _dataset(g, hidden, 96, 80, n_cams=2)
    return ds, GaussianTrainer(ds, TrainingConfig(iterations=100, cap_max=cap))


@pytest.mark.parametrize("n,cap", [(4000, 5000), (9000, 10000), (20000, 7000)])
def test_decimation_keeps_half_the_cap(n, cap):
    g = scene.synth_gaussians(0, 17, n, 3)
    hidden = scene.synth_gaussians(0, 18, 300, 3)
    ds, tr = _trainer(g, hidden, cap)
    k = int(cap * 0.5)
    assert tr.splat_count == k
    st = _state(tr)
    idx = _kept_indices(g, st, k)
    assert (np.diff(idx) > 0).all()                 # distinct, in index order
    shapes = {"means": 3, "scales": 3, "quats": 4, "f_dc": 3, "f_rest": 45, "opac": 1}
    for key in GROUPS:
        want = np.asarray(g[key], np.float32).reshape(n, shapes[key])[idx].reshape(-1)
        assert np.array_equal(st[key], want), key  # rows copied unchanged
        assert not st["m." + key].any() and not st["v." + key].any()
    # uniform choice: every tenth of the index range keeps about k / n of its rows
    p = k / n
    for d in range(10):
        lo, hi = d * n // 10, (d + 1) * n // 10
        got = int(((idx >= lo) & (idx < hi)).sum())
        sd = np.sqrt((hi - lo) * p * (1 - p))
        assert abs(got - (hi - lo) * p) <= 5 * sd + 1, (d, got, (hi - lo) * p)
    # deterministic: a second trainer on the same seeds keeps the same rows
    tr2 = GaussianTrainer(ds, TrainingConfig(iterations=100, cap_max=cap))
    assert np.array_equal(_state(tr2)["means"], st["means"])


def test_decimation_threshold_is_strict():
    cap = 4000
    hidden = scene.synth_gaussians(0, 18, 300, 3)
    g = scene.synth_gaussians(0, 19, 3000, 3)  # exactly 75% of the cap: all kept
    _, tr = _trainer(g, hidden, cap)
    assert tr.splat_count == 3000
    assert np.array_equal(_state(tr)["means"], g["means"].reshape(-1))
    g = scene.synth_gaussians(0, 19, 3001, 3)  # one more: decimated to 2000
    _, tr = _trainer(g, hidden, cap)
    assert tr.splat_count == 2000

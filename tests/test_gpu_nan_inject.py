"""The reference's NaN-injection self-test (MSPLAT_NAN_INJECT, core/src/model.cpp:638-659):
at the end of a relocation pass 100 draws among the active Gaussians get NaN DC colours.
With the non-finite gates in place (projection zombie gate metal:1813-1833, Adam gates
metal:1697-1705) the poisoned Gaussians render invisible (the image equals the oracle's
render of the state without them), the gate counters record them, training stays finite,
and the next relocation pass culls them as non-finite."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from msplat_amd import GaussianTrainer, TrainingConfig, _lib, scene
from parity_util import assert_exact, small_scene
from test_gpu_trainer import GROUPS, _dataset, _host_state, _state

pytestmark = pytest.mark.gpu


def _nan_counts(tr):
    out = np.zeros(16, np.uint32)
    assert _lib.lib().msplat_engine_read(tr.engine_handle, b"nan_counts", out.ctypes.data, out.nbytes) == 64
    return out


def test_nan_inject_is_contained_and_culled():
    W, H = 96, 80
    n = 3000
    g, hidden, _ = small_scene(n=n, n_hidden=1000, W=W, H=H)
    ds, cams = _dataset(g, hidden, W, H, n_cams=3)
    cap = 4200  # n <= 0.75 cap: no seed decimation
    tr = GaussianTrainer(ds, TrainingConfig(iterations=10000, sh_degree_interval=1, cap_max=cap))
    tr.set_option("nan_inject_step", 500)
    tr.step()
    tr.step()
    info = tr.run_mcmc(500)
    n1 = tr.splat_count
    assert n1 == info["num_active"] > n
    st = _state(tr)
    fdc = st["f_dc"].reshape(n1, 3)
    bad = np.flatnonzero(~np.isfinite(fdc).all(1))
    assert 1 <= bad.size <= 100 and np.isnan(fdc[bad]).all()
    for k in GROUPS:
        if k != "f_dc":
            assert np.isfinite(st[k]).all(), k
    # invisible: the render equals the oracle's render of the state without those rows
    keep = np.setdiff1d(np.arange(n1), bad)
    full = _host_state(st, 3)
    clean = {k: full.view()[k][keep] for k in GROUPS}
    clean["degree"] = 3
    before = _nan_counts(tr)
    deg = min(tr.iteration, 3)  # SH degrees in use at the trainer's iteration (interval 1)
    for ci in range(2):
        cam = cams[ci]["cam"]
        img = tr.render(ci)
        assert_exact(f"render {ci} (poisoned state, oracle)", img,
                     O.render(full, cam, deg, scene.DEFAULT_BG, want_ids=False)["img"])
        assert_exact(f"render {ci} (state without the poisoned rows)", img,
                     O.render(O.HostState(clean), cam, deg, scene.DEFAULT_BG, want_ids=False)["img"])
    after = _nan_counts(tr)
    assert after[7] > before[7] and after[12] > before[12]   # zombie gate: non-finite colour
    # training goes on with finite values everywhere else
    for _ in range(3):
        tr.step()
        assert np.isfinite(tr.last_loss)
    st = _state(tr)
    fdc = st["f_dc"].reshape(n1, 3)
    assert np.array_equal(np.flatnonzero(~np.isfinite(fdc).all(1)), bad)
    for k in GROUPS:
        if k != "f_dc":
            assert np.isfinite(st[k]).all(), k
    # the next pass culls them
    info2 = tr.run_mcmc(600)
    assert info2["nonfinite"] == bad.size
    st = _state(tr)
    for k in GROUPS:
        assert np.isfinite(st[k]).all(), k

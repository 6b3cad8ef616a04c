"""Multi-view batched steps (msplat_trainer_step_batch) against the oracle's batched step
from the same state: per-view losses 1e-5, parameters and Adam moments 1e-4 after one
update of the gradients summed over views of different sizes; a one-view batch equals the
plain step; batched steps recorded as a HIP graph and replayed (new cameras, targets and
step scalars each replay, eager work and synchronisation in between) match the oracle at
every replay, for chunked and monolithic views."""
import numpy as np
import pytest

import oracle_py as O
from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, scene
from parity_util import assert_rel, compare_state, small_scene
from test_gpu_trainer import GROUPS, _host_state, _state

pytestmark = pytest.mark.gpu

SIZES = [(96, 80), (128, 96), (160, 144), (112, 112)]


def _uneven_dataset(g, hidden, n_cams=6, sizes=SIZES):
    c2w = scene.synth_cameras(31, n_cams, 3.0, 0.0)
    cams = []
    for i, c in enumerate(c2w):
        W, H = sizes[i % len(sizes)]
        cam = scene.prepare_camera(c, float(W), float(W), W / 2.0, H / 2.0, W, H)
        img = O.render(O.HostState(hidden), cam, 3, scene.DEFAULT_BG, want_ids=False)["img"]
        cams.append({"c2w": c, "fx": float(W), "fy": float(W), "cx": W / 2.0, "cy": H / 2.0,
                     "image": scene.quantize(img), "cam": cam})
    return Dataset.from_arrays(cams, g), cams


def _batch_hps(pre, cams, idx):
    """The step scalars of a batch: the per-camera scalars taken before the step, with the
    loss normalised by 3 H W B."""
    hps = []
    for ci in idx:
        hp = pre[ci]
        c = cams[ci]["cam"]
        hp.loss_inv_n = float(np.float32(1.0) / np.float32(c.height * c.width * 3 * len(idx)))
        hps.append(hp)
    return hps


def _check_batched_steps(tr, cams, views, steps):
    for step in steps:
        before = _state(tr)
        host = _host_state(before, 3)
        host64 = _host_state(before, 3, dtype=np.float64)
        pre = [tr.step_params(ci, step, step) for ci in range(len(cams))]
        s = tr.step_batch(views)
        assert s.iteration == step
        idx, losses = tr.last_batch(losses=True)
        assert len(idx) == views
        hps = _batch_hps(pre, cams, idx)
        bc, bg = [cams[i]["cam"] for i in idx], [cams[i]["image"] for i in idx]
        loss, _ = O.train_step_batch(host, bc, bg, hps)
        O.train_step_batch(host64, bc, bg, hps)
        assert_rel(f"batch loss@{step}", np.mean(losses), loss, 1e-5)
        compare_state(f"@{step}", _state(tr), host, host64)


@pytest.mark.parametrize("views", [3, 5])
def test_batched_steps_match_oracle(views):
    g, hidden, _ = small_scene(n=2500, n_hidden=1200)
    ds, cams = _uneven_dataset(g, hidden)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * 2500))
    _check_batched_steps(tr, cams, views, (1, 2, 3))


# (sizes, views): one size, so that every step replays the first recording; and mixed
# sizes, monolithic (>= 400 tiles) beside chunked views
GRAPH_CASES = {"one_size": ([(96, 80)], 3), "mixed": ([(320, 320), (96, 80), (128, 96)], 3)}


@pytest.mark.parametrize("case", list(GRAPH_CASES))
def test_graph_batched_steps_match_oracle(case):
    sizes, views = GRAPH_CASES[case]
    g, hidden, _ = small_scene(n=2500, n_hidden=1200)
    ds, cams = _uneven_dataset(g, hidden, n_cams=6, sizes=sizes)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * 2500))
    tr.set_option("batch_graph", 1)
    # SH degrees in use grow 1, 2, 3 over steps 1-3 (a new recording each), then stay
    _check_batched_steps(tr, cams, views, (1, 2, 3, 4, 5))
    info = tr.batch_info()
    assert info["captures"] >= 3 and info["captures"] + info["replays"] == 5
    if case == "one_size":
        assert info["captures"] == 3 and info["replays"] == 2
    # eager work on the same trainer between replays
    tr.render(0)
    _check_batched_steps(tr, cams, views, (6,))


def test_sequential_batch_matches_oracle():
    """The baseline mode: views one after another through the single-view kernels."""
    g, hidden, _ = small_scene(n=2500, n_hidden=1200)
    ds, cams = _uneven_dataset(g, hidden, n_cams=6, sizes=[(320, 320), (96, 80), (128, 96)])
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * 2500))
    tr.set_option("batch_sequential", 1)
    _check_batched_steps(tr, cams, 3, (1, 2, 3))


def test_autotuned_graph_steps_match_oracle():
    """Every autotuner candidate trains identically: stage one's raster choices (pixels per
    lane x persistent residency), then stage two's list choices (histogram rank chunks x
    bucket waves) around the best, all recorded as graphs sized for the larger view (views of
    96x80 and 640x480 alternate in each slot), then stage three, the best so far launched
    eagerly."""
    g, hidden, _ = small_scene(n=2500, n_hidden=1200)
    ds, cams = _uneven_dataset(g, hidden, n_cams=6, sizes=[(640, 480), (96, 80)])
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * 2500))
    tr.set_option("batch_graph", 1)
    tr.set_option("autotune", 1)
    _check_batched_steps(tr, cams, 2, tuple(range(1, 56)))
    info = tr.autotune_info()
    cands = [c[:5] for c in info["timed"]]
    assert len(cands) == 12 and len(set(cands)) == 12
    assert {c[0] for c in cands[:7]} == {1, 4} and all(c[2:] == (0, 0, 1) for c in cands[:7])
    def best(timed):  # a later candidate wins only by more than 1% (float32, as the library)
        b = timed[0]
        for c in timed:
            if np.float32(c[5]) < np.float32(0.99) * np.float32(b[5]):
                b = c
        return b
    assert cands[0][:2] == (1, 8 * (cands[3][1] // 16))  # the default residency first
    best1 = best(info["timed"][:7])[:2]
    assert all(c[:2] == best1 for c in cands[7:11])
    assert {c[2:] for c in cands[7:11]} == {(384, 0, 1), (1536, 0, 1), (0, 2, 1), (384, 2, 1)}
    assert cands[11] == best(info["timed"][:11])[:4] + (0,)
    assert info["chosen"] in cands
    assert all(c[5] > 0 for c in info["timed"])


def test_four_pixels_per_lane_match_oracle():
    """The one-wave-per-tile raster kernels (4 pixels per lane) in plain and batched steps."""
    g, hidden, _ = small_scene(n=2500, n_hidden=1200)
    ds, cams = _uneven_dataset(g, hidden, n_cams=4, sizes=[(320, 320), (336, 320)])
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * 2500))
    tr.set_option("raster_ppl", 4)
    _check_batched_steps(tr, cams, 2, (1, 2))


def test_one_view_batch_equals_plain_step():
    g, hidden, _ = small_scene(n=2000, n_hidden=1000)
    ds, cams = _uneven_dataset(g, hidden, n_cams=4)
    cfg = TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=4000)
    a, b = GaussianTrainer(ds, cfg), GaussianTrainer(ds, cfg)
    for _ in range(3):
        a.step()
        b.step_batch(1)
        assert a.last_camera == b.last_camera
        sa, sb = _state(a), _state(b)
        for key in sa:
            assert_rel(key, sb[key], sa[key], 1e-6)

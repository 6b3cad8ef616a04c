"""GPU parity of the trainer behind the C API (msplat_c_api.h) against the CPU oracle:
per-step state after msplat_trainer_step (parameters and Adam moments 1e-4, loss 1e-5),
the GPU-resident MCMC relocation/growth (sources, copies and counters bit-exact; Eq. 9
rows 1e-6), the reference's Eq. 9 known-answer cases on the GPU kernel, held-out
evaluation metrics, and the render entry points."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, scene
from msplat_amd import _lib
from parity_util import assert_exact, assert_rel, compare_state, small_scene
from test_relocation_fixture import check_eq9

pytestmark = pytest.mark.gpu

GROUPS = ("means", "scales", "quats", "f_dc", "f_rest", "opac")


def _dataset(g, hidden, W, H, n_cams=4, test_every=0, fx=None):
    c2w = scene.synth_cameras(11 + 7, n_cams, 3.0, 0.0)
    fx = fx or float(W)
    cams = []
    for i, c in enumerate(c2w):
        cam = scene.prepare_camera(c, fx, fx, W / 2.0, H / 2.0, W, H)
        img = O.render(O.HostState(hidden), cam, 3, scene.DEFAULT_BG, want_ids=False)["img"]
        cams.append({"c2w": c, "fx": fx, "fy": fx, "cx": W / 2.0, "cy": H / 2.0, "image": scene.quantize(img),
                     "test": bool(test_every) and i % test_every == 0, "cam": cam})
    return Dataset.from_arrays(cams, g), cams


def _state(tr, n=None):
    out = {}
    for gi, k in enumerate(GROUPS):
        for which, pre in ((0, ""), (1, "m."), (2, "v.")):
            out[pre + k] = tr.read_group(gi, which)
    return out


def _host_state(st, degree, cap=None, dtype=np.float32):
    n = st["means"].size // 3
    nr = (degree + 1) ** 2 - 1
    shape = {"means": (n, 3), "scales": (n, 3), "quats": (n, 4), "f_dc": (n, 3), "f_rest": (n, nr, 3), "opac": (n,)}
    g = {k: st[k].reshape(shape[k]) for k in GROUPS}
    g["degree"] = degree
    mom = ({k: st["m." + k].reshape(shape[k]) for k in GROUPS}, {k: st["v." + k].reshape(shape[k]) for k in GROUPS})
    return O.HostState(g, moments=mom, capacity=cap, dtype=dtype)


def test_trainer_steps_match_oracle():
    W, H = 160, 144
    g, hidden, _ = small_scene(W=W, H=H)
    ds, cams = _dataset(g, hidden, W, H)
    train_cams = [c for c in cams if not c["test"]]
    cfg = TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * g["means"].shape[0])
    tr = GaussianTrainer(ds, cfg)
    for step in (1, 2, 3, 4):
        before = _state(tr)
        hps = [tr.step_params(ci, step, step) for ci in range(len(train_cams))]
        s = tr.step()
        assert s.iteration == step and s.splat_count == g["means"].shape[0]
        ci = tr.last_camera
        host = _host_state(before, 3)
        host64 = _host_state(before, 3, dtype=np.float64)
        gt = train_cams[ci]["image"]
        orc = O.train_step(host, train_cams[ci]["cam"], gt, hps[ci], want_ids=False)
        O.train_step(host64, train_cams[ci]["cam"], gt, hps[ci], want_ids=False)
        after = _state(tr)
        assert_rel(f"loss@{step}", tr.last_loss, orc["loss"], 1e-5)
        compare_state(f"@{step}", after, host, host64)
    n_adam, lr = tr.optim_state()
    assert n_adam == 4
    assert abs(lr[0] - scene.means_lr(4)) <= 1e-6 * scene.means_lr(4)


def _mcmc_scene(n=4000, degree=3, seed=5):
    g = scene.synth_gaussians(0, seed, n, degree)
    rng = np.random.default_rng(seed)
    dead = rng.choice(n, n // 5, replace=False)
    g["opac"][dead] = -8.0                       # sigmoid < 0.005
    far = rng.choice(np.setdiff1d(np.arange(n), dead), 30, replace=False)
    g["means"][far] *= 10.0                      # beyond the cull radius 3
    g["f_rest"][far[:3], 2, 1] = np.nan          # non-finite rows are dead too
    g["scales"][dead[:5], 0] = np.inf
    return g


@pytest.mark.parametrize("grow_to_cap", [False, True], ids=["grow_5pct", "grow_to_cap"])
def test_gpu_mcmc_matches_oracle(grow_to_cap):
    """Relocation + growth on the device against the oracle from the same pre-pass state.
    The initial set is 75% of the cap (above that, fresh seeds are decimated first,
    model.cpp:176-204); "grow_to_cap" runs passes until the next growth hits the cap."""
    W, H = 96, 80
    g = _mcmc_scene()
    n = g["means"].shape[0]
    hidden = scene.synth_gaussians(0, 99, 500, 3)
    ds, _ = _dataset(g, hidden, W, H, n_cams=2)
    cap = (4 * n) // 3 + 2  # n <= 0.75 cap: no seed decimation
    tr = GaussianTrainer(ds, TrainingConfig(iterations=10000, cap_max=cap))
    tr.step()  # non-zero moments so that zeroing is observable
    step, count = 500, 0
    if grow_to_cap:
        while int(np.float32(1.05) * np.float32(tr.splat_count)) < cap:
            count = tr.run_mcmc(step)["relocation_count"]
            step += 100
            tr.step()
    n0 = tr.splat_count
    before = _state(tr)
    host = _host_state(before, 3, cap=cap)
    info_o = (C.c_int32 * 6)()
    O.lib().oracle_mcmc_step(host.struct(), cap, 3.0, count, info_o)
    info = tr.run_mcmc(step)
    want = dict(zip(("num_active", "dead", "grown", "nonfinite", "culled", "relocation_count"), list(info_o)))
    assert info == want
    if not grow_to_cap:
        assert info["dead"] > n // 5 and info["culled"] == 30 - 3 and info["nonfinite"] == 3 + 5
    n2 = info["num_active"]
    assert n2 == min(cap, int(np.float32(1.05) * np.float32(n0)))
    after = _state(tr)
    for k in GROUPS:
        got = after[k].reshape(n2, -1)
        want_p = host.p[k][:n2].reshape(n2, -1)
        if k in ("scales", "opac"):
            # Eq. 9 rewrites source rows in double (device pow/log vs host libm)
            assert_rel(k, got, want_p, 1e-6)
        else:
            assert_exact(k, got, want_p)
        assert_exact("m." + k, after["m." + k].reshape(n2, -1), host.m[k][:n2].reshape(n2, -1))
        assert_exact("v." + k, after["v." + k].reshape(n2, -1), host.v[k][:n2].reshape(n2, -1))
    # the relocation counter advanced once per draw performed (relocation if any row died,
    # growth if any row was added)
    assert info["relocation_count"] == count + int(info["dead"] > 0) + int(info["grown"] > 0)
    # outside the schedule nothing happens (model.cpp:498-501)
    assert tr.run_mcmc(step + 50)["num_active"] == n2


def test_gpu_eq9_known_answers():
    L = _lib.lib()

    def eq9(o, n):
        oo = np.array([o], np.float32)
        cnt = np.array([n], np.int32)
        on = np.zeros(1, np.float32)
        sn = np.zeros(1, np.float32)
        assert L.msplat_eq9_device(1, _lib.fptr(oo), cnt.ctypes.data_as(C.POINTER(C.c_int)), _lib.fptr(on),
                                   _lib.fptr(sn)) == 0
        return float(on[0]), float(sn[0])
    check_eq9(eq9)
    # and the GPU kernel equals the oracle's arithmetic case by case
    OL = O.lib()
    for o in (0.05, 0.5, 0.9, 0.998):
        for n in (1, 2, 7, 51, 80):
            on, sn = eq9(o, n)
            so = np.ones(3, np.float32)
            sw = np.zeros(3, np.float32)
            ow = C.c_float()
            OL.oracle_eq9(float(np.float32(o)), so.ctypes.data, n, C.byref(ow), sw.ctypes.data)
            assert abs(on - ow.value) <= 1e-7 * max(1.0, abs(ow.value)), (o, n)
            assert abs(sn - sw[0]) <= 1e-6 * abs(sw[0]), (o, n)


def test_evaluate_and_render_entry_points(tmp_path):
    W, H = 128, 96
    g, hidden, _ = small_scene(n=2000, n_hidden=1000, W=W, H=H)
    ds, cams = _dataset(g, hidden, W, H, n_cams=4, test_every=2)
    assert ds.num_train == 2 and ds.num_test == 2
    tr = GaussianTrainer(ds, TrainingConfig(iterations=3, sh_degree_interval=1))
    tr.train()
    assert tr.iteration == 3
    m = tr.evaluate()
    assert m["num_test"] == 2 and m["num_gaussians"] == g["means"].shape[0]
    test_cams = [c for c in cams if c["test"]]
    ps, ss, ls = 0.0, 0.0, 0.0
    for i, c in enumerate(test_cams):
        img = tr.render(i, use_test=True)
        assert img.shape == (H, W, 3)
        ref = O.eval_metrics(img, c["image"])
        ps += ref["psnr"]
        ss += ref["ssim"]
        ls += ref["l1"]
    assert abs(m["psnr"] - ps / 2) < 1e-4
    assert abs(m["ssim"] - ss / 2) < 1e-6
    assert abs(m["l1"] - ls / 2) < 1e-7
    # render_from_pose with the camera's own pose reproduces render(); RGBA8 truncates it
    pose = ds.camera_pose(1)
    a = tr.render(1)
    b = tr.render_from_pose(pose, 1)
    assert_exact("render_from_pose", b, a)
    rgba = tr.render_from_pose_rgba(pose, 1)
    assert rgba.shape == (H, W, 4) and (rgba[..., 3] == 255).all()
    want = (np.minimum(np.maximum(a, 0.0), 1.0) * np.float32(255.0)).astype(np.uint8)
    assert_exact("rgba", rgba[..., :3], want)
    # out-of-range indices give empty buffers (msplat_api.mm:169-170, 188-189, 211-213)
    assert tr.render(99).size == 0 and tr.render_from_pose(pose, 7).size == 0

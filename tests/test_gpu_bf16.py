"""bfloat16 SH storage (f_dc / f_rest coefficients and their Adam moments in bf16, fp32
masters and arithmetic) against the oracle run in the same storage mode: renders bit-exact,
per-step state within 1e-4 (fp32 arrays) and within one bf16 ulp per element for the bf16
moments (a last-bit fp32 difference can move a value across a bf16 rounding boundary),
MCMC copies of the bf16 rows exact; the checkpoint round trip in bf16 mode is kept off the
GPU with the fp32 checkpoint case (see its skip reason)."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from msplat_amd import GaussianTrainer, TrainingConfig, scene
from parity_util import assert_exact, assert_rel, compare_state, oracle_pair, small_scene
from test_gpu_trainer import GROUPS, _dataset, _host_state, _mcmc_scene, _state

pytestmark = pytest.mark.gpu


def bf16_round(x):
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def assert_bf16_close(name, got, want, tol=1e-4):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = max(np.abs(want).max(), 1e-30)
    ulp = np.abs(want) * 2.0 ** -7
    bad = np.abs(got - want) > tol * scale + ulp
    assert not bad.any(), f"{name}: {bad.sum()} elements beyond one bf16 ulp"
    return float(np.abs(got - want).max() / scale)


def test_bf16_round_matches_device_convention():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.1415927, 1e-20, 65504.0], np.float32)
    r = bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0078125)
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_bf16_train_steps_match_oracle():
    W, H = 160, 144
    g, hidden, _ = small_scene(W=W, H=H)
    ds, cams = _dataset(g, hidden, W, H)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * g["means"].shape[0]))
    tr.set_option("sh_bf16", 1)
    for step in (1, 2, 3, 4):
        before = _state(tr)
        hps = [tr.step_params(ci, step, step) for ci in range(len(cams))]
        tr.step()
        ci = tr.last_camera
        g0 = _host_state(before, 3)
        mom = (g0.m, g0.v)
        gd = dict(g0.view(), degree=3)
        host, orc, host64, _, href, _ = oracle_pair(gd, mom, cams[ci]["cam"], cams[ci]["image"], hps[ci],
                                                    sh_bf16=True, want_ids=False, fp32_ref=True)
        after = _state(tr)
        assert_rel(f"loss@{step}", tr.last_loss, orc["loss"], 1e-5)
        for k in ("f_dc", "f_rest"):
            for pre in ("m.", "v."):
                assert (bf16_round(after[pre + k]) == after[pre + k]).all(), f"{pre}{k} not bf16"
        # per element: fp32 arrays 1e-4, bf16 moments one bf16 step, both plus 4x the FP64
        # error of the all-fp32 oracle from the same state
        compare_state(f"@{step}", after, host, host64, bf16_groups=("f_dc", "f_rest"), href=href)
    # render parity at SH degree 3: the forward reads the bf16 shadows
    st = _host_state(_state(tr), 3)
    st.sh_bf16 = True
    for ci in range(2):
        ref = O.render(st, cams[ci]["cam"], 3, scene.DEFAULT_BG, want_ids=False)["img"]
        assert_exact(f"bf16 render {ci}", tr.render(ci), ref)


def test_bf16_storage_switch():
    W, H = 96, 80
    g, hidden, _ = small_scene(n=1500, n_hidden=600, W=W, H=H)
    ds, _ = _dataset(g, hidden, W, H, n_cams=2)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=100, sh_degree_interval=1, cap_max=4000))
    tr.step()
    tr.step()
    fp = _state(tr)
    tr.set_option("sh_bf16", 1)
    bf = _state(tr)
    for k in GROUPS:
        assert_exact(k, bf[k], fp[k])  # masters untouched
        for pre in ("m.", "v."):
            want = bf16_round(fp[pre + k]) if k in ("f_dc", "f_rest") else fp[pre + k]
            assert_exact(pre + k, bf[pre + k], want)
    tr.set_option("sh_bf16", 0)
    back = _state(tr)
    for k in GROUPS:
        assert_exact(k, back[k], bf[k])


@pytest.mark.skip(reason="runs the save / fresh trainer / load sequence of test_checkpoint_layout_and_round_trip, "
                         "whose first GPU run ended with the box faulted and whose cause is not shown; it stays "
                         "off the GPU with that case (the CPU emulation, tests/cpu_emu, covers it)")
def test_bf16_checkpoint_round_trip(tmp_path):
    W, H = 96, 80
    g, hidden, _ = small_scene(n=1500, n_hidden=600, W=W, H=H)
    ds, _ = _dataset(g, hidden, W, H, n_cams=2)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=100, sh_degree_interval=1, cap_max=4000))
    tr.step()
    tr.step()
    tr.set_option("sh_bf16", 1)
    bf = _state(tr)
    path = str(tmp_path / "b.msplat")
    tr.save_checkpoint(path)
    tr2 = GaussianTrainer(ds, TrainingConfig(iterations=100, sh_degree_interval=1, cap_max=4000))
    tr2.set_option("sh_bf16", 1)
    tr2.load_checkpoint(path)
    st2 = _state(tr2)
    for key in bf:
        assert_exact(key, st2[key], bf[key])
    # the two trainers continue identically (same state, same shadows)
    tr.set_option("sh_bf16", 1)
    a = tr.render(0)
    b = tr2.render(0)
    assert_exact("render after load", b, a)


def test_bf16_mcmc_matches_oracle():
    W, H = 96, 80
    g = _mcmc_scene()
    n = g["means"].shape[0]
    hidden = scene.synth_gaussians(0, 99, 500, 3)
    ds, _ = _dataset(g, hidden, W, H, n_cams=2)
    cap = (4 * n) // 3 + 2  # n <= 0.75 cap: no seed decimation
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, cap_max=cap, sh_degree_interval=1))
    tr.set_option("sh_bf16", 1)
    tr.step()
    host = _host_state(_state(tr), 3, cap=cap)
    info_o = (C.c_int32 * 6)()
    O.lib().oracle_mcmc_step(host.struct(), cap, 3.0, 0, info_o)
    info = tr.run_mcmc(500)
    assert info["num_active"] == info_o[0] and info["dead"] == info_o[1] and info["grown"] == info_o[2]
    n2 = info["num_active"]
    after = _state(tr)
    for k in GROUPS:
        got = after[k].reshape(n2, -1)
        want = host.p[k][:n2].reshape(n2, -1)
        if k in ("scales", "opac"):
            assert_rel(k, got, want, 1e-6)
        else:
            assert_exact(k, got, want)
        assert_exact("m." + k, after["m." + k].reshape(n2, -1), host.m[k][:n2].reshape(n2, -1))
        assert_exact("v." + k, after["v." + k].reshape(n2, -1), host.v[k][:n2].reshape(n2, -1))
    # the copied rows' shadows follow their masters: a render equals the oracle's
    cams_ = ds.camera(0)
    cam = scene.prepare_camera(cams_["c2w"], cams_["fx"], cams_["fy"], cams_["cx"], cams_["cy"], W, H)
    st = _host_state(after, 3)
    st.sh_bf16 = True
    ref = O.render(st, cam, 1, scene.DEFAULT_BG, want_ids=False)["img"]
    assert_exact("render after relocation", tr.render(0), ref)

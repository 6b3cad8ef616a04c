"""GPU parity at the benchmark workloads, one step from an identical state against the FP32
oracle (per element, rtol 1e-4 for gradients and state, absolute floor 4x each output's FP64
error; images, final T and contributor indices bit-exact; the loss 1e-5) with the FP64 error
of every output recorded:

- C0: 50k uniform Gaussians, 256x256 views, fx 256: the depth-chunked raster path with
  K = 110 chunks of 512 (reference msplat_metal.mm:1204-1224);
- C1: 1.5M Gaussians of a 64-component clustered mixture, 1600x1040, fx 1300: the
  heavy-tailed list build (tiles over 2048 entries), monolithic raster;
- C3: 3M clustered Gaussians at 1920x1080 with SH coefficients and moments in bf16;
- C2: one 8-view batch of 2M clustered Gaussians at 640x480, 1024x768, 1600x1040 and
  1920x1080, recorded as a HIP graph;
- C4: the learned relocation draw at 2M Gaussians after two 8-view batches of 1600x1040.

Tile counts, offsets, bins and sorted per-tile id lists are compared bit for bit wherever
the kernel-level engine exposes them (C0, C1)."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, scene
from parity_util import (GROUPS, assert_close, assert_exact, assert_rel, compare_state, oracle_pair)

pytestmark = pytest.mark.gpu

TOL_IMG = 1e-5


def _gpu():
    from msplat_amd.device import DeviceArray, DeviceState, Engine
    return DeviceArray, DeviceState, Engine


def _ground_truth(hidden, cam):
    """The hidden set rendered on the device at SH degree 3, quantised to uint8."""
    _, DeviceState, Engine = _gpu()
    eng = Engine()
    eng.render(DeviceState(hidden), cam, 3, scene.DEFAULT_BG)
    img = eng.raster(cam.width, cam.height)["img"]
    eng.close()
    return scene.quantize(img)


def _engine_step_parity(g, hidden, cams, steps, expect_k=None, min_max_tile=None):
    """Kernel-level engine steps (msplat_engine_train_step, debug buffers on): every step
    starts the FP32 and FP64 oracles from the device state it starts from."""
    DeviceArray, DeviceState, Engine = _gpu()
    N = g["means"].shape[0]
    eng = Engine()
    dst = DeviceState(g)
    for step, ci, deg in steps:
        cam = cams[ci]
        W, H = cam.width, cam.height
        print(f"step {step}: camera {ci} {W}x{H}, N={N}, SH degrees {deg}", flush=True)
        gt = _ground_truth(hidden, cam)
        gt_dev = DeviceArray(gt.shape, np.uint8, gt)
        hp = scene.step_params(step, scene.LR_INIT, W, H, step, 3, deg)
        host = dst.to_host()
        mom = dst.moments_host()
        h32, o32, h64, o64 = oracle_pair(host, mom, cam, gt, hp)
        eng.train_step(dst, cam, gt_dev, hp, debug=True)
        eng.sync()
        tag = f"@{step}"
        # projection: integer outputs and packed records bit for bit
        pj = eng.projection(N)
        assert_exact("radii" + tag, pj["radii"], o32["radii"])
        assert_exact("tiles_hit" + tag, pj["tiles_hit"], o32["tiles_hit"])
        vis = o32["radii"] > 0
        for k in ("xys", "depths", "conics", "colors", "aabb"):
            assert_exact(k + tag, pj[k][vis], o32[k][vis])
        st = eng.stats()
        assert st["num_isect"] == o32["num_isect"] and st["overflow"] == 0
        assert st["k_max"] == o32["k_max"]
        if expect_k is not None:
            assert st["k_max"] == expect_k
        lists = eng.lists(st["num_tiles"], st["num_isect"])
        assert_exact("tile_counts" + tag, lists["tile_counts"], o32["tile_counts"])
        assert_exact("tile_bins" + tag, lists["tile_bins"], o32["tile_bins"])
        assert_exact("ids" + tag, lists["ids"], o32["ids"])
        if min_max_tile is not None:
            assert int(o32["tile_counts"].max()) > min_max_tile
        ras = eng.raster(W, H)
        assert_exact("final_idx" + tag, ras["final_idx"], o32["final_idx"])
        assert_exact("img" + tag, ras["img"], o32["img"])
        assert_exact("final_T" + tag, ras["final_T"], o32["final_T"])
        assert_close("img vs fp64" + tag, ras["img"], o32["img"], o64["img"], rtol=TOL_IMG)
        assert abs(eng.loss() - o32["loss"]) <= TOL_IMG * abs(o32["loss"])
        assert_close("loss" + tag, [eng.loss()], [o32["loss"]], [o64["loss"]], rtol=TOL_IMG)
        gr = eng.grads(N, W, H)
        assert_exact("v_rendered" + tag, gr["v_rendered"], o32["v_rendered"])
        for k in ("v_xy", "v_conic", "v_rgb", "v_mean", "v_scale", "v_quat", "v_opac"):
            assert_close(k + tag, gr[k], o32[k], o64[k])
        got = {}
        after = dst.to_host()
        m_gpu, v_gpu = dst.moments_host()
        for k in GROUPS:
            got[k], got["m." + k], got["v." + k] = after[k], m_gpu[k], v_gpu[k]
        compare_state(tag, got, h32, h64)
        assert list(eng.nan_counts()[:6]) == [0] * 6
    eng.close()


def test_c0_chunked_50k_256():
    g = scene.synth_gaussians(0, 1, 50000, 3)
    hidden = scene.synth_gaussians(0, 2, 20000, 3)
    c2w = scene.synth_cameras(5, 16, 3.0, 0.0)
    cams = [scene.prepare_camera(c, 256.0, 256.0, 128.0, 128.0, 256, 256) for c in c2w]
    # SH degree 0 (C0's 200 steps at the default interval of 1000), then 3, then with the
    # regularisers of step >= 500
    _engine_step_parity(g, hidden, cams, [(1, 0, 0), (2, 5, 3), (600, 11, 3)], expect_k=110)


def test_c1_clustered_1p5m_1600x1040():
    g = scene.synth_gaussians(1, 1, 1500000, 3)
    hidden = scene.synth_gaussians(1, 2, 20000, 3)
    c2w = scene.synth_cameras(5, 64, 3.0, 0.15)
    cams = [scene.prepare_camera(c, 1300.0, 1300.0, 800.0, 520.0, 1600, 1040) for c in c2w]
    _engine_step_parity(g, hidden, cams, [(1, 0, 1), (600, 17, 3)], expect_k=1, min_max_tile=2048)


def _trainer_state(tr):
    out = {}
    for gi, k in enumerate(GROUPS):
        for which, pre in ((0, ""), (1, "m."), (2, "v.")):
            out[pre + k] = tr.read_group(gi, which)
    return out


def _shaped(st, n, degree=3):
    nr = (degree + 1) ** 2 - 1
    shape = {"means": (n, 3), "scales": (n, 3), "quats": (n, 4), "f_dc": (n, 3), "f_rest": (n, nr, 3), "opac": (n,)}
    g = {k: st[k].reshape(shape[k]) for k in GROUPS}
    g["degree"] = degree
    mom = ({k: st["m." + k].reshape(shape[k]) for k in GROUPS}, {k: st["v." + k].reshape(shape[k]) for k in GROUPS})
    return g, mom


def _dataset(g, hidden, sizes, f_at_1600, cam_seed, jitter=0.15):
    c2w = scene.synth_cameras(cam_seed, len(sizes), 3.0, jitter)
    cams = []
    for c, (W, H) in zip(c2w, sizes):
        f = f_at_1600 * W / 1600.0
        cam = scene.prepare_camera(c, f, f, W / 2.0, H / 2.0, W, H)
        cams.append({"c2w": c, "fx": f, "fy": f, "cx": W / 2.0, "cy": H / 2.0, "image": _ground_truth(hidden, cam),
                     "cam": cam})
    return Dataset.from_arrays(cams, g), cams


def test_c3_bf16_3m_1920x1080():
    n = 3000000
    g = scene.synth_gaussians(1, 1, n, 3)
    hidden = scene.synth_gaussians(1, 2, 20000, 3)
    ds, cams = _dataset(g, hidden, [(1920, 1080)] * 2, 1300.0, 7)
    cap = (4 * n) // 3 + 4  # n <= 0.75 cap: no seed decimation
    tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, sh_degree_interval=1, cap_max=cap))
    tr.set_option("sh_bf16", 1)
    tr.step()
    tr.step()
    for step in (3,):  # SH degree 3 in use from step 3 (interval 1)
        before = _trainer_state(tr)
        hps = [tr.step_params(ci, step, step) for ci in range(len(cams))]
        tr.step()
        ci = tr.last_camera
        print(f"step {step}: camera {ci}", flush=True)
        g0, mom = _shaped(before, n)
        h32, o32, h64, o64, hr, orr = oracle_pair(g0, mom, cams[ci]["cam"], cams[ci]["image"], hps[ci],
                                                  sh_bf16=True, want_ids=False, fp32_ref=True)
        tag = f"@{step}"
        assert abs(tr.last_loss - o32["loss"]) <= TOL_IMG * abs(o32["loss"])
        assert_close("loss" + tag, [tr.last_loss], [o32["loss"]], [o64["loss"]], rtol=TOL_IMG, ref32=[orr["loss"]])
        compare_state(tag, _trainer_state(tr), h32, h64, bf16_groups=("f_dc", "f_rest"), href=hr)


def test_c2_graph_batch_2m_four_sizes():
    n = 2000000
    g = scene.synth_gaussians(1, 1, n, 3)
    hidden = scene.synth_gaussians(1, 2, 20000, 3)
    sizes = [(640, 480), (1024, 768), (1600, 1040), (1920, 1080)] * 2
    ds, cams = _dataset(g, hidden, sizes, 1300.0, 9)
    cap = (4 * n) // 3 + 4
    tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, sh_degree_interval=1, cap_max=cap))
    tr.set_option("batch_graph", 1)
    tr.step_batch(8)
    tr.step_batch(8)
    step = 3  # SH degree 3 in use (a new recording at each of degrees 1, 2, 3)
    before = _trainer_state(tr)
    pre = [tr.step_params(ci, step, step) for ci in range(len(cams))]
    tr.step_batch(8)
    idx, losses = tr.last_batch(losses=True)
    assert sorted(idx) == list(range(8))
    assert tr.batch_info()["captures"] == 3
    hps = []
    for ci in idx:
        hp = pre[ci]
        c = cams[ci]["cam"]
        hp.loss_inv_n = float(np.float32(1.0) / np.float32(c.height * c.width * 3 * 8))
        hps.append(hp)
    g0, mom = _shaped(before, n)
    h32 = O.HostState(g0, moments=mom)
    h64 = O.HostState(g0, dtype=np.float64, moments=mom)
    ccams = [cams[i]["cam"] for i in idx]
    gts = [cams[i]["image"] for i in idx]
    loss32, _ = O.train_step_batch(h32, ccams, gts, hps)
    loss64, _ = O.train_step_batch(h64, ccams, gts, hps)
    assert_rel("batch loss", np.mean(losses), loss32, TOL_IMG)
    assert_close("batch loss", [np.mean(losses)], [loss32], [loss64], rtol=TOL_IMG)
    compare_state("@batch", _trainer_state(tr), h32, h64)


def test_c4_learned_draw_2m():
    """Two 8-view batches of 1600x1040 gather the proposal's statistics; the relocation pass
    then draws its sources from the network's logits and places copies by its offsets."""
    n = 2000000
    g = scene.synth_gaussians(1, 1, n, 3)
    rng = np.random.default_rng(4)
    dead = rng.choice(n, n // 5, replace=False)
    g["opac"][dead] = -8.0
    hidden = scene.synth_gaussians(1, 2, 20000, 3)
    ds, cams = _dataset(g, hidden, [(1600, 1040)] * 8, 1300.0, 11)
    cap = (4 * n) // 3 + 4
    tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, cap_max=cap))
    tr.set_option("learned_relocation", 1)
    th = O.mlp_init(3)
    th[5248:5504] *= 4.0
    tr.proposal_params(th)
    tr.set_option("relocation_mode", 2)
    tr.step_batch(8)
    tr.step_batch(8)
    before = _trainer_state(tr)
    g0, mom = _shaped(before, n)
    host = O.HostState(g0, moments=mom, capacity=cap)
    info = tr.run_mcmc(500)
    feats = tr.proposal_read("features", np.uint16).reshape(-1, 16)[:n]
    out = np.ascontiguousarray(tr.proposal_read("out").reshape(-1, 4)[:n])
    ref_out = O.mlp_forward(th, feats)
    # bf16 operands rounded where the oracle rounds them; fp32 sums in another order can move a
    # hidden activation across a bf16 rounding boundary, and a flip in the first hidden layer
    # can flip others in the second (the output weights here are scaled by 4): all but 1e-4 of
    # the outputs within one bf16 step (2^-8) of (1 + |ref|), every output within 2^-4 of it
    d = np.abs(out.astype(np.float64) - ref_out)
    scale = 1.0 + np.abs(ref_out)
    print(f"proposal outputs: max |d| {d.max():.3e}, beyond 2^-8 (1 + |ref|): {int((d > scale / 256).sum())} "
          f"of {d.size}")
    assert np.mean(d > scale / 256) <= 1e-4
    assert (d <= scale / 16).all()
    info_o = (C.c_int32 * 6)()
    O.lib().oracle_mcmc_step_learned(host.struct(), cap, 3.0, 0, out.ctypes.data, info_o)
    want = dict(zip(("num_active", "dead", "grown", "nonfinite", "culled", "relocation_count"), list(info_o)))
    assert info == want and info["dead"] >= n // 5
    n2 = info["num_active"]
    after = _trainer_state(tr)
    for k in GROUPS:
        got_p = after[k].reshape(n2, -1)
        want_p = host.p[k][:n2].reshape(n2, -1)
        if k in ("scales", "opac"):
            assert_close(k, got_p, want_p, rtol=1e-6)   # Eq. 9 in double on both sides
        elif k == "means":
            assert_close(k, got_p, want_p, rtol=1e-5)   # copies offset by exp(Eq. 9 log-scale)
        else:
            assert_exact(k, got_p, want_p)
        assert_exact("m." + k, after["m." + k].reshape(n2, -1), host.m[k][:n2].reshape(n2, -1))
        assert_exact("v." + k, after["v." + k].reshape(n2, -1), host.v[k][:n2].reshape(n2, -1))

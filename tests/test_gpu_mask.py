"""Masked training (the mask / has_mask inputs of msplat_train_step, reference
msplat_metal.metal:3602-3653 and 3815-3834; loss normalised by the unmasked pixels,
model.cpp:929-934) on the GPU against the CPU oracle from the same state: the images and
integer outputs bit for bit, the loss-gradient image bit for bit, gradients and the
optimiser state per element (1e-4 relative plus 4x the output's FP64 error), through the
engine entry point, the trainer's dataset masks and the multi-view batched step."""
import numpy as np
import pytest

import oracle_py as O
from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, scene
from parity_util import assert_close, assert_exact, assert_rel, compare_state, oracle_pair, small_scene
from test_gpu_batch import _batch_hps, _uneven_dataset
from test_gpu_parity import _compare_forward, _gpu, _render_gt
from test_gpu_trainer import _dataset, _host_state, _state

pytestmark = pytest.mark.gpu


def _mask(W, H, seed):
    """A masked block, scattered masked pixels, values either side of 128 and a masked
    border strip (the SSIM halo reaches past it)."""
    rng = np.random.default_rng(seed)
    m = np.where(rng.random((H, W)) < 0.15, rng.integers(0, 128, (H, W)), rng.integers(128, 256, (H, W)))
    m[H // 4: H // 2, W // 3: (2 * W) // 3] = 127
    m[:, W - 3:] = 0
    m[0, 0], m[0, 1] = 128, 127
    return m.astype(np.uint8)


@pytest.mark.parametrize("W,H", [(170, 134), (336, 336)], ids=["ragged_chunked", "monolithic"])
def test_masked_train_step_parity(W, H):
    DeviceState, Engine = _gpu()
    from msplat_amd.device import DeviceArray
    g, hidden, cams = small_scene(W=W, H=H)
    N = g["means"].shape[0]
    eng = Engine()
    dst = DeviceState(g)
    for step, ci in ((1, 1), (2, 2), (600, 3)):
        cam = cams[ci]
        _, img = _render_gt(Engine, DeviceState, hidden, cam)
        gt = scene.quantize(img)
        gt_dev = DeviceArray(gt.shape, np.uint8, gt)
        m = _mask(W, H, step)
        m_dev = DeviceArray(m.shape, np.uint8, m)
        hp = scene.step_params(step, scene.LR_INIT, W, H, step, 3, 3)
        hp.loss_inv_n = float(np.float32(1.0) / np.float32((m >= 128).sum() * 3))
        host = dst.to_host()
        mom = dst.moments_host()
        ost, orc, o64s, o64 = oracle_pair(host, mom, cam, gt, hp, mask=m)
        eng.train_step(dst, cam, gt_dev, hp, debug=True, mask=m_dev)
        eng.sync()
        _compare_forward(eng, orc, N, cam)
        assert_rel(f"loss@{step}", eng.loss(), orc["loss"], 1e-5)
        gr = eng.grads(N, W, H)
        assert_exact("v_rendered", gr["v_rendered"], orc["v_rendered"])
        # 6 pixels or more inside the masked block every 11x11 window that holds a pixel is
        # centred on a masked pixel and its own L1 term is dropped: no gradient at all
        assert (gr["v_rendered"][H // 4 + 6: H // 2 - 6, W // 3 + 6: (2 * W) // 3 - 6] == 0).all()
        for k in ("v_xy", "v_conic", "v_rgb", "v_mean", "v_scale", "v_quat", "v_opac"):
            assert_close(f"{k}@{step} masked", gr[k], orc[k], o64[k])
        after = dst.to_host()
        m_gpu, v_gpu = dst.moments_host()
        got = {}
        for k in O.GROUP_SHAPES:
            got[k], got["m." + k], got["v." + k] = after[k], m_gpu[k], v_gpu[k]
        compare_state(f"@{step} masked", got, ost, o64s)


def test_all_white_mask_is_the_unmasked_step():
    """A mask that keeps every pixel multiplies by one: the same bits as the unmasked step."""
    DeviceState, Engine = _gpu()
    from msplat_amd.device import DeviceArray
    W, H = 170, 134
    g, hidden, cams = small_scene(W=W, H=H)
    N = g["means"].shape[0]
    _, img = _render_gt(Engine, DeviceState, hidden, cams[1])
    gt = scene.quantize(img)
    gt_dev = DeviceArray(gt.shape, np.uint8, gt)
    ones = DeviceArray((H, W), np.uint8, np.full((H, W), 255, np.uint8))
    hp = scene.step_params(1, scene.LR_INIT, W, H, 1, 3, 3)
    outs = []
    for mask in (None, ones):
        eng, dst = Engine(), DeviceState(g)
        eng.train_step(dst, cams[1], gt_dev, hp, debug=True, mask=mask)
        eng.sync()
        outs.append((eng.loss(), eng.grads(N, W, H)["v_rendered"], dst.to_host()))
    assert outs[0][0] == outs[1][0]
    assert_exact("v_rendered", outs[1][1], outs[0][1])
    # the updated state is equal up to the order of the raster backward's float atomics,
    # which differs between any two runs
    for k in O.GROUP_SHAPES:
        a, b = outs[1][2][k], outs[0][2][k]
        assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b) + 1e-9), k


def test_trainer_dataset_masks():
    """Masks attached to dataset cameras: binarised at 128, all-white dropped, the step's
    loss normalised by the kept pixels (trainer state against the oracle), and a mask that
    keeps nothing trains as unmasked (reference: has_mask needs unmasked pixels > 0)."""
    W, H = 160, 144
    g, hidden, _ = small_scene(W=W, H=H)
    ds, cams = _dataset(g, hidden, W, H)
    masks = [_mask(W, H, 10 + i) for i in range(len(cams))]
    masks[2] = np.full((H, W), 200, np.uint8)   # all kept: dropped
    masks[3] = np.full((H, W), 40, np.uint8)    # nothing kept: trains unmasked
    kept = [ds.set_mask(i, masks[i]) for i in range(len(cams))]
    assert kept[2] == W * H and kept[3] == 0
    assert kept[0] == int((masks[0] >= 128).sum())
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * g["means"].shape[0]))
    seen = set()
    for step in range(1, 9):
        before = _state(tr)
        hps = [tr.step_params(ci, step, step) for ci in range(len(cams))]
        tr.step()
        ci = tr.last_camera
        seen.add(ci)
        use = masks[ci] if 0 < kept[ci] < W * H else None
        want_inv = np.float32(1.0) / np.float32((kept[ci] if use is not None else W * H) * 3)
        assert hps[ci].loss_inv_n == float(want_inv)
        host = _host_state(before, 3)
        host64 = _host_state(before, 3, dtype=np.float64)
        orc = O.train_step(host, cams[ci]["cam"], cams[ci]["image"], hps[ci], want_ids=False, mask=use)
        O.train_step(host64, cams[ci]["cam"], cams[ci]["image"], hps[ci], want_ids=False, mask=use)
        assert_rel(f"loss@{step}", tr.last_loss, orc["loss"], 1e-5)
        compare_state(f"@{step} cam {ci}", _state(tr), host, host64)
    assert len(seen) >= 3


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_masked_batched_steps(graph):
    """Batched steps over masked and unmasked views of different sizes (per-view loss
    normalised by 3 U_v B), eager and recorded as a HIP graph and replayed."""
    g, hidden, _ = small_scene(n=2500, n_hidden=1200)
    ds, cams = _uneven_dataset(g, hidden)
    masks = {}
    for i, c in enumerate(cams):
        if i % 3 != 2:
            m = _mask(c["cam"].width, c["cam"].height, 40 + i)
            assert ds.set_mask(i, m) == int((m >= 128).sum())
            masks[i] = m
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, sh_degree_interval=1, cap_max=2 * 2500))
    if graph:
        tr.set_option("batch_graph", 1)
    for step in (1, 2, 3, 4, 5):   # SH degrees grow over steps 1-3; 4 and 5 replay a recording
        before = _state(tr)
        host = _host_state(before, 3)
        host64 = _host_state(before, 3, dtype=np.float64)
        pre = [tr.step_params(ci, step, step) for ci in range(len(cams))]
        tr.step_batch(3)
        idx, losses = tr.last_batch(losses=True)
        hps = _batch_hps(pre, cams, idx)
        for k, ci in enumerate(idx):
            if ci in masks:
                hps[k].loss_inv_n = float(np.float32(1.0) / np.float32(int((masks[ci] >= 128).sum()) * 3 * len(idx)))
        bc, bg = [cams[i]["cam"] for i in idx], [cams[i]["image"] for i in idx]
        bm = [masks.get(i) for i in idx]
        loss, _ = O.train_step_batch(host, bc, bg, hps, masks=bm)
        O.train_step_batch(host64, bc, bg, hps, masks=bm)
        assert_rel(f"batch loss@{step}", np.mean(losses), loss, 1e-5)
        compare_state(f"@{step} masked batch", _state(tr), host, host64)
    if graph:
        info = tr.batch_info()
        assert info["replays"] >= 1

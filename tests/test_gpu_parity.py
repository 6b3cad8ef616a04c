"""GPU parity: one training step (and one render) of the HIP path against the CPU oracle,
from an identical state. Integer outputs (tile rectangles, counts, offsets, sorted per-tile
id lists, contributor indices) must match bit for bit; images and the loss within 1e-5
relative; gradients and the optimiser state per element within 1e-4 relative plus an
absolute floor of 4x the output's FP64 error (float atomics make the GPU summation order
differ from the oracle's; tests/parity_util.py).
"""
import numpy as np
import pytest

from msplat_amd import scene
import oracle_py as O
from parity_util import assert_close, assert_exact, assert_rel, compare_state, oracle_pair, small_scene

pytestmark = pytest.mark.gpu

TOL_IMG = 1e-5
TOL_GRAD = 1e-4


def _gpu():
    from msplat_amd.device import DeviceState, Engine
    return DeviceState, Engine


def _render_gt(Engine, DeviceState, hidden, cam):
    eng = Engine()
    dst = DeviceState(hidden)
    eng.render(dst, cam, 3, scene.DEFAULT_BG)
    img = eng.raster(cam.width, cam.height)["img"]
    return eng, img


def _compare_forward(eng, orc, N, cam, check_bits=True):
    pj = eng.projection(N)
    assert_exact("radii", pj["radii"], orc["radii"])
    assert_exact("tiles_hit", pj["tiles_hit"], orc["tiles_hit"])
    vis = orc["radii"] > 0
    for k in ("xys", "depths", "conics", "colors", "aabb"):
        assert_exact(k, pj[k][vis], orc[k][vis])
    st = eng.stats()
    assert st["num_isect"] == orc["num_isect"]
    assert st["k_max"] == orc["k_max"]
    lists = eng.lists(st["num_tiles"], st["num_isect"])
    assert_exact("tile_counts", lists["tile_counts"], orc["tile_counts"])
    assert_exact("tile_bins", lists["tile_bins"], orc["tile_bins"])
    assert_exact("ids", lists["ids"], orc["ids"])
    ras = eng.raster(cam.width, cam.height)
    assert_exact("final_idx", ras["final_idx"], orc["final_idx"])
    assert_rel("img", ras["img"], orc["img"], TOL_IMG)
    assert_rel("final_T", ras["final_T"], orc["final_T"], TOL_IMG)
    if check_bits:
        assert_exact("img bits", ras["img"], orc["img"])
        assert_exact("final_T bits", ras["final_T"], orc["final_T"])


@pytest.mark.parametrize("W,H", [(160, 144), (336, 336)], ids=["chunked_K", "monolithic"])
def test_render_parity(W, H):
    DeviceState, Engine = _gpu()
    g, hidden, cams = small_scene(W=W, H=H)
    eng, img = _render_gt(Engine, DeviceState, hidden, cams[0])
    orc = O.render(O.HostState(hidden), cams[0], 3, scene.DEFAULT_BG)
    _compare_forward(eng, orc, hidden["means"].shape[0], cams[0])


def _cluster_scene(n, spread_px, W, H, seed, n_big=0):
    """n small Gaussians at depths 2.4-2.6 in front of camera 0 of small_scene, within
    spread_px pixels of the image centre: most of a depth chunk's Gaussians land in the same
    few tiles, so the per-chunk bucket segments are long."""
    g, _, cams = small_scene(n=n, n_hidden=16, W=W, H=H)
    c2w = scene.synth_cameras(11 + 7, 4, 3.0, 0.0)[0]
    R = c2w[:3, :3] * np.array([1.0, -1.0, -1.0], np.float32)  # camera axes: x right, y down, z forward
    rng = np.random.default_rng(seed)
    z = rng.uniform(2.4, 2.6, n)
    u = rng.uniform(-spread_px, spread_px, (n, 2))
    pc = np.stack([u[:, 0] * z / W, u[:, 1] * z / W, z], 1)  # fx = W
    g["means"] = (pc @ R.T + c2w[:3, 3]).astype(np.float32)
    g["scales"][:] = np.float32(-5.5)
    g["scales"][:n_big] = np.float32(-1.7)  # ~100 tiles each (the list capacity is 48 per Gaussian)
    return g, cams[0]


@pytest.fixture
def bucket_waves(request):
    """Forces the bucket kernel's waves per workgroup for one test (0: by LDS)."""
    from msplat_amd import _lib
    L = _lib.lib()
    assert L.msplat_set_list_bucket_waves(request.param) == 0
    yield request.param
    L.msplat_set_list_bucket_waves(0)


# tiles holding 1500 and more entries, 150+ and 100+ (chunk segments, one tile's entries from
# one depth chunk of 125 ranks, of up to 125 entries), and Gaussians covering more than 64
# tiles each, written in rank order by 1, 2 or 4 waves per chunk
@pytest.mark.parametrize("bucket_waves", [1, 2, 4], indirect=True, ids=["1wave", "2waves", "4waves"])
@pytest.mark.parametrize("n,spread,n_big,min_count",
                         [(5000, 6.0, 0, 1500), (4000, 40.0, 0, 150), (4000, 60.0, 0, 100),
                          (4000, 60.0, 600, 600)],
                         ids=["one_tile", "few_tiles", "many_tiles", "rects_over_64_tiles"])
def test_render_parity_clustered(n, spread, n_big, min_count, bucket_waves):
    DeviceState, Engine = _gpu()
    W = H = 336
    g, cam = _cluster_scene(n, spread, W, H, seed=3, n_big=n_big)
    eng = Engine()
    eng.render(DeviceState(g), cam, 3, scene.DEFAULT_BG)
    orc = O.render(O.HostState(g), cam, 3, scene.DEFAULT_BG)
    assert orc["tile_counts"].max() >= min_count
    _compare_forward(eng, orc, n, cam)


@pytest.mark.parametrize("bucket_waves", [1, 2, 4], indirect=True, ids=["1wave", "2waves", "4waves"])
def test_list_overflow_is_contained(bucket_waves):
    """More tile hits than the list capacity (48 per Gaussian, msplat_metal.mm:518): tile
    counts and clamped bins stay exact, the tiles whose lists fit hold the oracle's lists, the
    tile that crosses the capacity holds only the empty record (id N), tiles past it are empty,
    and every tile that lost entries renders as background."""
    DeviceState, Engine = _gpu()
    W = H = 336
    n = 64
    g, cam = _cluster_scene(n, 40.0, W, H, seed=5, n_big=n)  # 64 Gaussians of ~100 tiles each
    eng = Engine()
    eng.render(DeviceState(g), cam, 3, scene.DEFAULT_BG)
    orc = O.render(O.HostState(g), cam, 3, scene.DEFAULT_BG, ids_cap=1000 * n)  # the oracle's full lists
    st = eng.stats()
    cap = 48 * n
    assert st["num_isect"] == orc["num_isect"] > cap and st["overflow"] == 1
    T = st["num_tiles"]
    counts = orc["tile_counts"].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert_exact("tile_counts", eng.read("tile_counts", np.int32, (T,)), orc["tile_counts"])
    want_bins = np.stack([np.minimum(off, cap), np.minimum(off + counts, cap)], 1).astype(np.int32)
    assert_exact("tile_bins", eng.read("tile_bins", np.int32, (T, 2)), want_bins)
    ids = eng.read("ids", np.int32, (cap,))
    lost = []
    for t in range(T):
        a, c = int(off[t]), int(counts[t])
        if c == 0:
            continue
        if a + c <= cap:
            assert_exact(f"ids of tile {t}", ids[a:a + c], orc["ids"][a:a + c])
        else:
            lost.append(t)
            if a < cap:
                assert (ids[a:cap] == n).all(), f"tile {t}: entries below the capacity are not the empty record"
    assert len(lost) > 1
    img = eng.raster(W, H)["img"]
    tbx = (W + 15) // 16
    bg = np.array(scene.DEFAULT_BG, np.float32)
    for t in lost:
        ty, tx = divmod(t, tbx)
        tile = img[16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16]
        assert (tile == bg).all(), f"tile {t} does not show the background"


@pytest.mark.parametrize("W,H", [(160, 144), (336, 336)], ids=["chunked_K", "monolithic"])
def test_train_step_parity(W, H):
    DeviceState, Engine = _gpu()
    from msplat_amd.device import DeviceArray
    g, hidden, cams = small_scene(W=W, H=H)
    N = g["means"].shape[0]
    eng = Engine()
    dst = DeviceState(g)
    for step, ci in ((1, 1), (2, 2), (600, 3)):
        cam = cams[ci]
        geng, img = _render_gt(Engine, DeviceState, hidden, cam)
        gt = scene.quantize(img)
        gt_dev = DeviceArray(gt.shape, np.uint8, gt)
        hp = scene.step_params(step, scene.LR_INIT, W, H, step, 3, 3)
        host = dst.to_host()
        mom = dst.moments_host()
        ost, orc, o64s, o64 = oracle_pair(host, mom, cam, gt, hp)
        eng.train_step(dst, cam, gt_dev, hp, debug=True)
        eng.sync()
        _compare_forward(eng, orc, N, cam)
        assert abs(eng.loss() - orc["loss"]) <= TOL_IMG * abs(orc["loss"])
        gr = eng.grads(N, W, H)
        assert_exact("v_rendered", gr["v_rendered"], orc["v_rendered"])
        for k in ("v_xy", "v_conic", "v_rgb", "v_mean", "v_scale", "v_quat", "v_opac"):
            assert_close(f"{k}@{step}", gr[k], orc[k], o64[k], rtol=TOL_GRAD)
        after = dst.to_host()
        m_gpu, v_gpu = dst.moments_host()
        got = {}
        for k in O.GROUP_SHAPES:
            got[k], got["m." + k], got["v." + k] = after[k], m_gpu[k], v_gpu[k]
        compare_state(f"@{step}", got, ost, o64s, rtol=TOL_GRAD)
        assert list(eng.nan_counts()[:6]) == [0] * 6

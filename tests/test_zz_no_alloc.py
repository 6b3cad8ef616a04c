"""msplat_render_noalloc (the allocation-free render, scratch in a static device array)
against msplat_engine_render through the C ABI: images, final transmittance and contributor
indices bit for bit, on two calls in a row whose second is smaller (stale scratch contents
from the first call must not leak into the second), for the monolithic and the depth-chunked
raster paths; a render too large for the array is rejected."""
import ctypes as C

import numpy as np
import pytest

from msplat_amd import _lib, scene
from parity_util import assert_exact

pytestmark = pytest.mark.gpu


def _engine_render(g, cam):
    from msplat_amd.device import DeviceState, Engine
    eng = Engine()
    st = DeviceState(g)
    eng.render(st, cam, 3, scene.DEFAULT_BG)
    out = eng.raster(cam.width, cam.height)
    eng.close()
    return out


def _noalloc_render(g, cam, deg=3):
    from msplat_amd.device import DeviceArray, DeviceState
    L = _lib.lib()
    st = DeviceState(g)
    W, H = cam.width, cam.height
    img = DeviceArray((H, W, 3), np.float32)
    fT = DeviceArray((H, W), np.float32)
    fi = DeviceArray((H, W), np.int32)
    bg = (C.c_float * 3)(*scene.DEFAULT_BG)
    rc = L.msplat_render_noalloc(st.struct(), cam, deg, bg, C.c_void_p(img.ptr), C.c_void_p(fT.ptr),
                                 C.c_void_p(fi.ptr), None)
    assert rc == 0, rc
    L.msplat_sync()
    return {"img": img.download(), "final_T": fT.download(), "final_idx": fi.download()}


def _scene(n, W, H, seed):
    g = scene.synth_gaussians(0, seed, n, 3)
    c2w = scene.synth_cameras(seed + 7, 2, 3.0, 0.0)[0]
    cam = scene.prepare_camera(c2w, float(W), float(W), W / 2.0, H / 2.0, W, H)
    return g, cam


@pytest.mark.parametrize("sizes", [((6000, 400, 352), (2500, 336, 320)),     # monolithic (>= 400 tiles)
                                   ((4000, 160, 144), (1500, 96, 80))],       # depth-chunked (K > 1)
                         ids=["monolithic", "chunked"])
def test_noalloc_render_matches_engine_render(sizes):
    for k, (n, W, H) in enumerate(sizes):
        print(f"call {k}: {n} Gaussians, {W}x{H}", flush=True)
        g, cam = _scene(n, W, H, 11 + k)
        want = _engine_render(g, cam)
        got = _noalloc_render(g, cam)
        for key in ("img", "final_T", "final_idx"):
            assert_exact(f"{key} (call {k})", got[key], want[key])


def test_noalloc_render_rejects_oversize():
    from msplat_amd.device import DeviceArray, DeviceState
    L = _lib.lib()
    g, cam = _scene(400000, 1920, 1080, 5)
    st = DeviceState(g)
    img = DeviceArray((1080, 1920, 3), np.float32)
    fT = DeviceArray((1080, 1920), np.float32)
    fi = DeviceArray((1080, 1920), np.int32)
    bg = (C.c_float * 3)(*scene.DEFAULT_BG)
    rc = L.msplat_render_noalloc(st.struct(), cam, 3, bg, C.c_void_p(img.ptr), C.c_void_p(fT.ptr),
                                 C.c_void_p(fi.ptr), None)
    assert rc == -3
    L.msplat_sync()
    # and a render that fits still works after the rejection
    g2, cam2 = _scene(2000, 336, 320, 6)
    assert_exact("img after a rejection", _noalloc_render(g2, cam2)["img"], _engine_render(g2, cam2)["img"])

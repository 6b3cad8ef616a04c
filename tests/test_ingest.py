"""Dataset ingestion on the host (include/msplat_io.h): image decoding against PIL, the area
resize against an independent integral formulation, undistortion against the Brown-Conrady
forward model, the COLMAP / Nerfstudio / Polycam loaders against poses built with scipy,
the kNN scales against brute force, and the seeding against numpy's MT19937 stream. The
reference's loaders (core/src/loaders/*, core/src/input_data.cpp) are followed; CoreGraphics'
decoders are not available here, so JPEG parity is to within a stated decoder tolerance."""
import ctypes as C
import io
import os
import struct

import numpy as np
import pytest
from PIL import Image
from scipy.spatial.transform import Rotation

from ingest_fixtures import write_colmap, write_json, write_ply, write_png
from msplat_amd import _lib


def L():
    return _lib.lib()


def _p(a):
    return a.ctypes.data


def decode(path):
    w, h = C.c_int(), C.c_int()
    assert L().msplat_image_decode(str(path).encode(), C.byref(w), C.byref(h), None) == 0
    out = np.zeros((h.value, w.value, 3), np.uint8)
    assert L().msplat_image_decode(str(path).encode(), C.byref(w), C.byref(h), _p(out)) == 0
    return out


# ------------------------------------------------------------------------------------ PNG
def _over_black(rgb, a):
    return ((rgb.astype(np.uint32) * a[..., None] + 127) // 255).astype(np.uint8)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "1", "I;16"])
def test_png_decode_matches_pil(tmp_path, mode):
    rng = np.random.default_rng(1)
    h, w = 37, 53
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    if mode == "RGB":
        im = Image.fromarray(rgba[..., :3], "RGB")
        want = rgba[..., :3]
    elif mode == "RGBA":
        im = Image.fromarray(rgba, "RGBA")
        want = _over_black(rgba[..., :3], rgba[..., 3])
    elif mode == "L":
        im = Image.fromarray(rgba[..., 0], "L")
        want = np.repeat(rgba[..., :1], 3, 2)
    elif mode == "LA":
        im = Image.fromarray(rgba[..., :2], "LA")
        want = _over_black(np.repeat(rgba[..., :1], 3, 2), rgba[..., 1])
    elif mode == "P":
        im = Image.fromarray(rgba[..., :3], "RGB").quantize(37)
        pal = np.array(im.getpalette()[:37 * 3], np.uint8).reshape(-1, 3)
        want = pal[np.asarray(im)]
    elif mode == "1":
        bits = rgba[..., 0] > 127
        im = Image.fromarray(bits.astype(np.uint8) * 255, "L").convert("1")
        want = np.repeat((bits.astype(np.uint8) * 255)[..., None], 3, 2)
    else:
        v = (rgba[..., 0].astype(np.uint16) << 8) | rgba[..., 1]
        im = Image.fromarray(v.astype(np.uint16))
        want = np.repeat(((v.astype(np.uint32) * 255 + 32767) // 65535).astype(np.uint8)[..., None], 3, 2)
    path = tmp_path / "a.png"
    im.save(path)
    assert np.array_equal(decode(path), want)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,depth", [(2, 8), (6, 8), (0, 1), (0, 2), (0, 4), (0, 16), (3, 2), (3, 4), (2, 16), (4, 8)])
def test_png_filters_depths_and_interlace(tmp_path, ctype, depth, interlace):
    """Every scanline filter (rows cycle through 0-4), bit depths below 8 packed MSB first,
    16-bit samples, Adam7 passes, a palette with tRNS alpha and a grey transparency key."""
    rng = np.random.default_rng(ctype * 100 + depth)
    h, w = 19, 23
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    s = rng.integers(0, top + 1, (h, w, ch))
    palette = trns = None
    if ctype == 3:
        palette = rng.integers(0, 256, (top + 1, 3))
        trns = bytes(rng.integers(0, 256, top + 1).astype(np.uint8))
    if ctype == 0 and depth == 4:
        trns = struct.pack(">H", int(s[0, 0, 0]))  # grey key: these pixels become black
    path = tmp_path / "f.png"
    write_png(path, s, ctype, depth, palette=palette, trns=trns, interlace=interlace)
    got = decode(path)

    def to8(v):
        v = v.astype(np.uint64)
        if depth == 16:
            return ((v * 255 + 32767) // 65535).astype(np.uint8)
        return (v * 255 // top).astype(np.uint8)
    if ctype == 3:
        rgb = np.asarray(palette, np.uint8)[s[..., 0]]
        a = np.frombuffer(trns, np.uint8)[s[..., 0]]
        want = _over_black(rgb, a)
    elif ctype == 0:
        want = np.repeat(to8(s)[..., :1], 3, 2)
        if trns is not None:
            want[s[..., 0] == s[0, 0, 0]] = 0
    elif ctype == 4:
        want = _over_black(np.repeat(to8(s[..., :1]), 3, 2), to8(s[..., 1]))
    elif ctype == 6:
        want = _over_black(to8(s[..., :3]), to8(s[..., 3]))
    else:
        want = to8(s)
    assert np.array_equal(got, want)


def test_corrupt_and_unknown_images_fail(tmp_path):
    p = tmp_path / "x.png"
    write_png(p, np.zeros((8, 8, 3), np.uint8), 2, 8)
    data = bytearray(open(p, "rb").read())
    data[60:70] = b"\xff" * 10  # inside the deflate stream
    open(p, "wb").write(bytes(data))
    w, h = C.c_int(), C.c_int()
    assert L().msplat_image_decode(str(p).encode(), C.byref(w), C.byref(h), None) == -1
    q = tmp_path / "y.bmp"
    Image.fromarray(np.zeros((4, 4, 3), np.uint8)).save(q)
    assert L().msplat_image_decode(str(q).encode(), C.byref(w), C.byref(h), None) == -1
    assert L().msplat_image_decode(str(tmp_path / "missing.png").encode(), C.byref(w), C.byref(h), None) == -1


# ----------------------------------------------------------------------------------- JPEG
def _photo(h, w, seed):
    """A smooth image with edges (blocks of colour + gradients + noise), like a photograph."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([128 + 100 * np.sin(x / 9.0 + c) * np.cos(y / 13.0 - c) for c in range(3)], -1)
    img[h // 3: h // 2, w // 4: w // 2] = [220, 30, 40]
    img += rng.normal(0, 6, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("sub,restart", [(0, 0), (1, 0), (2, 0), (2, 3), ("gray", 0)],
                         ids=["444", "422", "420", "420_restart", "grey"])
def test_jpeg_decode_within_tolerance_of_libjpeg(tmp_path, sub, restart):
    """Baseline JPEG against PIL's libjpeg-turbo decode (accurate integer IDCT, fancy
    upsampling): the decoders differ in IDCT rounding and chroma interpolation only, so
    every sample is held within 3 levels (8 for 4:2:0 chroma edges) and the mean within
    0.6."""
    img = _photo(77, 91, 3)
    path = tmp_path / "a.jpg"
    kw = {"quality": 92}
    if restart:
        kw["restart_marker_blocks"] = restart
    if sub == "gray":
        Image.fromarray(img[..., 1], "L").save(path, **kw)
    else:
        Image.fromarray(img, "RGB").save(path, subsampling=sub, **kw)
    got = decode(path).astype(np.int32)
    want = np.asarray(Image.open(path).convert("RGB")).astype(np.int32)
    d = np.abs(got - want)
    assert d.mean() <= 0.6, d.mean()
    assert d.max() <= (8 if sub in (1, 2) else 3), d.max()


def test_progressive_jpeg_is_rejected(tmp_path):
    path = tmp_path / "p.jpg"
    Image.fromarray(_photo(32, 32, 1)).save(path, progressive=True)
    w, h = C.c_int(), C.c_int()
    assert L().msplat_image_decode(str(path).encode(), C.byref(w), C.byref(h), None) == -1


# ---------------------------------------------------------------------------- resize area
def _area_resize(src, dw, dh):
    """Independent formulation: each output pixel is the integral of the piecewise-constant
    source over its footprint divided by the footprint's area, in float64."""
    sh, sw = src.shape[:2]
    ex = np.linspace(0, sw, dw + 1)
    ey = np.linspace(0, sh, dh + 1)

    def weights(edges, n):
        W = np.zeros((len(edges) - 1, n))
        for i in range(len(edges) - 1):
            for j in range(n):
                W[i, j] = max(0.0, min(edges[i + 1], j + 1) - max(edges[i], j))
        return W / W.sum(1, keepdims=True)
    return np.einsum("ij,jkc,lk->ilc", weights(ey, sh), src.astype(np.float64), weights(ex, sw))


@pytest.mark.parametrize("sw,sh,dw,dh", [(64, 48, 32, 24), (50, 37, 17, 11), (31, 29, 31, 29), (20, 10, 30, 25)])
def test_resize_area(sw, sh, dw, dh):
    src = np.random.default_rng(0).random((sh, sw, 3)).astype(np.float32)
    dst = np.zeros((dh, dw, 3), np.float32)
    assert L().msplat_image_resize_area(_p(src), sw, sh, _p(dst), dw, dh) == 0
    assert np.abs(dst - _area_resize(src, dw, dh)).max() < 2e-6
    if (sw, sh) == (dw, dh):
        assert np.array_equal(dst, src)


# ------------------------------------------------------------------------------ undistort
def test_undistort_identity_and_forward_model():
    rng = np.random.default_rng(2)
    h, w = 60, 80
    src = rng.random((h, w, 3)).astype(np.float32)
    fx, fy, cx, cy = 70.0, 72.0, 41.0, 29.0
    dst = np.zeros_like(src)
    ow, oh = C.c_int(), C.c_int()
    intr = np.zeros(4, np.float32)
    zero = np.zeros(5, np.float32)
    assert L().msplat_image_undistort(_p(src), w, h, fx, fy, cx, cy, _p(zero), _p(dst), C.byref(ow), C.byref(oh),
                                      _p(intr)) == 0
    # without distortion the inner rectangle runs from 0 to the last pixel's coordinate
    # (w - 1, h - 1), so the last column and row go (the reference calls it only for
    # distorted cameras)
    assert (ow.value, oh.value) == (w - 1, h - 1)
    # (sample positions ((x - cx) / fx) * fx + cx are x to within float rounding)
    assert np.abs(dst.reshape(-1)[: (w - 1) * (h - 1) * 3].reshape(h - 1, w - 1, 3) - src[: h - 1, : w - 1]).max() < 1e-4
    assert np.allclose(intr, [fx, fy, cx, cy])
    k1, k2, p1, p2, k3 = 0.15, 0.02, 0.002, -0.001, 0.0  # pincushion: the valid region shrinks
    dist = np.array([k1, k2, p1, p2, k3], np.float32)
    assert L().msplat_image_undistort(_p(src), w, h, fx, fy, cx, cy, _p(dist), _p(dst), C.byref(ow), C.byref(oh),
                                      _p(intr)) == 0
    W, H = ow.value, oh.value
    assert 0 < W < w and 0 < H < h
    out = dst.reshape(-1)[: W * H * 3].reshape(H, W, 3)
    ox, oy = cx - intr[2], cy - intr[3]  # crop offset
    assert float(ox).is_integer() and float(oy).is_integer() and intr[0] == fx and intr[1] == fy
    # forward model: output pixel (u, v) of the uncropped image samples the source at the
    # distorted position (bilinear)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    x = (xx + ox - cx) / fx
    y = (yy + oy - cy) / fy
    r2 = x * x + y * y
    rad = 1 + k1 * r2 + k2 * r2 * r2 + k3 * r2 ** 3
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    sx, sy = xd * fx + cx, yd * fy + cy
    # the crop keeps the whole output inside the source (no border)
    assert sx.min() >= -0.5 and sx.max() <= w - 0.5 and sy.min() >= -0.5 and sy.max() <= h - 0.5
    x0 = np.clip(np.floor(sx).astype(int), 0, w - 1)
    x1 = np.clip(np.floor(sx).astype(int) + 1, 0, w - 1)
    y0 = np.clip(np.floor(sy).astype(int), 0, h - 1)
    y1 = np.clip(np.floor(sy).astype(int) + 1, 0, h - 1)
    tx, ty = (sx - np.floor(sx))[..., None], (sy - np.floor(sy))[..., None]
    s = src.astype(np.float64)
    want = (s[y0, x0] * (1 - tx) + s[y0, x1] * tx) * (1 - ty) + (s[y1, x0] * (1 - tx) + s[y1, x1] * tx) * ty
    assert np.abs(out - want).max() < 2e-4


# ---------------------------------------------------------------------------- scenes: COLMAP
def _expected_c2w(q, t):
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()  # scipy: x, y, z, w
    c2w = np.eye(4)
    c2w[:3, :3] = R.T
    c2w[:3, 3] = -R.T @ np.asarray(t)
    c2w[:3, 1:3] *= -1  # OpenGL camera axes
    return c2w


def _scene(path, images=None):
    s = L().msplat_scene_load(str(path).encode(), images.encode() if images else None)
    assert s, "scene failed to load"
    return s


def _camera(s, i):
    c2w = np.zeros(16, np.float32)
    intr = np.zeros(4, np.float32)
    dist = np.zeros(5, np.float32)
    w, h = C.c_int(), C.c_int()
    path = C.create_string_buffer(4096)
    rel = C.create_string_buffer(4096)
    assert L().msplat_scene_camera(s, i, _p(c2w), _p(intr), _p(dist), C.byref(w), C.byref(h), path, 4096, rel, 4096) == 0
    return {"c2w": c2w.reshape(4, 4), "intr": intr, "dist": dist, "w": w.value, "h": h.value,
            "path": path.value.decode(), "rel": rel.value.decode()}


def _points(s):
    n = L().msplat_scene_num_points(s)
    xyz = np.zeros((n, 3), np.float32)
    rgb = np.zeros((n, 3), np.uint8)
    assert L().msplat_scene_points(s, _p(xyz), _p(rgb)) == 0
    return xyz, rgb


def _transform(s):
    sc = C.c_float()
    tr = np.zeros(3, np.float32)
    assert L().msplat_scene_transform(s, C.byref(sc), _p(tr)) == 0
    return sc.value, tr


def _colmap_scene(root, sparse_sub=True, n_img=5):
    rng = np.random.default_rng(7)
    cams = {1: (1, 40, 30, [35.0, 36.0, 20.0, 15.0]),                       # PINHOLE
            7: (4, 40, 30, [33.0, 34.0, 19.5, 14.5, -0.05, 0.01, 0.001, -0.002]),  # OPENCV
            9: (2, 40, 30, [30.0, 20.0, 15.0, 0.02])}                       # SIMPLE_RADIAL
    imgs, names = [], ["c.png", "a.png", "e.png", "b.png", "d.png"][:n_img]
    for k, name in enumerate(names):
        q = Rotation.random(random_state=k).as_quat()  # x, y, z, w
        q = [q[3], q[0], q[1], q[2]]
        t = rng.normal(0, 2, 3)
        imgs.append((k + 10, q, t, [1, 7, 9][k % 3], name, k))
    xyz = rng.normal(0, 1, (50, 3))
    rgb = rng.integers(0, 256, (50, 3), dtype=np.uint8)
    sparse = os.path.join(root, "sparse", "0") if sparse_sub else root
    write_colmap(sparse, cams, imgs, (xyz, rgb))
    os.makedirs(os.path.join(root, "images"), exist_ok=True)
    for k, name in enumerate(names):
        Image.fromarray(_photo(30, 40, k)).save(os.path.join(root, "images", name))
    return cams, imgs, xyz, rgb


def test_colmap_scene(tmp_path):
    cams, imgs, xyz, rgb = _colmap_scene(str(tmp_path))
    s = _scene(tmp_path)
    try:
        n = L().msplat_scene_num_cameras(s)
        assert n == 5
        order = sorted(imgs, key=lambda r: r[4])
        c2ws = np.stack([_expected_c2w(r[1], r[2]) for r in order])
        mean = c2ws[:, :3, 3].mean(0)
        scale = 1.0 / np.abs(c2ws[:, :3, 3] - mean).max()
        sc, tr = _transform(s)
        assert np.allclose(tr, mean, atol=1e-5) and abs(sc - scale) < 1e-5 * scale
        for i, r in enumerate(order):
            c = _camera(s, i)
            want = c2ws[i].copy()
            want[:3, 3] = (want[:3, 3] - mean) * scale
            assert np.allclose(c["c2w"], want, atol=2e-5), (i, c["c2w"], want)
            model, w, h, params = cams[r[3]]
            assert (c["w"], c["h"]) == (w, h)
            assert c["rel"] == r[4] and c["path"] == os.path.join(str(tmp_path), "images", r[4])
            fx, fy, cx, cy = (params[0], params[0], params[1], params[2]) if model == 2 else params[:4]
            assert np.allclose(c["intr"], [fx, fy, cx, cy])
            k1, k2, p1, p2 = {1: (0, 0, 0, 0), 4: tuple(params[4:8]), 2: (params[3], 0, 0, 0)}[model]
            assert np.allclose(c["dist"], [k1, k2, p1, p2, 0.0])
        p_xyz, p_rgb = _points(s)
        assert np.allclose(p_xyz, (xyz - mean) * scale, atol=1e-5) and np.array_equal(p_rgb, rgb)
    finally:
        L().msplat_scene_free(s)


def test_colmap_ply_points_and_image_folder(tmp_path):
    """cameras.bin directly in the folder, points from points3D.ply (binary with 16-bit
    colours: their high bytes), images from an explicit folder."""
    rng = np.random.default_rng(3)
    cams = {1: (0, 16, 12, [14.0, 8.0, 6.0])}  # SIMPLE_PINHOLE
    q = [1.0, 0.0, 0.0, 0.0]
    write_colmap(str(tmp_path), cams, [(1, q, [0, 0, 1.0], 1, "x.jpg", 0), (2, q, [0, 0, -1.0], 1, "y.jpg", 2)])
    xyz = rng.normal(0, 1, (9, 3)).astype(np.float32)
    rgb16 = rng.integers(0, 65536, (9, 3))
    write_ply(str(tmp_path / "points3D.ply"), xyz, rgb16, color_type="ushort")
    s = _scene(tmp_path, images="/data/imgs")
    try:
        assert _camera(s, 0)["path"] == "/data/imgs/x.jpg"
        _, p_rgb = _points(s)
        assert np.array_equal(p_rgb, (rgb16 >> 8).astype(np.uint8))
    finally:
        L().msplat_scene_free(s)


# ------------------------------------------------------------------------ scenes: Nerfstudio
def test_nerfstudio_scene(tmp_path):
    rng = np.random.default_rng(5)
    frames, mats = [], {}
    for k, name in enumerate(["img_2", "img_0", "img_1"]):
        M = np.eye(4)
        M[:3, :3] = Rotation.random(random_state=k + 3).as_matrix()
        M[:3, 3] = rng.normal(0, 3, 3)
        fr = {"file_path": "images/" + name, "transform_matrix": M.tolist()}
        if k == 1:
            fr.update({"fl_x": 50.0, "w": 64, "k1": 0.01})
        frames.append(fr)
        mats[name] = M
        os.makedirs(tmp_path / "images", exist_ok=True)
        Image.fromarray(_photo(24, 32, k)).save(tmp_path / "images" / (name + ".png"))
    xyz = rng.normal(0, 1, (12, 3)).astype(np.float32)
    rgb = rng.random((12, 3)).astype(np.float32)
    write_ply(str(tmp_path / "pc.ply"), xyz, rgb, ascii=True, color_type="float")
    write_json(tmp_path / "transforms.json", {"w": 32, "h": 24, "fl_x": 30.0, "fl_y": 31.0, "cx": 16.0, "cy": 12.0,
                                               "ply_file_path": "pc.ply", "frames": frames})
    s = _scene(tmp_path)
    try:
        assert L().msplat_scene_num_cameras(s) == 3
        names = sorted(mats)  # sorted by (resolved) path
        centers = np.stack([mats[n][:3, 3] for n in names])
        mean = centers.mean(0)
        scale = 1.0 / np.abs(centers - mean).max()
        for i, n in enumerate(names):
            c = _camera(s, i)
            assert c["path"].endswith("images/%s.png" % n) and c["rel"] == ""
            want = mats[n].copy()
            want[:3, 3] = (want[:3, 3] - mean) * scale
            assert np.allclose(c["c2w"], want, atol=2e-5)
            if n == "img_0":
                assert c["w"] == 64 and np.allclose(c["intr"], [50, 31, 16, 12]) and np.isclose(c["dist"][0], 0.01)
            else:
                assert c["w"] == 32 and np.allclose(c["intr"], [30, 31, 16, 12]) and not c["dist"].any()
        p_xyz, p_rgb = _points(s)
        assert np.allclose(p_xyz, (xyz - mean) * scale, atol=1e-5)
        assert np.array_equal(p_rgb, (rgb * np.float32(255)).astype(np.uint8))
    finally:
        L().msplat_scene_free(s)


# --------------------------------------------------------------------------- scenes: Polycam
def test_polycam_keyframes_and_cameras_json(tmp_path):
    rng = np.random.default_rng(9)
    kf = tmp_path / "a" / "keyframes" / "corrected_cameras"
    im = tmp_path / "a" / "keyframes" / "corrected_images"
    os.makedirs(kf)
    os.makedirs(im)
    Rs, Ts = {}, {}
    for stem in ["20", "10"]:
        R = Rotation.random(random_state=int(stem)).as_matrix()
        T = rng.normal(0, 1, 3)
        d = {"width": 32, "height": 24, "fx": 30.0, "fy": 30.5}
        for r in range(3):
            for c in range(3):
                d["R_%d%d" % (r, c)] = R[r, c]
        d.update({"t_20": T[0], "t_21": T[1], "t_22": T[2]})
        write_json(kf / (stem + ".json"), d)
        Image.fromarray(_photo(24, 32, 1)).save(im / (stem + ".jpg"))
        Rs[stem], Ts[stem] = R, T
    write_ply(str(tmp_path / "a" / "keyframes" / "point_cloud.ply"), rng.normal(0, 1, (5, 3)).astype(np.float32))
    s = _scene(tmp_path / "a")
    try:
        expect = {}
        for stem in ["10", "20"]:
            M = np.eye(4)
            M[:3, :3] = Rs[stem] * np.array([[1], [-1], [-1]])
            M[:3, 3] = Ts[stem] * np.array([1, -1, -1])
            expect[stem] = M
        centers = np.stack([expect[k][:3, 3] for k in ["10", "20"]])
        mean = centers.mean(0)
        scale = 1.0 / np.abs(centers - mean).max()
        for i, stem in enumerate(["10", "20"]):
            c = _camera(s, i)
            want = expect[stem].copy()
            want[:3, 3] = (want[:3, 3] - mean) * scale
            assert np.allclose(c["c2w"], want, atol=2e-5)
            assert c["path"].endswith("corrected_images/%s.jpg" % stem)
            assert np.allclose(c["intr"], [30, 30.5, 16, 12])
        _, p_rgb = _points(s)
        assert (p_rgb == 128).all()  # no colours: grey
    finally:
        L().msplat_scene_free(s)
    # layout 2: cameras.json with a top-level array
    b = tmp_path / "b"
    os.makedirs(b)
    M = np.eye(4)
    M[:3, 3] = [1, 2, 3]
    write_json(b / "cameras.json", [{"width": 10, "height": 8, "fx": 9.0, "fy": 9.0, "file_path": "f.png",
                                     "transform_matrix": M.tolist()},
                                    {"width": 10, "height": 8, "fx": 9.0, "fy": 9.0, "cx": 4.0, "file_path": "g.png",
                                     "transform_matrix": (M * 0 + np.eye(4)).tolist()}])
    write_ply(str(b / "sparse.ply"), rng.normal(0, 1, (4, 3)).astype(np.float32))
    s = _scene(b)
    try:
        c0, c1 = _camera(s, 0), _camera(s, 1)
        assert c0["path"] == str(b / "f.png") and np.allclose(c0["intr"], [9, 9, 5, 4]) and c1["intr"][2] == 4.0
        # centres (1, 2, 3) and 0: mean (0.5, 1, 1.5), then divided by 1.5
        assert np.allclose(c0["c2w"][:3, 3], [1 / 3, 2 / 3, 1]) and np.allclose(c1["c2w"][:3, 3], [-1 / 3, -2 / 3, -1])
    finally:
        L().msplat_scene_free(s)


def test_unrecognised_folder_fails(tmp_path):
    assert not L().msplat_scene_load(str(tmp_path).encode(), None)


# ------------------------------------------------------------------ camera images and masks
def test_camera_image_rescale_downscale_undistort(tmp_path):
    """Camera::loadImage: a file of another size rescales the intrinsics; downscale by area
    with the true ratio; distortion removed; round-to-nearest uint8."""
    _colmap_scene(str(tmp_path), n_img=3)
    # camera "c.png" (PINHOLE 40x30, second by name): replace the file by an 80x60 image
    big = _photo(60, 80, 11)
    Image.fromarray(big).save(tmp_path / "images" / "c.png")
    s = _scene(tmp_path)
    try:
        i = 1
        assert _camera(s, i)["rel"] == "c.png"
        assert L().msplat_scene_load_image(s, i, C.c_float(3.0)) == 0
        c = _camera(s, i)
        assert (c["w"], c["h"]) == (26, 20)  # int(80 / 3), int(60 / 3)
        sx, sy = 2.0 * 26 / 80, 2.0 * 20 / 60
        assert np.allclose(c["intr"], [35 * sx, 36 * sy, 20 * sx, 15 * sy], rtol=1e-6)
        img = np.zeros((20, 26, 3), np.uint8)
        assert L().msplat_scene_image(s, i, _p(img)) == 0
        want = np.clip(np.round(_area_resize(big.astype(np.float32) / np.float32(255), 26, 20) * 255), 0, 255)
        assert np.abs(img.astype(int) - want).max() <= 1
        # "a.png": OPENCV camera (k1 < 0: barrel), undistorted: the distortion is gone and
        # the size is the inner rectangle's
        assert _camera(s, 0)["rel"] == "a.png" and _camera(s, 0)["dist"].any()
        assert L().msplat_scene_load_image(s, 0, C.c_float(1.0)) == 0
        c = _camera(s, 0)
        assert c["w"] <= 40 and c["h"] <= 30 and not c["dist"].any()
    finally:
        L().msplat_scene_free(s)


def test_camera_masks(tmp_path):
    _colmap_scene(str(tmp_path), n_img=3)
    md = tmp_path / "masks"
    os.makedirs(md)
    m = np.full((30, 40), 255, np.uint8)
    m[5:15, 10:30] = 0
    m[0, 0], m[0, 1] = 128, 127
    # cameras by name: a.png, c.png, e.png
    Image.fromarray(m, "L").save(md / "a.png.png")          # "<name>.png"
    Image.fromarray(np.full((15, 20), 255, np.uint8), "L").save(md / "c.png")  # extension replaced; all white
    Image.fromarray(np.kron(m[::2, ::2], np.ones((4, 4), np.uint8)), "L").save(md / "e.png")  # 80x60: resized
    s = _scene(tmp_path)
    try:
        out = np.zeros((30, 40), np.uint8)
        kept = L().msplat_scene_load_mask(s, 0, str(md).encode(), _p(out))
        assert kept == int((m >= 128).sum()) and np.array_equal(out, np.where(m >= 128, 255, 0))
        assert L().msplat_scene_load_mask(s, 1, str(md).encode(), _p(out)) == 0   # all white: no mask
        kept = L().msplat_scene_load_mask(s, 2, str(md).encode(), _p(out))
        assert kept == int((m[::2, ::2] >= 128).sum()) * 4
        assert L().msplat_scene_load_mask(s, 0, str(tmp_path / "none").encode(), _p(out)) == 0
    finally:
        L().msplat_scene_free(s)


# ------------------------------------------------------------------------------ seeding
def _brute_knn(xyz, k=4):
    x = xyz.astype(np.float32)
    out = np.zeros(len(x), np.float32)
    for i in range(len(x)):
        d = x[i] - x
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]  # float32, axis order
        s = np.sort(dd)[:k]
        acc = np.float32(0)
        for j in range(1, k):
            acc = np.float32(acc + np.sqrt(s[j]))
        out[i] = acc / np.float32(k - 1)
    return out


@pytest.mark.parametrize("kind", ["uniform", "clustered", "duplicates", "flat"])
def test_knn_mean_distance_is_exact(kind):
    rng = np.random.default_rng(11)
    if kind == "uniform":
        xyz = rng.uniform(-1, 1, (3000, 3))
    elif kind == "clustered":
        xyz = np.concatenate([rng.normal(c, 0.01, (500, 3)) for c in rng.uniform(-5, 5, (6, 3))])
    elif kind == "duplicates":
        xyz = rng.uniform(-1, 1, (800, 3))
        xyz = np.concatenate([xyz, xyz[:200]])
    else:
        xyz = np.concatenate([rng.uniform(-1, 1, (1500, 2)), np.zeros((1500, 1))], 1)
    xyz = xyz.astype(np.float32)
    out = np.zeros(len(xyz), np.float32)
    assert L().msplat_knn_mean_distance(_p(xyz), len(xyz), 4, _p(out)) == 0
    assert np.array_equal(out, _brute_knn(xyz))


def _mt_uniform_floats(n):
    """std::mt19937(42) draws as uniform floats: each 32-bit draw converted to float32 and
    divided by 2^32 (generate_canonical<float> with one draw)."""
    raw = np.random.RandomState(42)._bit_generator.random_raw(n).astype(np.uint32)
    return raw.astype(np.float32) / np.float32(4294967296.0)


def test_seed_gaussians():
    rng = np.random.default_rng(4)
    n = 700
    xyz = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
    out = {k: np.zeros(s, np.float32) for k, s in
           {"means": (n, 3), "scales": (n, 3), "quats": (n, 4), "opac": (n,), "f_dc": (n, 3), "f_rest": (n, 15, 3)}.items()}
    out["f_rest"][:] = 7
    assert L().msplat_seed_gaussians(_p(xyz), _p(rgb), n, 3, C.c_float(0.1), C.c_float(0.5), _p(out["means"]),
                                     _p(out["scales"]), _p(out["quats"]), _p(out["opac"]), _p(out["f_dc"]),
                                     _p(out["f_rest"])) == 0
    assert np.array_equal(out["means"], xyz)
    ls = np.log(_brute_knn(xyz) * np.float32(0.1))
    assert np.allclose(out["scales"], np.repeat(ls[:, None], 3, 1), rtol=0, atol=1e-6)
    u = _mt_uniform_floats(3 * n).reshape(n, 3).astype(np.float64)
    su, s1u = np.sqrt(u[:, 0].astype(np.float32)), np.sqrt(np.float32(1) - u[:, 0].astype(np.float32))
    q = np.stack([s1u * np.sin(2 * np.pi * u[:, 1]), s1u * np.cos(2 * np.pi * u[:, 1]),
                  su * np.sin(2 * np.pi * u[:, 2]), su * np.cos(2 * np.pi * u[:, 2])], 1).astype(np.float32)
    assert np.abs(out["quats"] - q).max() <= 1e-7
    assert np.allclose(np.linalg.norm(out["quats"], axis=1), 1, atol=1e-6)
    assert np.array_equal(out["f_dc"], ((rgb / 255.0 - 0.5) / 0.28209479177387814).astype(np.float32))
    assert not out["f_rest"].any() and not out["opac"].any()

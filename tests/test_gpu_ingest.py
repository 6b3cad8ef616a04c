"""A scene folder end to end on the GPU (msplat_dataset_create on a COLMAP folder,
Dataset::Dataset in msplat_api.mm:25-41): the train / held-out split, every camera's pose,
intrinsics and uint8 ground truth equal to the host loader's (msplat_io.h), the trainer's
initial Gaussians equal to the seeding from the point cloud bit for bit, masks from a mask
folder, and training steps on it."""
import ctypes as C
import os

import numpy as np
import pytest
from PIL import Image
from scipy.spatial.transform import Rotation

from ingest_fixtures import write_colmap
from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib
from test_gpu_trainer import GROUPS, _state

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data


def _folder(root, n_cams=6, n_pts=2000):
    rng = np.random.default_rng(21)
    cams = {1: (1, 96, 72, [80.0, 80.0, 48.0, 36.0]), 2: (4, 96, 72, [78.0, 79.0, 47.0, 35.5, -0.02, 0.004, 0, 0])}
    imgs = []
    for k in range(n_cams):
        # cameras on a circle looking at the origin (world-to-camera rotation and translation)
        ang = 2 * np.pi * k / n_cams
        c = np.array([3 * np.cos(ang), 0.3 * k, 3 * np.sin(ang)])
        z = -c / np.linalg.norm(c)
        x = np.cross([0, 1, 0], z)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        R = np.stack([x, y, z])  # rows: camera axes in world coordinates (OpenCV)
        q = Rotation.from_matrix(R).as_quat()
        imgs.append((k + 1, [q[3], q[0], q[1], q[2]], -R @ c, 1 + k % 2, "img_%02d.png" % k, 0))
    xyz = rng.uniform(-1, 1, (n_pts, 3))
    rgb = rng.integers(0, 256, (n_pts, 3), dtype=np.uint8)
    write_colmap(os.path.join(root, "sparse", "0"), cams, imgs, (xyz, rgb))
    os.makedirs(os.path.join(root, "images"))
    for k in range(n_cams):
        y, x = np.mgrid[0:72, 0:96]
        img = np.stack([128 + 90 * np.sin(x / 7.0 + k + c) * np.cos(y / 9.0) for c in range(3)], -1)
        Image.fromarray(np.clip(img, 0, 255).astype(np.uint8)).save(os.path.join(root, "images", "img_%02d.png" % k))
    return n_pts


def test_colmap_folder_dataset_and_training(tmp_path):
    L = _lib.lib()
    n = _folder(str(tmp_path))
    ds = Dataset(str(tmp_path), eval_mode=True, test_every=3)
    assert ds.num_train == 4 and ds.num_test == 2  # cameras 0 and 3 held out
    sc = L.msplat_scene_load(str(tmp_path).encode(), None)
    try:
        for i in range(6):
            assert L.msplat_scene_load_image(sc, i, C.c_float(1.0)) == 0
            c2w = np.zeros(16, np.float32)
            intr = np.zeros(4, np.float32)
            w, h = C.c_int(), C.c_int()
            assert L.msplat_scene_camera(sc, i, _p(c2w), _p(intr), None, C.byref(w), C.byref(h), None, 0, None, 0) == 0
            img = np.zeros((h.value, w.value, 3), np.uint8)
            assert L.msplat_scene_image(sc, i, _p(img)) == 0
            test = i % 3 == 0
            cam = ds.camera(i // 3 if test else i - 1 - i // 3, test)
            assert np.array_equal(cam["c2w"].reshape(-1), c2w)
            assert np.allclose([cam["fx"], cam["fy"], cam["cx"], cam["cy"]], intr)
            assert (cam["width"], cam["height"]) == (w.value, h.value)
            assert np.array_equal(cam["image"], img)
        npts = L.msplat_scene_num_points(sc)
        xyz = np.zeros((npts, 3), np.float32)
        rgb = np.zeros((npts, 3), np.uint8)
        L.msplat_scene_points(sc, _p(xyz), _p(rgb))
    finally:
        L.msplat_scene_free(sc)
    seeds = {k: np.zeros(s, np.float32) for k, s in
             {"means": (n, 3), "scales": (n, 3), "quats": (n, 4), "opac": (n,), "f_dc": (n, 3), "f_rest": (n, 15, 3)}.items()}
    assert L.msplat_seed_gaussians(_p(xyz), _p(rgb), n, 3, C.c_float(0.1), C.c_float(0.5), _p(seeds["means"]),
                                   _p(seeds["scales"]), _p(seeds["quats"]), _p(seeds["opac"]), _p(seeds["f_dc"]),
                                   _p(seeds["f_rest"])) == 0
    tr = GaussianTrainer(ds, TrainingConfig(iterations=100, cap_max=4 * n, sh_degree_interval=1))
    st = _state(tr)
    for k in GROUPS:
        assert np.array_equal(st[k], seeds[k].reshape(-1)), k
    # masks for two cameras by their COLMAP names
    md = tmp_path / "masks"
    os.makedirs(md)
    m = np.full((72, 96), 255, np.uint8)
    m[20:40, 30:70] = 0
    Image.fromarray(m, "L").save(md / "img_01.png.png")
    Image.fromarray(m[::-1], "L").save(md / "img_02.png")
    assert ds.load_masks(str(md)) == 2
    for _ in range(12):
        s = tr.step()
        assert np.isfinite(tr.last_loss) and s.splat_count == n
    m_eval = tr.evaluate()
    assert np.isfinite(m_eval["psnr"])

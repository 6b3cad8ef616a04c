"""Python API of msplat (the reference's `msplat._core`, python/msplat/_core.pyi:6-167),
implemented over the C API of libmsplat_core.so.

    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig
    ds = Dataset("synthetic:c0", eval_mode=True)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=200))
    tr.train(lambda s: print(s.iteration, s.splat_count), callback_every=50)

Every call goes through the same C entry points a Swift or C caller uses; the library
raises nothing across the boundary, so failures surface here as RuntimeError.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import fptr

DEFAULT_BG = [0.6130, 0.0101, 0.3984]


@dataclass
class TrainingConfig:
    """Training options (reference _core.pyi:6-48). The options of the reference's older
    ADC densification (num_downscales ... split_screen_size, output, save_every) are kept
    for signature compatibility and, as in the reference's MCMC trainer, unused."""
    iterations: int = 30000
    sh_degree: int = 3
    sh_degree_interval: int = 1000
    ssim_weight: float = 0.2
    num_downscales: int = 2
    resolution_schedule: int = 3000
    refine_every: int = 100
    warmup_length: int = 500
    reset_alpha_every: int = 30
    densify_grad_thresh: float = 0.0002
    densify_size_thresh: float = 0.01
    stop_screen_size_at: int = 4000
    split_screen_size: float = 0.05
    keep_crs: bool = False
    downscale_factor: float = 1.0
    output: str = "splat.ply"
    save_every: int = -1
    bg_color: list = field(default_factory=lambda: list(DEFAULT_BG))
    # MCMC options (the C API's MsplatConfig; not in the reference's Python stub)
    cap_max: int = 1000000
    noise_lr: float = 5e5
    opacity_reg: float = 0.01
    scale_reg: float = 0.01
    cull_radius: float = 3.0

    def to_c(self) -> _lib.MsplatConfig:
        c = _lib.default_config()
        c.iterations = int(self.iterations)
        c.shDegree = int(self.sh_degree)
        c.shDegreeInterval = int(self.sh_degree_interval)
        c.ssimWeight = float(self.ssim_weight)
        c.capMax = int(self.cap_max)
        c.noiseLr = float(self.noise_lr)
        c.opacityReg = float(self.opacity_reg)
        c.scaleReg = float(self.scale_reg)
        c.cullRadius = float(self.cull_radius)
        c.keepCrs = bool(self.keep_crs)
        c.downscaleFactor = float(self.downscale_factor)
        for k in range(3):
            c.bgColor[k] = float(self.bg_color[k])
        return c


class TrainingStats:
    """Per-step statistics (reference _core.pyi:50-66)."""

    __slots__ = ("_s",)

    def __init__(self, s: _lib.MsplatStats):
        self._s = s

    @property
    def iteration(self) -> int:
        return int(self._s.iteration)

    @property
    def splat_count(self) -> int:
        return int(self._s.splatCount)

    @property
    def ms_per_step(self) -> float:
        return float(self._s.msPerStep)

    def __repr__(self):
        return f"TrainingStats(iteration={self.iteration}, splat_count={self.splat_count}, " \
               f"ms_per_step={self.ms_per_step:.3f})"


class Dataset:
    """Cameras with uint8 ground truth on the GPU (reference _core.pyi:68-90).

    `path` is a scene folder (COLMAP, Nerfstudio or Polycam; include/msplat_io.h) or
    "synthetic:<preset>[:key=value,...]" (include/msplat_api_ext.h); use
    Dataset.from_arrays for data already in memory."""

    def __init__(self, path: str, downscale_factor: float = 1.0, eval_mode: bool = False, test_every: int = 8):
        L = _lib.lib()
        self._h = L.msplat_dataset_create(path.encode(), float(downscale_factor), bool(eval_mode), int(test_every))
        if not self._h:
            raise RuntimeError(f"could not create dataset from {path!r} (reason on stderr)")

    @classmethod
    def from_arrays(cls, cameras, gaussians: dict):
        """cameras: iterable of dicts with c2w (4x4, OpenGL), fx, fy, cx, cy, image (H, W, 3
        uint8), optional test flag and optional mask (H, W uint8 training mask, white = use);
        gaussians: means, scales (log), quats, opac (logit), f_dc, f_rest, degree."""
        L = _lib.lib()
        self = cls.__new__(cls)
        self._h = L.msplat_dataset_create_empty()
        if not self._h:
            raise RuntimeError("could not create an empty dataset")
        for cam in cameras:
            img = np.ascontiguousarray(cam["image"], np.uint8)
            h, w = img.shape[:2]
            c2w = np.ascontiguousarray(np.asarray(cam["c2w"], np.float32).reshape(16))
            r = L.msplat_dataset_add_camera(self._h, fptr(c2w), cam["fx"], cam["fy"], cam["cx"], cam["cy"], w, h,
                                            img.ctypes.data_as(C.POINTER(C.c_uint8)), bool(cam.get("test", False)))
            if r < 0:
                raise RuntimeError("could not add camera")
            if cam.get("mask") is not None:
                self.set_mask(r, cam["mask"], bool(cam.get("test", False)))
        g = {k: np.ascontiguousarray(gaussians[k], np.float32) for k in
             ("means", "scales", "quats", "opac", "f_dc", "f_rest")}
        n = g["means"].shape[0]
        if L.msplat_dataset_set_init_gaussians(self._h, n, int(gaussians["degree"]), fptr(g["means"]),
                                               fptr(g["scales"]), fptr(g["quats"]), fptr(g["opac"]),
                                               fptr(g["f_dc"]), fptr(g["f_rest"])) != 0:
            raise RuntimeError("could not set initial Gaussians")
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().msplat_dataset_destroy(h)
            self._h = None

    @property
    def num_train(self) -> int:
        return _lib.lib().msplat_dataset_num_train(self._h)

    @property
    def num_test(self) -> int:
        return _lib.lib().msplat_dataset_num_test(self._h)

    def camera(self, index: int, test: bool = False, image: bool = True) -> dict:
        """Pose, intrinsics, size and (optionally) the uint8 ground truth of one camera."""
        L = _lib.lib()
        c2w = np.zeros(16, np.float32)
        intr = np.zeros(4, np.float32)
        w, h = C.c_int(), C.c_int()
        if L.msplat_dataset_camera_info(self._h, int(index), bool(test), fptr(c2w), fptr(intr), C.byref(w),
                                        C.byref(h), None) != 0:
            raise IndexError(index)
        out = {"c2w": c2w.reshape(4, 4), "fx": float(intr[0]), "fy": float(intr[1]), "cx": float(intr[2]),
               "cy": float(intr[3]), "width": w.value, "height": h.value}
        if image:
            img = np.zeros((h.value, w.value, 3), np.uint8)
            L.msplat_dataset_camera_info(self._h, int(index), bool(test), None, None, None, None,
                                         img.ctypes.data_as(C.POINTER(C.c_uint8)))
            out["image"] = img
        return out

    def load_masks(self, mask_dir: str) -> int:
        """Training masks of a scene-folder dataset from `mask_dir` ("<name>.png" or the
        name with its extension replaced by .png, white = use); returns the number of
        cameras masked."""
        r = _lib.lib().msplat_dataset_load_masks(self._h, str(mask_dir).encode())
        if r < 0:
            raise RuntimeError("could not load the masks (reason on stderr)")
        return int(r)

    def set_mask(self, index: int, mask: np.ndarray, test: bool = False) -> int:
        """Training mask (H, W uint8, white = use, black = ignore) of one camera, binarised at
        one half; an all-white mask is dropped. Returns the count of kept pixels."""
        m = np.ascontiguousarray(mask, np.uint8)
        info = self.camera(index, test, image=False)
        if m.shape != (info["height"], info["width"]):
            raise ValueError(f"mask shape {m.shape} != {(info['height'], info['width'])}")
        r = _lib.lib().msplat_dataset_set_mask(self._h, int(index), bool(test), m.ctypes.data_as(C.POINTER(C.c_uint8)))
        if r < 0:
            raise RuntimeError("could not set the mask")
        return int(r)

    def camera_pose(self, index: int) -> np.ndarray:
        out = np.zeros(16, np.float32)
        _lib.lib().msplat_dataset_camera_pose(self._h, int(index), fptr(out))
        return out.reshape(4, 4)


def _take_pixels(buf: _lib.MsplatPixelBuffer) -> np.ndarray:
    if not buf.data:
        return np.zeros((0, 0, 3), np.float32)
    h, w = buf.height, buf.width
    arr = np.ctypeslib.as_array(buf.data, shape=(h * w * 3,)).copy().reshape(h, w, 3)
    C.CDLL(None).free(C.cast(buf.data, C.c_void_p))
    return arr


class GaussianTrainer:
    """3D Gaussian Splatting trainer on the GPU (reference _core.pyi:92-167)."""

    def __init__(self, dataset: Dataset, config: TrainingConfig):
        L = _lib.lib()
        self._ds = dataset  # the dataset must outlive the trainer
        self._cfg = config
        self._h = L.msplat_trainer_create(dataset._h, config.to_c())
        if not self._h:
            raise RuntimeError("could not create trainer (reason on stderr)")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().msplat_trainer_destroy(h)
            self._h = None

    def step(self) -> TrainingStats:
        s = _lib.lib().msplat_trainer_step(self._h)
        if s.iteration == 0:
            raise RuntimeError("training step failed (reason on stderr)")
        return TrainingStats(s)

    def step_batch(self, views: int) -> TrainingStats:
        """One multi-view iteration over the next `views` cameras (msplat_trainer_step_batch)."""
        s = _lib.lib().msplat_trainer_step_batch(self._h, int(views))
        if s.iteration == 0:
            raise RuntimeError("batched step failed (reason on stderr)")
        return TrainingStats(s)

    def last_batch(self, losses: bool = False):
        """(camera indices, per-view losses or None) of the most recent batched step."""
        L = _lib.lib()
        n = L.msplat_trainer_last_batch(self._h, None, None, 0)
        cams = (C.c_int32 * max(n, 1))()
        ls = (C.c_double * max(n, 1))()
        L.msplat_trainer_last_batch(self._h, cams, ls if losses else None, n)
        return [int(c) for c in cams[:n]], ([float(x) for x in ls[:n]] if losses else None)

    def proposal_info(self) -> dict:
        """Counters of the learned relocation proposal (msplat_trainer_proposal_info)."""
        out = (C.c_double * 12)()
        if _lib.lib().msplat_trainer_proposal_info(self._h, out) != 0:
            raise RuntimeError("learned relocation is not enabled")
        keys = ("events", "learned_events", "plain_events", "samples", "val_bce", "base_bce", "val_off", "base_off",
                "trained", "delta_plain", "delta_learned", "heldout_loss")
        return {k: float(v) for k, v in zip(keys, out)}

    def proposal_params(self, theta=None):
        """The proposal network's parameters; replaces them when `theta` is given."""
        L = _lib.lib()
        if theta is not None:
            th = np.ascontiguousarray(theta, np.float32)
            if L.msplat_trainer_proposal_params(self._h, fptr(th), 1) != 0:
                raise RuntimeError("could not set the proposal parameters")
            return th
        th = np.zeros(5508, np.float32)
        if L.msplat_trainer_proposal_params(self._h, fptr(th), 0) != 0:
            raise RuntimeError("learned relocation is not enabled")
        return th

    def proposal_read(self, name: str, dtype=np.float32) -> np.ndarray:
        """A proposal buffer ("out", "features" as uint16, "stats") as a flat array."""
        L = _lib.lib()
        n = L.msplat_trainer_proposal_read(self._h, name.encode(), None, 0)
        if n < 0:
            raise RuntimeError(f"proposal buffer {name!r} unavailable")
        out = np.zeros(n // np.dtype(dtype).itemsize, dtype)
        L.msplat_trainer_proposal_read(self._h, name.encode(), out.ctypes.data, n)
        return out

    def autotune_info(self) -> dict:
        """Timed candidates of the batched-step autotuner, each (pixels per lane, raster
        workgroups, list rank chunks, bucket waves, recorded graph 1 / eager 0, score: its
        steps' time relative to a least-squares fit of step time on pixels over all timed
        steps, lower is faster), and the chosen (pixels per lane, workgroups, chunks, waves,
        graph) once tuning has finished."""
        cands = (C.c_int32 * (5 * 32))()
        ms = np.zeros(32, np.float32)
        ch = (C.c_int32 * 5)()
        n = _lib.lib().msplat_trainer_autotune_info(self._h, cands, fptr(ms), 32, ch)
        return {"timed": [tuple(int(cands[5 * k + j]) for j in range(5)) + (float(ms[k]),) for k in range(max(n, 0))],
                "chosen": tuple(int(ch[j]) for j in range(5))}

    def batch_info(self) -> dict:
        """Recorded-graph counters of the batched steps (msplat_trainer_batch_info)."""
        out = (C.c_int64 * 4)()
        if _lib.lib().msplat_trainer_batch_info(self._h, out) != 0:
            raise RuntimeError("batch_info failed")
        return {"captures": int(out[0]), "replays": int(out[1]), "launch": (int(out[2]), int(out[3]))}

    def train(self, callback=None, callback_every: int = 100) -> None:
        while self.iteration < self._cfg.iterations:
            s = self.step()
            if callback is not None and callback_every > 0 and s.iteration % callback_every == 0:
                callback(s)

    def evaluate(self) -> dict:
        m = _lib.lib().msplat_trainer_evaluate(self._h)
        return {"psnr": float(m.psnr), "ssim": float(m.ssim), "l1": float(m.l1), "num_test": int(m.numTest),
                "num_gaussians": int(m.numGaussians)}

    def render(self, cam_idx: int, use_test: bool = False) -> np.ndarray:
        return _take_pixels(_lib.lib().msplat_trainer_render(self._h, int(cam_idx), bool(use_test)))

    def render_from_pose(self, cam_to_world, ref_cam_idx: int = 0) -> np.ndarray:
        c2w = np.ascontiguousarray(np.asarray(cam_to_world, np.float32).reshape(16))
        return _take_pixels(_lib.lib().msplat_trainer_render_pose(self._h, fptr(c2w), int(ref_cam_idx)))

    def render_from_pose_rgba(self, cam_to_world, ref_cam_idx: int = 0) -> np.ndarray:
        L = _lib.lib()
        c2w = np.ascontiguousarray(np.asarray(cam_to_world, np.float32).reshape(16))
        w, h = C.c_int(), C.c_int()
        L.msplat_trainer_render_pose_to_buffer(self._h, fptr(c2w), int(ref_cam_idx), None, C.byref(w), C.byref(h))
        out = np.zeros((h.value, w.value, 4), np.uint8)
        if out.size:
            L.msplat_trainer_render_pose_to_buffer(self._h, fptr(c2w), int(ref_cam_idx),
                                                   out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(w), C.byref(h))
        return out

    def export_ply(self, path: str) -> None:
        _lib.lib().msplat_trainer_export_ply(self._h, path.encode())

    def export_splat(self, path: str) -> None:
        _lib.lib().msplat_trainer_export_splat(self._h, path.encode())

    def save_checkpoint(self, path: str) -> None:
        _lib.lib().msplat_trainer_save_checkpoint(self._h, path.encode())

    def load_checkpoint(self, path: str) -> None:
        if _lib.lib().msplat_trainer_load_checkpoint(self._h, path.encode()) < 0:
            raise RuntimeError(f"could not load checkpoint {path!r} (reason on stderr)")

    @property
    def splat_count(self) -> int:
        return _lib.lib().msplat_trainer_splat_count(self._h)

    @property
    def iteration(self) -> int:
        return _lib.lib().msplat_trainer_iteration(self._h)

    # ---- introspection (C API extensions) ----
    def read_group(self, group: int, which: int = 0) -> np.ndarray:
        L = _lib.lib()
        n = L.msplat_trainer_read_group(self._h, group, which, None, 0)
        if n < 0:
            raise RuntimeError("read_group failed")
        out = np.zeros(max(n, 0), np.float32)
        if n:
            L.msplat_trainer_read_group(self._h, group, which, fptr(out), n)
        return out

    def step_params(self, cam_idx: int, step: int, adam_step: int) -> _lib.MsplatStepParams:
        hp = _lib.MsplatStepParams()
        if _lib.lib().msplat_trainer_step_params(self._h, cam_idx, step, adam_step, hp) != 0:
            raise RuntimeError("step_params failed")
        return hp

    @property
    def last_camera(self) -> int:
        return _lib.lib().msplat_trainer_last_camera(self._h)

    @property
    def last_loss(self) -> float:
        return _lib.lib().msplat_trainer_last_loss(self._h)

    def run_mcmc(self, step: int) -> dict:
        info = (C.c_int32 * 6)()
        if _lib.lib().msplat_trainer_run_mcmc(self._h, step, info) != 0:
            raise RuntimeError("run_mcmc failed")
        keys = ("num_active", "dead", "grown", "nonfinite", "culled", "relocation_count")
        return dict(zip(keys, [int(x) for x in info]))

    def set_option(self, name: str, value) -> None:
        """Options beyond TrainingConfig: "sh_bf16" (bfloat16 SH coefficients and moments),
        "batch_graph" (record batched steps as HIP graphs and replay them), "persist_grid"
        (resident workgroups of the persistent raster kernels, 0 = default),
        "learned_relocation" (train and use the relocation proposal network),
        "relocation_mode" (0 judged by held-out loss, 1 plain, 2 learned), "raster_ppl"
        (pixels per lane of the persistent raster kernels, 1 or 4), "autotune"
        (time the persistent raster residencies over the first batched steps and keep the
        fastest), "batch_sequential" (batched steps view by view through the single-view
        path), "mcmc" (relocation / growth on their schedule; default 1)."""
        if _lib.lib().msplat_trainer_set_option(self._h, name.encode(), float(value)) != 0:
            raise RuntimeError(f"option {name!r} rejected (reason on stderr)")

    def optim_state(self):
        n = C.c_int()
        lr = np.zeros(6, np.float32)
        _lib.lib().msplat_trainer_optim_state(self._h, C.byref(n), fptr(lr))
        return n.value, lr

    @property
    def engine_handle(self):
        return _lib.lib().msplat_trainer_engine(self._h)


class Viewer:
    """Render-only viewer over a Gaussian PLY (msplat_viewer_* of the C API)."""

    def __init__(self, ply_path: str, width: int, height: int, fov_y: float):
        self._h = _lib.lib().msplat_viewer_create(ply_path.encode(), int(width), int(height), float(fov_y))
        if not self._h:
            raise RuntimeError(f"could not load {ply_path!r} (reason on stderr)")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().msplat_viewer_destroy(h)
            self._h = None

    def render(self, cam_to_world) -> np.ndarray:
        L = _lib.lib()
        w, h = self.size
        out = np.zeros((h, w, 4), np.uint8)
        c2w = np.ascontiguousarray(np.asarray(cam_to_world, np.float32).reshape(16))
        L.msplat_viewer_render(self._h, fptr(c2w), out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    def resize(self, width: int, height: int):
        _lib.lib().msplat_viewer_resize(self._h, int(width), int(height))

    def set_fov(self, fov_y: float):
        _lib.lib().msplat_viewer_set_fov(self._h, float(fov_y))

    def set_background(self, r: float, g: float, b: float):
        _lib.lib().msplat_viewer_set_background(self._h, float(r), float(g), float(b))

    @property
    def size(self):
        w, h = C.c_int(), C.c_int()
        _lib.lib().msplat_viewer_get_size(self._h, C.byref(w), C.byref(h))
        return w.value, h.value

    @property
    def splat_count(self) -> int:
        return _lib.lib().msplat_viewer_splat_count(self._h)

    def bounds(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        _lib.lib().msplat_viewer_bounds(self._h, fptr(lo), fptr(hi))
        return lo, hi

    def density_center(self):
        c = np.zeros(3, np.float32)
        _lib.lib().msplat_viewer_density_center(self._h, fptr(c))
        return c


def sync() -> None:
    _lib.lib().msplat_sync()


def cleanup() -> None:
    _lib.lib().msplat_cleanup()

"""Device-resident training state and the kernel-level engine, driven from Python.

Device memory is allocated through the library's own HIP runtime (msplat_dev_* in
include/msplat_kernels.h); numpy arrays are the host side. Every kernel that runs is one
of the library's hand-written HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

GROUPS = ("means", "scales", "quats", "f_dc", "f_rest", "opac")
# layout of the engine's "grad2d" buffer: one record of G2D floats per Gaussian (the raster
# backward's screen-space gradients: colour, conic, mean2d, opacity logit)
G2D = 9
G2D_RGB, G2D_CONIC, G2D_XY, G2D_OPAC = slice(0, 3), slice(3, 6), slice(6, 8), 8


class DeviceArray:
    """A float32/uint8/... array in device memory (owned, freed on close)."""

    def __init__(self, shape, dtype=np.float32, host: np.ndarray | None = None):
        self.L = _lib.lib()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = self.L.msplat_dev_alloc(max(self.nbytes, 256))
        if not self.ptr:
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")
        if host is not None:
            self.upload(host)
        else:
            self.zero()

    def upload(self, host: np.ndarray, offset_rows: int = 0):
        h = np.ascontiguousarray(host, self.dtype)
        row = self.nbytes // max(self.shape[0], 1) if self.shape else self.nbytes
        if h.nbytes + offset_rows * row > self.nbytes:
            raise ValueError("upload larger than the device array")
        if h.nbytes and self.L.msplat_dev_upload(C.c_void_p(self.ptr + offset_rows * row), h.ctypes.data, h.nbytes):
            raise RuntimeError("device upload failed")

    def download(self, rows: int | None = None) -> np.ndarray:
        shape = self.shape if rows is None else (rows,) + self.shape[1:]
        out = np.empty(shape, self.dtype)
        if out.nbytes and self.L.msplat_dev_download(out.ctypes.data, self.ptr, out.nbytes):
            raise RuntimeError("device download failed")
        return out

    def zero(self):
        if self.L.msplat_dev_memset(self.ptr, 0, max(self.nbytes, 256)):
            raise RuntimeError("device memset failed")

    def close(self):
        if self.ptr:
            self.L.msplat_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def require_device():
    if _lib.lib().msplat_dev_count() <= 0:
        raise RuntimeError("msplat_amd: no HIP device is available")


class DeviceState:
    """Parameters + Adam moments on the GPU (cap-sized rows, first N active)."""

    def __init__(self, g: dict, capacity: int | None = None, moments=None):
        require_device()
        n = g["means"].shape[0]
        self.N = n
        self.degree = int(g["degree"])
        cap = capacity or n
        self.cap = cap

        def up(a):
            a = np.ascontiguousarray(a, np.float32)
            d = DeviceArray((cap,) + a.shape[1:], np.float32)
            if n:
                d.upload(a)
            return d
        self.p = {k: up(g[k]) for k in GROUPS}
        if moments is None:
            self.m = {k: DeviceArray(self.p[k].shape) for k in GROUPS}
            self.v = {k: DeviceArray(self.p[k].shape) for k in GROUPS}
        else:
            self.m = {k: up(moments[0][k]) for k in GROUPS}
            self.v = {k: up(moments[1][k]) for k in GROUPS}

    def struct(self) -> _lib.MsplatDeviceState:
        s = _lib.MsplatDeviceState()
        s.N, s.degree = self.N, self.degree
        s.means, s.scales, s.quats = self.p["means"].ptr, self.p["scales"].ptr, self.p["quats"].ptr
        s.f_dc, s.f_rest, s.opac = self.p["f_dc"].ptr, self.p["f_rest"].ptr, self.p["opac"].ptr
        for gi, k in enumerate(GROUPS):
            s.m[gi] = self.m[k].ptr
            s.v[gi] = self.v[k].ptr
        return s

    def to_host(self) -> dict:
        out = {k: self.p[k].download(self.N) for k in GROUPS}
        out["degree"] = self.degree
        return out

    def moments_host(self):
        return ({k: self.m[k].download(self.N) for k in GROUPS},
                {k: self.v[k].download(self.N) for k in GROUPS})


class Engine:
    """Owner of one MsplatEngine (kernel-level ABI)."""

    def __init__(self):
        require_device()
        self.L = _lib.lib()
        self.h = self.L.msplat_engine_create()
        if not self.h:
            raise RuntimeError("msplat_engine_create failed (no usable HIP device)")

    def close(self):
        if self.h:
            self.L.msplat_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train_step(self, st: DeviceState, cam, gt: DeviceArray, hp, debug=False, mask: DeviceArray = None):
        """One step; mask: optional device uint8 [H, W] training mask (below 128: ignored)."""
        if mask is None:
            rc = self.L.msplat_engine_train_step(self.h, st.struct(), cam, gt.ptr, hp, 1 if debug else 0)
        else:
            rc = self.L.msplat_engine_train_step_masked(self.h, st.struct(), cam, gt.ptr, mask.ptr, hp,
                                                        1 if debug else 0)
        if rc != 0:
            raise RuntimeError(f"msplat_engine_train_step failed ({rc})")

    def render(self, st: DeviceState, cam, degrees_to_use, bg):
        bgarr = (C.c_float * 3)(*bg)
        rc = self.L.msplat_engine_render(self.h, st.struct(), cam, degrees_to_use, bgarr)
        if rc != 0:
            raise RuntimeError(f"msplat_engine_render failed ({rc})")

    def sync(self):
        if self.L.msplat_engine_sync(self.h) != 0:
            raise RuntimeError("HIP error on the engine stream")

    def read(self, name: str, dtype, shape) -> np.ndarray:
        out = np.zeros(shape, dtype)
        n = self.L.msplat_engine_read(self.h, name.encode(), out.ctypes.data, out.nbytes)
        if n < 0:
            raise KeyError(name)
        return out

    def stats(self) -> dict:
        a = (C.c_int64 * 6)()
        self.L.msplat_engine_stats(self.h, a)
        return {"num_isect": a[0], "num_padded": a[1], "overflow": a[2], "max_tile": a[3], "k_max": a[4],
                "num_tiles": a[5]}

    def loss(self) -> float:
        return self.L.msplat_engine_loss(self.h)

    def projection(self, N: int) -> dict:
        A = self.read("recA", np.float32, (N, 4))
        B = self.read("recB", np.float32, (N, 4))
        Cc = self.read("recC", np.float32, (N, 2))
        return {
            "radii": self.read("radii", np.int32, (N,)),
            "tiles_hit": self.read("tiles_hit", np.int32, (N,)),
            "depths": self.read("depths", np.float32, (N,)),
            "aabb": self.read("aabb", np.float32, (N, 2)),
            "xys": A[:, :2].copy(),
            "opac_sig": A[:, 2].copy(),
            "conics": np.stack([A[:, 3], B[:, 0], B[:, 1]], 1),
            "colors": np.stack([B[:, 2], B[:, 3], Cc[:, 0]], 1),
        }

    def lists(self, num_tiles: int, num_isect: int) -> dict:
        return {
            "tile_counts": self.read("tile_counts", np.int32, (num_tiles,)),
            "tile_bins": self.read("tile_bins", np.int32, (num_tiles, 2)),
            "ids": self.read("ids", np.int32, (num_isect,)),
        }

    def raster(self, W: int, H: int) -> dict:
        return {
            "img": self.read("img", np.float32, (H, W, 3)),
            "final_T": self.read("final_T", np.float32, (H, W)),
            "final_idx": self.read("final_idx", np.int32, (H, W)),
        }

    def grads(self, N: int, W: int, H: int) -> dict:
        g2d = self.read("grad2d", np.float32, (N, G2D))
        return {
            "v_rendered": self.read("v_rendered", np.float32, (H, W, 3)),
            "v_xy": g2d[:, G2D_XY].copy(),
            "v_conic": g2d[:, G2D_CONIC].copy(),
            "v_rgb": g2d[:, G2D_RGB].copy(),
            "v_mean": self.read("dbg_v_mean", np.float32, (N, 3)),
            "v_scale": self.read("dbg_v_scale", np.float32, (N, 3)),
            "v_quat": self.read("dbg_v_quat", np.float32, (N, 4)),
            "v_opac": self.read("dbg_v_opac", np.float32, (N,)),
        }

    def nan_counts(self) -> np.ndarray:
        return self.read("nan_counts", np.uint32, (16,))

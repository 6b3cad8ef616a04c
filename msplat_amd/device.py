"""Device-resident training state and the kernel-level engine, driven from Python.

PyTorch is used only to own device memory (tensors on cuda:0 under ROCm); every kernel
that runs is one of the library's hand-written HIP kernels, called through the C ABI of
include/msplat_kernels.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

GROUPS = ("means", "scales", "quats", "f_dc", "f_rest", "opac")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("msplat_amd: no HIP device is available")
    return torch


class DeviceState:
    """Parameters + Adam moments on the GPU (cap-sized rows, first N active)."""

    def __init__(self, g: dict, capacity: int | None = None, moments=None):
        torch = _torch()
        n = g["means"].shape[0]
        self.N = n
        self.degree = int(g["degree"])
        cap = capacity or n
        self.cap = cap
        dev = torch.device("cuda:0")

        def up(a):
            a = np.ascontiguousarray(a, np.float32)
            t = torch.zeros((cap,) + a.shape[1:], dtype=torch.float32, device=dev)
            if n:
                t[:n] = torch.from_numpy(a).to(dev)
            return t
        self.p = {k: up(g[k]) for k in GROUPS}
        if moments is None:
            self.m = {k: torch.zeros_like(self.p[k]) for k in GROUPS}
            self.v = {k: torch.zeros_like(self.p[k]) for k in GROUPS}
        else:
            self.m = {k: up(moments[0][k]) for k in GROUPS}
            self.v = {k: up(moments[1][k]) for k in GROUPS}

    def struct(self) -> _lib.MsplatDeviceState:
        s = _lib.MsplatDeviceState()
        s.N, s.degree = self.N, self.degree
        s.means, s.scales, s.quats = self.p["means"].data_ptr(), self.p["scales"].data_ptr(), self.p["quats"].data_ptr()
        s.f_dc, s.f_rest, s.opac = self.p["f_dc"].data_ptr(), self.p["f_rest"].data_ptr(), self.p["opac"].data_ptr()
        for gi, k in enumerate(GROUPS):
            s.m[gi] = self.m[k].data_ptr()
            s.v[gi] = self.v[k].data_ptr()
        return s

    def to_host(self) -> dict:
        _torch().cuda.synchronize()
        out = {k: self.p[k][: self.N].cpu().numpy() for k in GROUPS}
        out["degree"] = self.degree
        return out

    def moments_host(self):
        _torch().cuda.synchronize()
        return ({k: self.m[k][: self.N].cpu().numpy() for k in GROUPS},
                {k: self.v[k][: self.N].cpu().numpy() for k in GROUPS})


class Engine:
    """Owner of one MsplatEngine (kernel-level ABI)."""

    def __init__(self):
        _torch()
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

    def train_step(self, st: DeviceState, cam, gt, hp, debug=False, sync_inputs=True):
        if sync_inputs:
            _torch().cuda.synchronize()  # inputs written on torch's stream
        s = st.struct()
        rc = self.L.msplat_engine_train_step(self.h, s, cam, gt.data_ptr(), hp, 1 if debug else 0)
        if rc != 0:
            raise RuntimeError(f"msplat_engine_train_step failed ({rc})")

    def render(self, st: DeviceState, cam, degrees_to_use, bg):
        _torch().cuda.synchronize()
        s = st.struct()
        bgarr = (C.c_float * 3)(*bg)
        rc = self.L.msplat_engine_render(self.h, s, cam, degrees_to_use, bgarr)
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
        pkC = self.read("pkC", np.float32, (num_isect, 2))
        return {
            "tile_counts": self.read("tile_counts", np.int32, (num_tiles,)),
            "tile_bins": self.read("tile_bins", np.int32, (num_tiles, 2)),
            "ids": pkC[:, 1].view(np.int32).copy(),
        }

    def raster(self, W: int, H: int) -> dict:
        return {
            "img": self.read("img", np.float32, (H, W, 3)),
            "final_T": self.read("final_T", np.float32, (H, W)),
            "final_idx": self.read("final_idx", np.int32, (H, W)),
        }

    def grads(self, N: int, W: int, H: int) -> dict:
        return {
            "v_rendered": self.read("v_rendered", np.float32, (H, W, 3)),
            "v_xy": self.read("v_xy", np.float32, (N, 2)),
            "v_conic": self.read("v_conic", np.float32, (N, 3)),
            "v_rgb": self.read("v_rgb", np.float32, (N, 3)),
            "v_mean": self.read("dbg_v_mean", np.float32, (N, 3)),
            "v_scale": self.read("dbg_v_scale", np.float32, (N, 3)),
            "v_quat": self.read("dbg_v_quat", np.float32, (N, 4)),
            "v_opac": self.read("dbg_v_opac", np.float32, (N,)),
        }

    def nan_counts(self) -> np.ndarray:
        return self.read("nan_counts", np.uint32, (16,))

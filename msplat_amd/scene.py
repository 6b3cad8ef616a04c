"""Synthetic scenes and step hyper-parameters (host side, numpy).

Thin wrappers over the library's seeded generators (scene.cpp) plus the scalar bookkeeping
of Model::fullIteration (reference core/src/model.cpp:913-958) and Model::schedulersStep
(model.cpp:256-261), used by the Python API, the tests and the benchmark.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from ._lib import fptr

# Adam learning rates per group: means, scales, quats, f_dc, f_rest, opacity (model.cpp:223)
LR_INIT = (0.00016, 0.005, 0.001, 0.0025, 0.000125, 0.05)
MEANS_LR_FINAL = 0.0000016
BETA1, BETA2, EPS = 0.9, 0.999, 1e-8
DEFAULT_BG = (0.6130, 0.0101, 0.3984)


def num_rest(degree: int) -> int:
    return (degree + 1) * (degree + 1) - 1


def synth_gaussians(kind: int, seed: int, n: int, degree: int = 3) -> dict:
    L = _lib.lib()
    nr = num_rest(degree)
    g = {
        "means": np.zeros((n, 3), np.float32),
        "scales": np.zeros((n, 3), np.float32),
        "quats": np.zeros((n, 4), np.float32),
        "opac": np.zeros((n,), np.float32),
        "f_dc": np.zeros((n, 3), np.float32),
        "f_rest": np.zeros((n, nr, 3), np.float32),
    }
    rc = L.msplat_synth_gaussians(kind, seed, n, degree, fptr(g["means"]), fptr(g["scales"]), fptr(g["quats"]),
                                  fptr(g["opac"]), fptr(g["f_dc"]), fptr(g["f_rest"]))
    if rc != 0:
        raise ValueError("bad synthetic Gaussian arguments")
    g["degree"] = degree
    return g


def synth_cameras(seed: int, count: int, radius: float = 3.0, jitter: float = 0.0) -> np.ndarray:
    L = _lib.lib()
    c2w = np.zeros((count, 16), np.float32)
    if L.msplat_synth_cameras(seed, count, radius, jitter, fptr(c2w)) != 0:
        raise ValueError("bad camera arguments")
    return c2w.reshape(count, 4, 4)


def prepare_camera(c2w, fx, fy, cx, cy, width, height) -> _lib.MsplatCamera:
    L = _lib.lib()
    c = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
    cam = _lib.MsplatCamera()
    L.msplat_camera_prepare(fptr(c), fx, fy, cx, cy, width, height, cam)
    return cam


def quantize(img: np.ndarray) -> np.ndarray:
    L = _lib.lib()
    src = np.ascontiguousarray(img, np.float32)
    out = np.zeros(src.shape, np.uint8)
    L.msplat_quantize_rgb(fptr(src), src.size, out.ctypes.data_as(_lib.C.POINTER(_lib.C.c_uint8)))
    return out


def means_lr(step: int, max_steps: int = 30000) -> float:
    """Position LR after schedulersStep(step) (model.cpp:256-261), float arithmetic."""
    t = np.float32(min(max(np.float32(step) / np.float32(max_steps), 0.0), 1.0))
    a = np.float32(math.log(np.float32(LR_INIT[0]))) * (np.float32(1.0) - t)
    b = np.float32(math.log(np.float32(MEANS_LR_FINAL))) * t
    return float(np.float32(math.exp(np.float32(a + b))))


def step_params(adam_step: int, lrs, width: int, height: int, step: int, degree: int, degrees_to_use: int,
                ssim_weight: float = 0.2, opacity_reg: float = 0.01, scale_reg: float = 0.01,
                noise_lr: float = 5e5, cull_radius: float = 3.0, bg=DEFAULT_BG, sgld: bool = True,
                reg_invis_scale: float = 1.0) -> _lib.MsplatStepParams:
    """The scalars Model::fullIteration hands to msplat_train_step (model.cpp:913-958, 971)."""
    hp = _lib.MsplatStepParams()
    f = np.float32
    bc1 = f(1.0) - f(f(BETA1) ** f(adam_step))
    bc2 = f(1.0) - f(f(BETA2) ** f(adam_step))
    for gi in range(6):
        hp.step_size[gi] = float(f(f(lrs[gi]) / bc1))
        hp.bc2_sqrt[gi] = float(f(np.sqrt(bc2)))
    hp.beta1, hp.beta2, hp.eps = BETA1, BETA2, EPS
    hp.ssim_weight = ssim_weight
    hp.loss_inv_n = float(f(1.0) / f(height * width * 3))
    hp.op_reg = opacity_reg if step >= 500 else 0.0
    hp.sc_reg = scale_reg if step >= 500 else 0.0
    hp.reg_invis_scale = reg_invis_scale
    for c in range(3):
        hp.background[c] = bg[c]
    hp.noise_lr = noise_lr
    hp.xyz_lr = lrs[0]
    hp.max_disp = float(f(cull_radius) * f(0.0167)) if cull_radius > 0 else 0.05
    hp.sgld_seed = step
    hp.sgld_enable = 1 if sgld else 0
    hp.degree = degree
    hp.degrees_to_use = degrees_to_use
    return hp

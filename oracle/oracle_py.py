"""ctypes binding of the CPU restatement (TEST INFRASTRUCTURE).

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def _state_struct(ft):
    P = C.POINTER(ft)

    class S(C.Structure):
        _fields_ = [("N", C.c_int32), ("degree", C.c_int32), ("means", P), ("scales", P), ("quats", P),
                    ("f_dc", P), ("f_rest", P), ("opac", P), ("m", P * 6), ("v", P * 6), ("sh_bf16", C.c_int32)]
    return S


def _out_struct(ft):
    P = C.POINTER(ft)
    I = C.POINTER(C.c_int32)

    class O(C.Structure):
        _fields_ = [("xys", P), ("depths", P), ("conics", P), ("colors", P), ("aabb", P),
                    ("radii", I), ("tiles_hit", I), ("tile_counts", I), ("tile_bins", I), ("ids", I),
                    ("ids_cap", C.c_int64), ("num_isect", C.c_int64), ("num_padded", C.c_int64),
                    ("overflow", C.c_int32), ("k_max", C.c_uint32),
                    ("img", P), ("final_T", P), ("final_idx", I), ("v_rendered", P), ("loss_sum", C.c_double),
                    ("v_xy", P), ("v_conic", P), ("v_rgb", P), ("v_opac", P), ("v_mean", P), ("v_scale", P),
                    ("v_quat", P), ("nan_adam", C.c_uint32 * 6)]
    return O


StateF32, StateF64 = _state_struct(C.c_float), _state_struct(C.c_double)
OutF32, OutF64 = _out_struct(C.c_float), _out_struct(C.c_double)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        from msplat_amd._lib import MsplatCamera, MsplatStepParams  # struct layouts only
        L = C.CDLL(LIB_PATH)
        L.oracle_train_step_f32.argtypes = [C.POINTER(StateF32), C.POINTER(MsplatCamera), C.c_void_p,
                                            C.POINTER(MsplatStepParams), C.POINTER(OutF32), C.c_void_p]
        L.oracle_train_step_f64.argtypes = [C.POINTER(StateF64), C.POINTER(MsplatCamera), C.c_void_p,
                                            C.POINTER(MsplatStepParams), C.POINTER(OutF64), C.c_void_p]
        L.oracle_render_f32.argtypes = [C.POINTER(StateF32), C.POINTER(MsplatCamera), C.c_int,
                                        C.POINTER(C.c_float), C.POINTER(OutF32)]
        L.oracle_render_f64.argtypes = [C.POINTER(StateF64), C.POINTER(MsplatCamera), C.c_int,
                                        C.POINTER(C.c_float), C.POINTER(OutF64)]
        L.oracle_ssim_loss_f64.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                           C.c_void_p]
        L.oracle_ssim_loss_f64.restype = C.c_double
        L.oracle_ssim_loss_f32.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                           C.c_void_p]
        L.oracle_ssim_loss_f32.restype = C.c_double
        L.oracle_ssim_loss_masked_f64.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                                  C.c_double, C.c_void_p]
        L.oracle_ssim_loss_masked_f64.restype = C.c_double
        L.oracle_ssim_loss_masked_f32.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                                  C.c_float, C.c_void_p]
        L.oracle_ssim_loss_masked_f32.restype = C.c_double
        L.oracle_sgld_noise_f32.argtypes = [C.c_int, C.c_uint32, C.c_void_p]
        L.oracle_pcg_hash32.argtypes = [C.c_uint32, C.c_uint32]
        L.oracle_pcg_hash32.restype = C.c_uint32
        L.oracle_ms_expf.argtypes = [C.c_float]
        L.oracle_ms_expf.restype = C.c_float
        L.oracle_ms_logf.argtypes = [C.c_float]
        L.oracle_ms_logf.restype = C.c_float
        L.oracle_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.oracle_mlp_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p]
        L.oracle_mlp_loss_exact.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.oracle_mlp_loss_exact.restype = C.c_double
        L.oracle_ms_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.oracle_k_max.argtypes = [C.c_int64, C.c_int]
        L.oracle_k_max.restype = C.c_uint32
        L.oracle_eq9.argtypes = [C.c_float, C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_void_p]
        L.oracle_mcmc_step.argtypes = [C.POINTER(StateF32), C.c_int32, C.c_float, C.c_uint32, C.c_void_p]
        L.oracle_mcmc_step_learned.argtypes = [C.POINTER(StateF32), C.c_int32, C.c_float, C.c_uint32, C.c_void_p,
                                               C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_threads.restype = C.c_int
        vp = C.c_void_p
        L.oracle_raster2d_f64.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int,
                                          vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_raster2d_f64.restype = C.c_int
        L.oracle_project_f64.argtypes = [C.c_int, vp, vp, vp, C.POINTER(MsplatCamera), vp, vp, vp, vp]
        L.oracle_project_vjp_f64.argtypes = [C.c_int, vp, vp, vp, C.POINTER(MsplatCamera), vp, vp, vp, vp, vp, vp]
        L.oracle_sh_color_f64.argtypes = [C.c_int, vp, vp, vp, vp]
        L.oracle_ssim_window.argtypes = [vp]
        L.oracle_train_step_batch_f32.argtypes = [C.POINTER(StateF32), vp, vp, vp, C.c_int, vp, vp, vp]
        L.oracle_train_step_batch_f32.restype = C.c_double
        L.oracle_train_step_batch_f64.argtypes = [C.POINTER(StateF64), vp, vp, vp, C.c_int, vp, vp, vp]
        L.oracle_train_step_batch_f64.restype = C.c_double
        L.oracle_eval_sums.argtypes = [C.c_int, C.c_int, vp, vp, vp]
        _lib = L
    return _lib


GROUP_SHAPES = ("means", "scales", "quats", "f_dc", "f_rest", "opac")


class HostState:
    """Host copy of the trainable state (params + Adam moments), numpy arrays of a given dtype."""

    def __init__(self, g: dict, dtype=np.float32, moments=None, capacity=None, sh_bf16=False):
        self.dtype = dtype
        self.sh_bf16 = bool(sh_bf16)
        self.degree = int(g["degree"])
        n = g["means"].shape[0]
        self.N = n
        cap = capacity or n
        self.cap = cap

        def alloc(a):
            shape = (cap,) + a.shape[1:]
            out = np.zeros(shape, dtype)
            out[:n] = a
            return out
        self.p = {k: alloc(np.asarray(g[k])) for k in GROUP_SHAPES}
        if moments is None:
            self.m = {k: np.zeros_like(self.p[k]) for k in GROUP_SHAPES}
            self.v = {k: np.zeros_like(self.p[k]) for k in GROUP_SHAPES}
        else:
            self.m = {k: alloc(np.asarray(moments[0][k])) for k in GROUP_SHAPES}
            self.v = {k: alloc(np.asarray(moments[1][k])) for k in GROUP_SHAPES}

    def struct(self):
        S = StateF32 if self.dtype == np.float32 else StateF64
        ft = C.c_float if self.dtype == np.float32 else C.c_double
        s = S()
        s.N, s.degree = self.N, self.degree
        s.sh_bf16 = 1 if self.sh_bf16 else 0

        def ptr(a):
            return a.ctypes.data_as(C.POINTER(ft))
        s.means, s.scales, s.quats = ptr(self.p["means"]), ptr(self.p["scales"]), ptr(self.p["quats"])
        s.f_dc, s.f_rest, s.opac = ptr(self.p["f_dc"]), ptr(self.p["f_rest"]), ptr(self.p["opac"])
        for gi, k in enumerate(GROUP_SHAPES):
            s.m[gi] = ptr(self.m[k])
            s.v[gi] = ptr(self.v[k])
        return s

    def view(self):
        return {k: self.p[k][: self.N] for k in GROUP_SHAPES}


def _run(fn_name, st: HostState, cam, gt, hp, want_ids=True, render_args=None, ids_cap=None, mask=None):
    L = lib()
    f64 = st.dtype == np.float64
    O = OutF64 if f64 else OutF32
    ft = C.c_float if not f64 else C.c_double
    N = st.N
    W, H = cam.width, cam.height
    T = ((W + 15) // 16) * ((H + 15) // 16)
    out = {
        "xys": np.zeros((N, 2), st.dtype), "depths": np.zeros(N, st.dtype), "conics": np.zeros((N, 3), st.dtype),
        "colors": np.zeros((N, 3), st.dtype), "aabb": np.zeros((N, 2), st.dtype),
        "radii": np.zeros(N, np.int32), "tiles_hit": np.zeros(N, np.int32),
        "tile_counts": np.zeros(T, np.int32), "tile_bins": np.zeros((T, 2), np.int32),
        "img": np.zeros((H, W, 3), st.dtype), "final_T": np.zeros((H, W), st.dtype),
        "final_idx": np.zeros((H, W), np.int32),
    }
    train = render_args is None
    if train:
        out.update({"v_rendered": np.zeros((H, W, 3), st.dtype), "v_xy": np.zeros((N, 2), st.dtype),
                    "v_conic": np.zeros((N, 3), st.dtype), "v_rgb": np.zeros((N, 3), st.dtype),
                    "v_opac": np.zeros(N, st.dtype), "v_mean": np.zeros((N, 3), st.dtype),
                    "v_scale": np.zeros((N, 3), st.dtype), "v_quat": np.zeros((N, 4), st.dtype)})
    # the lists are copied out only when they fit (48 N, the list capacity, by default)
    ids_cap = (int(N) * 48 if ids_cap is None else int(ids_cap)) if want_ids else 0
    ids = np.zeros(max(ids_cap, 1), np.int32)
    o = O()
    for k, a in out.items():
        if a.dtype == np.int32:
            setattr(o, k, a.ctypes.data_as(C.POINTER(C.c_int32)))
        else:
            setattr(o, k, a.ctypes.data_as(C.POINTER(ft)))
    o.ids = ids.ctypes.data_as(C.POINTER(C.c_int32)) if want_ids else None
    o.ids_cap = ids_cap
    s = st.struct()
    if train:
        getattr(L, fn_name)(s, cam, gt.ctypes.data, hp, o, None if mask is None else mask.ctypes.data)
    else:
        deg_use, bg = render_args
        bgarr = (C.c_float * 3)(*bg)
        getattr(L, fn_name)(s, cam, deg_use, bgarr, o)
    out["num_isect"] = o.num_isect
    out["num_padded"] = o.num_padded
    out["overflow"] = o.overflow
    out["k_max"] = o.k_max
    if want_ids:
        out["ids"] = ids[: o.num_isect].copy()
    if train:
        out["loss_sum"] = o.loss_sum
        out["loss"] = o.loss_sum / (W * H)
        out["nan_adam"] = list(o.nan_adam)
    return out


def train_step(st: HostState, cam, gt: np.ndarray, hp, want_ids=True, mask=None):
    """One training step on the host state (updated in place); returns the intermediates.
    mask: optional uint8 [H, W] training mask (pixels below 128 own no loss)."""
    gt = np.ascontiguousarray(gt, np.uint8)
    mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    return _run("oracle_train_step_f64" if st.dtype == np.float64 else "oracle_train_step_f32", st, cam, gt, hp,
                want_ids, mask=mask)


def render(st: HostState, cam, degrees_to_use: int, bg, want_ids=True, ids_cap=None):
    return _run("oracle_render_f64" if st.dtype == np.float64 else "oracle_render_f32", st, cam, None, None,
                want_ids, render_args=(degrees_to_use, bg), ids_cap=ids_cap)


def train_step_batch(st: HostState, cams, gts, hps, want_images=False, masks=None):
    """One multi-view batched step on the host state (updated in place): per-view cameras,
    uint8 ground truths, step scalars and optional training masks (uint8 [H, W] or None per
    view). Returns (mean per-view loss, per-view images)."""
    from msplat_amd._lib import MsplatCamera, MsplatStepParams
    L = lib()
    B = len(cams)
    f64 = st.dtype == np.float64
    cam_arr = (MsplatCamera * B)(*cams)
    hp_arr = (MsplatStepParams * B)(*hps)
    gts = [np.ascontiguousarray(g, np.uint8) for g in gts]
    gt_ptrs = (C.c_void_p * B)(*[g.ctypes.data for g in gts])
    imgs = [np.zeros((c.height, c.width, 3), st.dtype) for c in cams] if want_images else None
    img_ptrs = (C.c_void_p * B)(*[a.ctypes.data for a in imgs]) if want_images else None
    fn = L.oracle_train_step_batch_f64 if f64 else L.oracle_train_step_batch_f32
    mk = None
    if masks is not None:
        masks = [None if m is None else np.ascontiguousarray(m, np.uint8) for m in masks]
        mk = (C.c_void_p * B)(*[None if m is None else m.ctypes.data for m in masks])
    loss = fn(st.struct(), C.addressof(cam_arr), C.addressof(gt_ptrs), C.addressof(hp_arr), B,
              C.addressof(img_ptrs) if want_images else None, None, C.addressof(mk) if mk is not None else None)
    return loss, imgs


def eval_sums(img: np.ndarray, gt: np.ndarray) -> np.ndarray:
    """{sum (r-g)^2, sum |r-g|, sum SSIM} of a float render [H, W, 3] against uint8 ground truth."""
    img = np.ascontiguousarray(img, np.float32)
    gt = np.ascontiguousarray(gt, np.uint8)
    out = np.zeros(3, np.float64)
    lib().oracle_eval_sums(img.shape[0], img.shape[1], img.ctypes.data, gt.ctypes.data, out.ctypes.data)
    return out


def eval_metrics(img: np.ndarray, gt: np.ndarray) -> dict:
    """psnr / ssim / l1 of one view as Trainer::evaluate computes them (float results)."""
    s = eval_sums(img, gt)
    n = float(img.size)
    return {"psnr": float(np.float32(10.0) * np.float32(np.log10(1.0 / (s[0] / n)))),
            "ssim": float(np.float32(s[2] / n)), "l1": float(np.float32(s[1] / n))}


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(n)


# ---- learned relocation proposal (restatement of this project's own network) ----------
MLP_NPARAM = 5508
MLP_BLOCKS = {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248), "W3": (5248, 5504),
              "b3": (5504, 5508)}


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-initialised parameter vector (biases zero) as the trainer seeds it."""
    rng = np.random.default_rng(seed)
    th = np.zeros(MLP_NPARAM, np.float32)
    for name, fan_in in (("W1", 16), ("W2", 64), ("W3", 64)):
        a, b = MLP_BLOCKS[name]
        th[a:b] = (rng.standard_normal(b - a) * np.sqrt(2.0 / fan_in)).astype(np.float32)
    return th


def mlp_forward(theta: np.ndarray, X: np.ndarray, exact: bool = False) -> np.ndarray:
    """[N 4] outputs; X is [N 16] uint16 (bf16 bit patterns)."""
    theta = np.ascontiguousarray(theta, np.float32)
    X = np.ascontiguousarray(X, np.uint16)
    out = np.zeros((X.shape[0], 4), np.float64)
    lib().oracle_mlp_forward(theta.ctypes.data, X.ctypes.data, X.shape[0], int(exact), out.ctypes.data)
    return out


def mlp_backward(theta: np.ndarray, X: np.ndarray, tgt: np.ndarray, exact: bool = False):
    """(d loss / d theta [5508], (logit loss, offset loss))."""
    theta = np.ascontiguousarray(theta, np.float32)
    X = np.ascontiguousarray(X, np.uint16)
    tgt = np.ascontiguousarray(tgt, np.float32)
    grad = np.zeros(MLP_NPARAM, np.float64)
    loss = np.zeros(2, np.float64)
    lib().oracle_mlp_backward(theta.ctypes.data, X.ctypes.data, tgt.ctypes.data, X.shape[0], int(exact),
                              grad.ctypes.data, loss.ctypes.data)
    return grad, (float(loss[0]), float(loss[1]))


def mlp_loss_exact(theta64: np.ndarray, X: np.ndarray, tgt: np.ndarray) -> float:
    theta64 = np.ascontiguousarray(theta64, np.float64)
    X = np.ascontiguousarray(X, np.uint16)
    tgt = np.ascontiguousarray(tgt, np.float32)
    return float(lib().oracle_mlp_loss_exact(theta64.ctypes.data, X.ctypes.data, tgt.ctypes.data, X.shape[0]))


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns, round to nearest even (the device convention)."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
    return r

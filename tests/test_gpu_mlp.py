"""The learned-proposal network's matrix-core kernels (msplat_mlp_* in msplat_kernels.h)
against the oracle's device-rounding restatement: outputs within 1e-2 of max(1, |output|)
and each parameter block's gradient within 3e-2 in relative norm (bf16 operands: the
oracle and the device round the same values, but fp32 sums in a different order can move
a hidden activation across a bf16 rounding boundary). Ragged batch sizes, an empty batch,
and the exact (double) error of the device result are covered."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from msplat_amd import _lib

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr)
        L = _lib.lib()
        self.p = L.msplat_dev_alloc(max(self.arr.nbytes, 16))
        assert self.p
        assert L.msplat_dev_upload(self.p, self.arr.ctypes.data, self.arr.nbytes) == 0

    def get(self):
        out = np.empty_like(self.arr)
        assert _lib.lib().msplat_dev_download(out.ctypes.data, self.p, out.nbytes) == 0
        return out

    def __del__(self):
        _lib.lib().msplat_dev_free(self.p)


def _bind():
    L = _lib.lib()
    L.msplat_mlp_packed_bytes.restype = C.c_size_t
    L.msplat_mlp_pack.argtypes = [C.c_void_p, C.c_void_p]
    L.msplat_mlp_forward.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.msplat_mlp_backward.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _case(n, seed):
    rng = np.random.default_rng(seed)
    th = O.mlp_init(seed)
    th[1024:1088] = rng.normal(0, 0.1, 64)
    th[5184:5248] = rng.normal(0, 0.1, 64)
    th[5504:5508] = rng.normal(0, 0.1, 4)
    X = O.bf16_bits(rng.normal(0, 1, (n, 16)).astype(np.float32))
    tgt = np.zeros((n, 8), np.float32)
    if n:
        tgt[:, 0] = rng.integers(0, 2, n)
        tgt[:, 1] = 1.0 / n
        tgt[:, 2:5] = rng.normal(0, 1, (n, 3))
        tgt[:, 5] = 0.25 / n
    return th, X, tgt


@pytest.mark.parametrize("n", [0, 1, 37, 1000, 4099])
def test_mlp_kernels_match_oracle(n):
    L = _bind()
    th, X, tgt = _case(n, 5 + n)
    d_th = Dev(th)
    d_pk = Dev(np.zeros(L.msplat_mlp_packed_bytes(), np.uint8))
    assert L.msplat_mlp_pack(d_th.p, d_pk.p) == 0
    d_x = Dev(X if n else np.zeros((1, 16), np.uint16))
    d_out = Dev(np.zeros((max(n, 1), 4), np.float32))
    assert L.msplat_mlp_forward(n, d_x.p, d_pk.p, d_out.p) == 0
    d_t = Dev(tgt if n else np.zeros((1, 8), np.float32))
    d_g = Dev(np.zeros(O.MLP_NPARAM, np.float32))
    d_l = Dev(np.zeros(2, np.float64))
    assert L.msplat_mlp_backward(n, d_x.p, d_t.p, d_pk.p, d_g.p, d_l.p) == 0
    g = d_g.get()
    loss = d_l.get()
    if n == 0:
        assert not g.any() and not loss.any()
        return
    out = d_out.get()
    ref = O.mlp_forward(th, X, exact=False)
    scale = max(1.0, float(np.abs(ref).max()))
    assert float(np.abs(out - ref).max()) <= 1e-2 * scale
    exact = O.mlp_forward(th, X, exact=True)
    print(f"n={n}: device vs exact (double) max |d out| = {float(np.abs(out - exact).max()):.3e}")
    gref, lref = O.mlp_backward(th, X, tgt, exact=False)
    for name, (a, b) in O.MLP_BLOCKS.items():
        rel = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-12)
        assert rel <= 3e-2, (name, rel)
    assert abs(loss[0] - lref[0]) <= 1e-2 * abs(lref[0]) + 1e-6
    assert abs(loss[1] - lref[1]) <= 1e-2 * abs(lref[1]) + 1e-6

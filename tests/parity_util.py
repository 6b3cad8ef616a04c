"""Helpers shared by the GPU parity tests: scene set-up and comparisons against the oracle.

Comparisons are per element: |got - want| <= rtol * |want| + floor, where `want` is the FP32
oracle and the absolute floor is FLOOR_K times that output's FP64 error, i.e. the largest
|FP32 oracle - FP64 oracle| over the array (the rounding noise any FP32 evaluation of the
output carries; float atomics sum in another order than the oracle's loops, so the device
cannot be closer to the FP32 oracle than that noise). Each comparison records the FP64 error
of the device result and of the FP32 oracle (max |x - FP64| / max |FP64|); the session prints
them at the end.
"""
from __future__ import annotations

import os

import numpy as np

from msplat_amd import scene

FLOOR_K = 4.0  # absolute floor = 4 x the output's FP64 error
# (test, output, n, device FP64 err, oracle FP64 err, worst |d| / (rtol |w| + floor),
#  max |device - oracle32| / max |oracle32|, fraction of elements beyond rtol |w| without the floor)
REPORT: list = []


def small_scene(n=3000, n_hidden=1500, W=160, H=144, degree=3, kind=0, seed=11, fx=None):
    g = scene.synth_gaussians(kind, seed, n, degree)
    hidden = scene.synth_gaussians(kind, seed + 1000, n_hidden, degree)
    c2w = scene.synth_cameras(seed + 7, 4, 3.0, 0.0)
    fx = fx or float(W)
    cams = [scene.prepare_camera(c, fx, fx, W / 2.0, H / 2.0, W, H) for c in c2w]
    return g, hidden, cams


def rel_err(a, b):
    """max |a - b| / max |b| (a summary figure; the assertions are per element)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.abs(b).max(), 1e-30) if b.size else 1.0
    return float(np.abs(a - b).max() / den) if b.size else 0.0


def assert_exact(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{name}: shape {a.shape} vs {b.shape}"
    bad = np.flatnonzero(a.reshape(-1) != b.reshape(-1))
    assert bad.size == 0, f"{name}: {bad.size} of {a.size} differ; first at {bad[:5]}: {a.reshape(-1)[bad[:5]]} vs {b.reshape(-1)[bad[:5]]}"


def _test_name():
    t = os.environ.get("PYTEST_CURRENT_TEST", "")
    return t.split(" ")[0].split("::")[-1] if t else ""


def assert_close(name, got, want, want64=None, rtol=1e-4, atol=0.0, k=FLOOR_K, ref32=None):
    """Per-element |got - want| <= rtol |want| + max(atol, k * e64), e64 = max |want - want64|.
    With `ref32` (an FP32 oracle result of the same output in another storage mode, e.g. the
    all-fp32 run beside a bf16-storage one) e64 = max |ref32 - want64| instead, so that the
    floor stays the arithmetic's rounding noise while the FP64 error reported is the
    device's against the exact result. Returns the device's FP64 error (or its error against
    `want` without an FP64 result)."""
    got = np.asarray(got, np.float64).reshape(-1)
    want = np.asarray(want, np.float64).reshape(-1)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    if want.size == 0:
        return 0.0
    assert (np.isfinite(got) == np.isfinite(want)).all(), f"{name}: non-finite values at other places"
    fin = np.isfinite(want)
    got, want = got[fin], want[fin]
    floor = float(atol)
    dev64 = orc64 = float("nan")
    if want64 is not None:
        w64 = np.asarray(want64, np.float64).reshape(-1)[fin]
        base = want if ref32 is None else np.asarray(ref32, np.float64).reshape(-1)[fin]
        e64 = float(np.abs(base - w64).max())
        floor = max(floor, k * e64)
        den = max(float(np.abs(w64).max()), 1e-30)
        dev64 = float(np.abs(got - w64).max()) / den
        orc64 = float(np.abs(want - w64).max()) / den
    err = np.abs(got - want)
    bound = rtol * np.abs(want) + floor
    ratio = err / np.maximum(bound, 1e-300)
    worst = int(np.argmax(ratio))
    den32 = max(float(np.abs(want).max()), 1e-30)
    REPORT.append((_test_name(), name, int(want.size), dev64, orc64, float(ratio[worst]), float(err.max()) / den32,
                   float(np.mean(err > rtol * np.abs(want)))))
    bad = err > bound
    assert not bad.any(), (f"{name}: {int(bad.sum())} of {want.size} elements beyond rtol {rtol:g} + floor {floor:.3e}; "
                           f"worst element {worst}: device {got[worst]!r} oracle {want[worst]!r}"
                           + (f" fp64 {float(np.asarray(want64, np.float64).reshape(-1)[fin][worst])!r}"
                              if want64 is not None else ""))
    return dev64 if want64 is not None else rel_err(got, want)


def assert_rel(name, a, b, tol):
    """Scalar (or whole-array summary) relative comparison, for losses and similar sums."""
    e = rel_err(a, b)
    assert e <= tol, f"{name}: relative error {e:.3e} > {tol:.1e}"
    return e


def report_lines():
    out = []
    for t, name, n, d64, o64, r, d32, frac in REPORT:
        out.append(f"{t:<40} {name:<22} n={n:<10d} dev-FP64 {d64:8.2e} orc32-FP64 {o64:8.2e} "
                   f"dev-orc32 {d32:8.2e} >rtol {frac:7.1e} worst/bound {r:5.2f}")
    return out


GROUPS = ("means", "scales", "quats", "f_dc", "f_rest", "opac")


def oracle_pair(g, moments, cam, gt, hp, sh_bf16=False, want_ids=True, fp32_ref=False, mask=None):
    """One oracle step from the same state in FP32 (in the given SH storage mode) and in FP64.
    Returns (state32, out32, state64, out64[, state32 all-fp32, out32 all-fp32])."""
    import oracle_py as O
    h32 = O.HostState(g, moments=moments, sh_bf16=sh_bf16)
    o32 = O.train_step(h32, cam, gt, hp, want_ids=want_ids, mask=mask)
    h64 = O.HostState(g, dtype=np.float64, moments=moments)
    o64 = O.train_step(h64, cam, gt, hp, want_ids=False, mask=mask)
    if not fp32_ref:
        return h32, o32, h64, o64
    hr = O.HostState(g, moments=moments)
    orr = O.train_step(hr, cam, gt, hp, want_ids=False, mask=mask)
    return h32, o32, h64, o64, hr, orr


def compare_state(tag, got, h32, h64, rtol=1e-4, bf16_groups=(), href=None):
    """Parameters and Adam moments after a step (got: {group, "m."+group, "v."+group} flat
    arrays of the active rows) against the oracle states. Moments of groups stored in bf16
    are held to one bf16 step (2^-7 relative) plus the floor."""
    n = h32.N
    for k in GROUPS:
        for pre, a32, a64, ar in (("", h32.p, h64.p, href.p if href else None),
                                  ("m.", h32.m, h64.m, href.m if href else None),
                                  ("v.", h32.v, h64.v, href.v if href else None)):
            r = 2.0 ** -7 if (pre and k in bf16_groups) else rtol
            assert_close(f"{pre}{k}{tag}", got[pre + k], a32[k][:n], a64[k][:n], rtol=r,
                         ref32=None if ar is None else ar[k][:n])

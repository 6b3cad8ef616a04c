"""Pins of the CPU oracle (the parity checker) against what the mathematics fixes:
closed forms, independent formulations, finite differences and sampling statistics.
Each test names the oracle function it pins; none re-types the oracle's formula."""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import ndimage, stats

import oracle_py as O
from msplat_amd import scene

HERE = os.path.dirname(os.path.abspath(__file__))
SSIM_C1, SSIM_C2 = 0.01 ** 2, 0.03 ** 2


def _p(a):
    return a.ctypes.data


def gauss_window():
    x = np.arange(11) - 5
    g = np.exp(-(x * x) / (2 * 1.5 ** 2))
    return g / g.sum()


# ---------------------------------------------------------------- SSIM (oracle ssim_loss)
def independent_ssim_loss(rend, gt_u8, w):
    """Zero-padded separable-Gaussian SSIM (pytorch-ssim formulation) + L1, FP64."""
    x = gt_u8.astype(np.float64) / 255.0
    y = rend.astype(np.float64)
    g = gauss_window()

    def blur(a):
        a = ndimage.correlate1d(a, g, axis=1, mode="constant", cval=0.0)
        return ndimage.correlate1d(a, g, axis=0, mode="constant", cval=0.0)
    ss = np.zeros(x.shape[:2])
    for c in range(3):
        mx, my = blur(x[..., c]), blur(y[..., c])
        sxx = np.maximum(blur(x[..., c] ** 2) - mx * mx, 0)
        syy = np.maximum(blur(y[..., c] ** 2) - my * my, 0)
        sxy = blur(x[..., c] * y[..., c]) - mx * my
        ss += ((2 * mx * my + SSIM_C1) * (2 * sxy + SSIM_C2)) / ((mx * mx + my * my + SSIM_C1) * (sxx + syy + SSIM_C2))
    l1 = np.abs(x - y).sum(-1)
    return float(np.mean(w * (1 - ss / 3) + (1 - w) * l1 / 3))


def _ssim_case(W=48, H=32, seed=0):
    rng = np.random.default_rng(seed)
    gt = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    base = gt.astype(np.float64) / 255.0
    rend = np.clip(base + rng.normal(0, 0.15, base.shape), 0.02, 0.98)
    return rend, gt


def test_ssim_loss_matches_independent_formulation():
    L = O.lib()
    rend, gt = _ssim_case()
    H, W = gt.shape[:2]
    v = np.zeros(rend.shape)
    loss = L.oracle_ssim_loss_f64(W, H, _p(np.ascontiguousarray(rend)), _p(gt), 0.2, 1.0 / (W * H * 3), _p(v))
    # window weights and C1, C2 are float constants in the reference: agreement to ~1e-8
    ref = independent_ssim_loss(rend, gt, 0.2)
    assert abs(loss / (W * H) - ref) < 1e-7 * ref


def test_ssim_gradient_finite_differences_and_border_term():
    L = O.lib()
    rend, gt = _ssim_case(seed=1)
    H, W = gt.shape[:2]
    w, inv_n = 0.2, 1.0 / (W * H * 3)
    v = np.zeros(rend.shape)
    L.oracle_ssim_loss_f64(W, H, _p(rend), _p(gt), w, inv_n, _p(v))
    g = gauss_window()
    rng = np.random.default_rng(2)
    eps = 1e-6
    for _ in range(40):
        yy, xx, c = rng.integers(0, H), rng.integers(0, W), rng.integers(0, 3)
        r1 = rend.copy(); r1[yy, xx, c] += eps
        r2 = rend.copy(); r2[yy, xx, c] -= eps
        fd = (L.oracle_ssim_loss_f64(W, H, _p(r1), _p(gt), w, inv_n, None)
              - L.oracle_ssim_loss_f64(W, H, _p(r2), _p(gt), w, inv_n, None)) / (2 * eps) / (W * H)
        # derivative fields computed from zeroed statistics at columns outside the image
        # (f2 = -2/C2, f3 = 2/C2) reach pixels within 5 columns of the side borders
        rows = sum(g[k] for k in range(11) if 0 <= yy - 5 + k < H)
        cols = sum(g[k] for k in range(11) if not (0 <= xx - 5 + k < W))
        x = gt[yy, xx, c] / 255.0
        extra = inv_n * (-w) * (2.0 / SSIM_C2) * (x - rend[yy, xx, c]) * rows * cols
        # float-valued window weights and C2 in the oracle: agreement to ~1e-6 of the term
        assert abs(v[yy, xx, c] - (fd + extra)) <= 2e-6 * abs(v[yy, xx, c]) + 1e-12


def independent_masked_loss_sum(rend, gt_u8, w, keep):
    """Sum over pixels of the masked pixel loss in FP64: a kept pixel adds
    w (1 - mean_c SSIM) + (1 - w) mean_c |x - y|, a masked one adds only w (the reference's
    loss value counts every pixel, msplat_metal.mm:1738, while the masked pixel's terms are
    multiplied by zero)."""
    x = gt_u8.astype(np.float64) / 255.0
    y = rend.astype(np.float64)
    g = gauss_window()

    def blur(a):
        a = ndimage.correlate1d(a, g, axis=1, mode="constant", cval=0.0)
        return ndimage.correlate1d(a, g, axis=0, mode="constant", cval=0.0)
    ss = np.zeros(x.shape[:2])
    for c in range(3):
        mx, my = blur(x[..., c]), blur(y[..., c])
        sxx = np.maximum(blur(x[..., c] ** 2) - mx * mx, 0)
        syy = np.maximum(blur(y[..., c] ** 2) - my * my, 0)
        sxy = blur(x[..., c] * y[..., c]) - mx * my
        ss += ((2 * mx * my + SSIM_C1) * (2 * sxy + SSIM_C2)) / ((mx * mx + my * my + SSIM_C1) * (sxx + syy + SSIM_C2))
    l1 = np.abs(x - y).sum(-1)
    k = keep.astype(np.float64)
    return float(np.sum(w + k * (-w * ss / 3 + (1 - w) * l1 / 3)))


def _mask_case(W, H, seed):
    """A mask with a masked block, scattered masked pixels and values either side of 128."""
    rng = np.random.default_rng(seed)
    m = np.where(rng.random((H, W)) < 0.2, rng.integers(0, 128, (H, W)), rng.integers(128, 256, (H, W)))
    m[H // 4: H // 2, W // 3: W // 2] = 127
    m[0, 0], m[0, 1] = 128, 127
    return m.astype(np.uint8)


def test_ssim_masked_loss_matches_independent_formulation():
    """oracle ssim_loss with a training mask (metal:3602-3653): masked pixels' terms are
    dropped from the sum; the threshold is 128."""
    L = O.lib()
    rend, gt = _ssim_case(seed=3)
    H, W = gt.shape[:2]
    m = _mask_case(W, H, 4)
    keep = m >= 128
    assert 0 < keep.sum() < W * H
    loss = L.oracle_ssim_loss_masked_f64(W, H, _p(rend), _p(gt), _p(m), 0.2, 1.0 / (3 * keep.sum()), None)
    ref = independent_masked_loss_sum(rend, gt, 0.2, keep)
    assert abs(loss - ref) < 1e-7 * ref
    # the unmasked case of the masked entry point is the unmasked loss, bit for bit
    ones = np.full((H, W), 255, np.uint8)
    v1, v2 = np.zeros(rend.shape), np.zeros(rend.shape)
    a = L.oracle_ssim_loss_masked_f64(W, H, _p(rend), _p(gt), _p(ones), 0.2, 1e-3, _p(v1))
    b = L.oracle_ssim_loss_f64(W, H, _p(rend), _p(gt), 0.2, 1e-3, _p(v2))
    assert a == b and np.array_equal(v1, v2)
    # an all-black mask leaves w per pixel
    zeros = np.zeros((H, W), np.uint8)
    assert L.oracle_ssim_loss_masked_f64(W, H, _p(rend), _p(gt), _p(zeros), 0.2, 1e-3, None) == pytest.approx(
        0.2 * W * H, rel=1e-12)


def test_ssim_masked_gradient_finite_differences():
    """The masked gradient (metal:3646-3648 fields, 3819-3834 L1 term) is the derivative of
    the masked objective: finite differences of the independent numpy loss, plus the
    reference's side-border term (unmasked: it comes from columns outside the image)."""
    L = O.lib()
    rend, gt = _ssim_case(seed=5)
    H, W = gt.shape[:2]
    m = _mask_case(W, H, 6)
    keep = m >= 128
    w, inv_n = 0.2, 1.0 / (3 * keep.sum())
    v = np.zeros(rend.shape)
    L.oracle_ssim_loss_masked_f64(W, H, _p(rend), _p(gt), _p(m), w, inv_n, _p(v))
    g = gauss_window()
    rng = np.random.default_rng(7)
    eps = 1e-6
    # samples at masked pixels, at kept pixels next to the masked block, and at random
    pts = [(H // 4 + 2, W // 3 + 3), (H // 2 - 1, W // 2 - 1), (H // 4 - 1, W // 3), (H // 2, W // 2)]
    pts += [(int(rng.integers(0, H)), int(rng.integers(0, W))) for _ in range(24)]
    n_masked = 0
    for yy, xx in pts:
        c = int(rng.integers(0, 3))
        r1 = rend.copy(); r1[yy, xx, c] += eps
        r2 = rend.copy(); r2[yy, xx, c] -= eps
        fd = (independent_masked_loss_sum(r1, gt, w, keep) - independent_masked_loss_sum(r2, gt, w, keep)) / (2 * eps)
        fd *= 3 * inv_n
        rows = sum(g[k] for k in range(11) if 0 <= yy - 5 + k < H)
        cols = sum(g[k] for k in range(11) if not (0 <= xx - 5 + k < W))
        x = gt[yy, xx, c] / 255.0
        extra = inv_n * (-w) * (2.0 / SSIM_C2) * (x - rend[yy, xx, c]) * rows * cols
        assert abs(v[yy, xx, c] - (fd + extra)) <= 5e-6 * abs(v[yy, xx, c]) + 1e-11, (yy, xx, c)
        n_masked += int(not keep[yy, xx])
    assert n_masked >= 2
    # a masked pixel deep inside the block has no L1 term and no own SSIM term: its
    # gradient is what its neighbours' windows give, with or without its own mask bit
    m2 = m.copy(); m2[H // 4 + 2, W // 3 + 3] = 255
    v2 = np.zeros(rend.shape)
    L.oracle_ssim_loss_masked_f64(W, H, _p(rend), _p(gt), _p(m2), w, inv_n, _p(v2))
    assert not np.allclose(v2[H // 4 + 2, W // 3 + 3], v[H // 4 + 2, W // 3 + 3])


# ------------------------------------------------------------- raster (oracle raster fwd/bwd)
def _raster_case(n, W, H, seed, single_tile=False):
    rng = np.random.default_rng(seed)
    op_mean, op_sd = (-4.5, 0.3) if single_tile else (-1.5, 0.7)  # no pixel saturates
    if single_tile:
        xy = rng.uniform(2, 14, (n, 2))
    else:
        xy = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    sx, sy = rng.uniform(1.0, 4.0, n), rng.uniform(1.0, 4.0, n)
    rho = rng.uniform(-0.5, 0.5, n)
    cxx, cyy, cxy = sx * sx + 0.3, sy * sy + 0.3, rho * sx * sy
    det = cxx * cyy - cxy * cxy
    conic = np.stack([cyy / det, -cxy / det, cxx / det], 1)
    rgb = rng.uniform(-0.4, 0.0, (n, 3))
    op = rng.normal(op_mean, op_sd, n)
    depth = rng.uniform(1, 5, n)
    return xy, conic, rgb, op, depth


def _raster(case, W, H, K, v_out=None):
    L = O.lib()
    xy, conic, rgb, op, depth = [np.ascontiguousarray(a, np.float64) for a in case]
    n = xy.shape[0]
    img = np.zeros((H, W, 3)); fT = np.zeros((H, W)); fi = np.zeros((H, W), np.int32)
    bg = (C.c_float * 3)(0.3, 0.2, 0.1)
    g = [np.zeros((n, 2)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)]
    kk = L.oracle_raster2d_f64(n, _p(xy), _p(conic), _p(rgb), _p(op), _p(depth), W, H, bg, K, _p(img), _p(fT),
                               _p(fi), _p(v_out) if v_out is not None else None, *[_p(a) for a in g])
    return img, fT, fi, g, kk


@pytest.mark.parametrize("n,W,H,K,single", [(60, 48, 32, 1, False), (1100, 16, 16, 0, True)],
                         ids=["monolithic", "chunked_K3"])
def test_raster_backward_finite_differences(n, W, H, K, single):
    case = _raster_case(n, W, H, seed=3, single_tile=single)
    rng = np.random.default_rng(4)
    v_out = rng.normal(size=(H, W, 3))
    img, fT, fi, (gxy, gcon, grgb, gop), kk = _raster(case, W, H, K, v_out)
    if single:
        # K rule: capacity 48 n over one tile -> 104 chunks, of which the list fills three
        assert kk == 104 and fT.min() > 1e-3
    eps = 1e-6
    checks = [(0, gxy, 2), (1, gcon, 3), (2, grgb, 3), (3, gop, 1)]
    idx = rng.choice(n, size=min(n, 25), replace=False)
    for arr_i, grad, width in checks:
        for i in idx:
            for k in range(width):
                cp = [a.copy() for a in case]
                cm = [a.copy() for a in case]
                if width == 1:
                    cp[arr_i][i] += eps; cm[arr_i][i] -= eps
                else:
                    cp[arr_i][i, k] += eps; cm[arr_i][i, k] -= eps
                lp = float((_raster(cp, W, H, K)[0] * v_out).sum())
                lm = float((_raster(cm, W, H, K)[0] * v_out).sum())
                fd = (lp - lm) / (2 * eps)
                got = grad[i] if width == 1 else grad[i, k]
                if arr_i == 1 and k == 1:
                    # the conic's off-diagonal b enters sigma once per symmetric entry
                    # (sigma = 0.5 d^T [[a, b], [b, c]] d): the VJP is per entry, half d/db
                    got = 2.0 * got
                assert abs(got - fd) <= 1e-5 * max(1.0, abs(fd)), (arr_i, i, k, got, fd)


def test_raster_forward_matches_bruteforce_compositing():
    """Independent per-pixel compositing in depth order (the reference's thresholds)."""
    W, H = 48, 32
    case = _raster_case(40, W, H, seed=5)
    img, fT, fi, _, _ = _raster(case, W, H, 1)
    xy, conic, rgb, op, depth = case
    order = np.lexsort((np.arange(len(depth)), np.array(depth, np.float32).view(np.uint32)))
    bg = np.array([0.3, 0.2, 0.1], np.float32).astype(np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    Tm = np.ones((H, W)); Cm = np.zeros((H, W, 3)); done = np.zeros((H, W), bool)
    # Gaussians outside a pixel's 16x16 tile list never reach it: build the tile membership
    ax = np.ceil(3 * np.sqrt(conic[:, 2] / (conic[:, 0] * conic[:, 2] - conic[:, 1] ** 2)))
    ay = np.ceil(3 * np.sqrt(conic[:, 0] / (conic[:, 0] * conic[:, 2] - conic[:, 1] ** 2)))
    for g in order:
        x0 = np.clip(np.trunc((xy[g, 0] / 16 - ax[g] / 16)), 0, (W + 15) // 16)
        x1 = np.clip(np.trunc((xy[g, 0] / 16 + ax[g] / 16 + 1)), 0, (W + 15) // 16)
        y0 = np.clip(np.trunc((xy[g, 1] / 16 - ay[g] / 16)), 0, (H + 15) // 16)
        y1 = np.clip(np.trunc((xy[g, 1] / 16 + ay[g] / 16 + 1)), 0, (H + 15) // 16)
        in_tile = (xs // 16 >= x0) & (xs // 16 < x1) & (ys // 16 >= y0) & (ys // 16 < y1)
        dx, dy = xy[g, 0] - xs, xy[g, 1] - ys
        sig = 0.5 * (conic[g, 0] * dx * dx + conic[g, 2] * dy * dy) + conic[g, 1] * dx * dy
        a = np.minimum(0.999, 1 / (1 + np.exp(-op[g])) * np.exp(-sig))
        ok = in_tile & ~done & (sig >= 0) & (sig < 5.55) & (a >= 1 / 255)
        nT = Tm * (1 - a)
        stop = ok & (nT <= 1e-4)
        done |= stop
        ok &= ~stop
        Cm += (ok * a * Tm)[..., None] * np.maximum(rgb[g] + 0.5, 0)
        Tm = np.where(ok, nT, Tm)
    ref = np.clip(Cm + Tm[..., None] * bg, 0, 1)
    assert np.abs(img - ref).max() < 1e-10
    assert np.abs(fT - Tm).max() < 1e-10


# ------------------------------------------------------------------- projection (forward/VJP)
def _cam(W=320, H=240, f=300.0, seed=9):
    c2w = scene.synth_cameras(seed, 1, 3.0, 0.0)[0]
    return scene.prepare_camera(c2w, f, f, W / 2, H / 2, W, H), c2w


def _proj(means, scales, quats, cam):
    L = O.lib()
    n = means.shape[0]
    xys, con, dep = np.zeros((n, 2)), np.zeros((n, 3)), np.zeros(n)
    rad = np.zeros(n, np.int32)
    L.oracle_project_f64(n, _p(means), _p(scales), _p(quats), cam, _p(xys), _p(con), _p(dep), _p(rad))
    return xys, con, dep, rad


def _gauss3d(n, seed):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-0.6, 0.6, (n, 3))
    scales = rng.normal(-2.5, 0.3, (n, 3))
    q = rng.normal(size=(n, 4))
    quats = q / np.linalg.norm(q, axis=1, keepdims=True)
    return means, scales, quats


def test_projection_closed_forms():
    cam, c2w = _cam()
    means, scales, quats = _gauss3d(200, 1)
    xys, con, dep, rad = _proj(means, scales, quats, cam)
    vis = rad > 0
    assert vis.sum() > 150
    V = np.array(cam.viewmat).reshape(4, 4)
    pc = (V[:3, :3] @ means.T).T + V[:3, 3]
    # pinhole: x = fx X/Z + cx - 0.5 (ndc2pix's half-pixel offset)
    assert np.abs(xys[vis, 0] - (cam.fx * pc[vis, 0] / pc[vis, 2] + cam.cx - 0.5)).max() < 1e-4
    assert np.abs(xys[vis, 1] - (cam.fy * pc[vis, 1] / pc[vis, 2] + cam.cy - 0.5)).max() < 1e-4
    assert np.allclose(dep[vis], pc[vis, 2])
    # EWA: cov2d = J W R S S R^T W^T J^T + 0.3 I; conic = cov2d^-1
    for i in np.flatnonzero(vis)[:50]:
        w_, x_, y_, z_ = quats[i]
        R = np.array([[1 - 2 * (y_ * y_ + z_ * z_), 2 * (x_ * y_ - w_ * z_), 2 * (x_ * z_ + w_ * y_)],
                      [2 * (x_ * y_ + w_ * z_), 1 - 2 * (x_ * x_ + z_ * z_), 2 * (y_ * z_ - w_ * x_)],
                      [2 * (x_ * z_ - w_ * y_), 2 * (y_ * z_ + w_ * x_), 1 - 2 * (x_ * x_ + y_ * y_)]])
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        S = np.diag(np.exp(scales[i]))
        X, Y, Z = pc[i]
        J = np.array([[cam.fx / Z, 0, -cam.fx * X / Z ** 2], [0, cam.fy / Z, -cam.fy * Y / Z ** 2]])
        cov2 = J @ V[:3, :3] @ R @ S @ S @ R.T @ V[:3, :3].T @ J.T + 0.3 * np.eye(2)
        Cn = np.array([[con[i, 0], con[i, 1]], [con[i, 1], con[i, 2]]])
        assert np.allclose(Cn @ cov2, np.eye(2), atol=1e-6)


def test_projection_vjp_finite_differences():
    L = O.lib()
    cam, c2w = _cam()
    n = 40
    means, scales, quats = _gauss3d(n, 2)
    xys, con, dep, rad = _proj(means, scales, quats, cam)
    vis = np.flatnonzero(rad > 0)
    rng = np.random.default_rng(3)
    v_con = rng.normal(size=(n, 3))
    zero2 = np.zeros((n, 2))
    vm, vs, vq = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 4))
    L.oracle_project_vjp_f64(n, _p(means), _p(scales), _p(quats), cam, _p(con), _p(zero2), _p(v_con), _p(vm),
                             _p(vs), _p(vq))
    eps = 1e-6

    # v_conic is per symmetric matrix entry (see the raster test): the scalar b counts twice
    wcon = v_con * np.array([1.0, 2.0, 1.0])

    def f(m, s, q):
        return (_proj(m, s, q, cam)[1] * wcon).sum(1)
    for i in vis[:15]:
        for k in range(3):
            mp, mm_ = means.copy(), means.copy(); mp[i, k] += eps; mm_[i, k] -= eps
            fd = (f(mp, scales, quats)[i] - f(mm_, scales, quats)[i]) / (2 * eps)
            assert abs(vm[i, k] - fd) <= 1e-5 * max(1.0, abs(fd))
            sp, sm = scales.copy(), scales.copy(); sp[i, k] += eps; sm[i, k] -= eps
            fd = (f(means, sp, quats)[i] - f(means, sm, quats)[i]) / (2 * eps)
            assert abs(vs[i, k] - fd) <= 1e-5 * max(1.0, abs(fd))
        # quaternion: the VJP is taken w.r.t. the normalised components; its tangential
        # part (I - q q^T) g is the derivative along the unit sphere
        q = quats[i]
        fdq = np.zeros(4)
        for k in range(4):
            qp, qm = quats.copy(), quats.copy(); qp[i, k] += eps; qm[i, k] -= eps
            fdq[k] = (f(means, scales, qp)[i] - f(means, scales, qm)[i]) / (2 * eps)
        tang = vq[i] - q * (q @ vq[i])
        assert np.abs(tang - fdq).max() <= 1e-5 * max(1.0, np.abs(fdq).max())


def test_projection_xy_vjp_on_axis():
    """The pixel-centre VJP holds the perspective divide fixed (reference project_pix_vjp);
    on the optical axis the divide's own derivative term vanishes, so it is the exact
    derivative there; off axis the difference is exactly that term."""
    L = O.lib()
    cam, c2w = _cam()
    V = np.array(cam.viewmat).reshape(4, 4)
    P = np.array(cam.projmat).reshape(4, 4)
    fwd = -c2w[:3, 2]  # OpenGL camera looks down -z
    n = 6
    means = np.array([c2w[:3, 3] + t * fwd for t in np.linspace(2.0, 4.0, n)])
    means[3:] += np.random.default_rng(0).normal(0, 0.2, (3, 3))  # off axis
    _, scales, quats = _gauss3d(n, 4)
    xys, con, dep, rad = _proj(means, scales, quats, cam)
    rng = np.random.default_rng(5)
    v_xy = rng.normal(size=(n, 2))
    vm, vs, vq = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 4))
    L.oracle_project_vjp_f64(n, _p(means), _p(scales), _p(quats), cam, _p(con), _p(v_xy), _p(np.zeros((n, 3))),
                             _p(vm), _p(vs), _p(vq))
    eps = 1e-6
    for i in range(n):
        fd = np.zeros(3)
        for k in range(3):
            mp, mm_ = means.copy(), means.copy(); mp[i, k] += eps; mm_[i, k] -= eps
            fd[k] = ((_proj(mp, scales, quats, cam)[0][i] - _proj(mm_, scales, quats, cam)[0][i]) * v_xy[i]).sum() / (2 * eps)
        ph = P @ np.append(means[i], 1.0)
        rw = 1.0 / (ph[3] + 1e-6)
        missing = -(0.5 * cam.width * v_xy[i, 0] * ph[0] + 0.5 * cam.height * v_xy[i, 1] * ph[1]) * rw * rw * P[3, :3]
        assert np.abs(vm[i] + missing - fd).max() <= 1e-5 * max(1.0, np.abs(fd).max())
        if i < 3:
            assert np.abs(vm[i] - fd).max() <= 1e-5 * max(1.0, np.abs(fd).max())


def test_sh_basis_orthonormal():
    """The 16 real SH basis functions (degree <= 3) are orthonormal on the sphere; a wrong
    constant, sign-independent term or polynomial breaks the Gram matrix."""
    L = O.lib()
    m = 12000
    k = np.arange(m) + 0.5
    th = np.arccos(1 - 2 * k / m)
    ph = np.pi * (1 + 5 ** 0.5) * k
    dirs = np.ascontiguousarray(np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], 1))
    B = np.zeros((m, 16))
    out = np.zeros(3)
    dc = np.zeros(3)
    rest = np.zeros((15, 3))
    # channel c of the colour is the basis value of coefficient (basis, c): use a different
    # basis per channel to evaluate three at a time
    for b0 in range(0, 16, 3):
        dc[:] = 0
        rest[:] = 0
        bs = list(range(b0, min(b0 + 3, 16)))
        for c, b in enumerate(bs):
            if b == 0:
                dc[c] = 1.0
            else:
                rest[b - 1, c] = 1.0
        for j in range(m):
            L.oracle_sh_color_f64(3, dirs[j].ctypes.data, dc.ctypes.data, rest.ctypes.data, out.ctypes.data)
            for c, b in enumerate(bs):
                B[j, b] = out[c]
    gram = (B.T @ B) * 4 * np.pi / m
    assert np.allclose(gram, np.eye(16), atol=2e-3)


# --------------------------------------------------------------------- lists (count/scan/sort)
def test_tile_lists_bruteforce():
    g, hidden, _ = (scene.synth_gaussians(0, 21, 1500, 3), None, None)
    c2w = scene.synth_cameras(22, 2, 3.0, 0.0)
    W, H = 200, 176
    cam = scene.prepare_camera(c2w[0], 200.0, 200.0, W / 2, H / 2, W, H)
    st = O.HostState(g, dtype=np.float64)
    r = O.render(st, cam, 3, scene.DEFAULT_BG)
    tbx, tby = (W + 15) // 16, (H + 15) // 16
    lists = [[] for _ in range(tbx * tby)]
    for i in np.flatnonzero(r["radii"] > 0):
        x, y = r["xys"][i]
        ax, ay = r["aabb"][i]
        x0 = int(min(max(0, np.trunc(x / 16 - ax / 16)), tbx)); x1 = int(min(max(0, np.trunc(x / 16 + ax / 16 + 1)), tbx))
        y0 = int(min(max(0, np.trunc(y / 16 - ay / 16)), tby)); y1 = int(min(max(0, np.trunc(y / 16 + ay / 16 + 1)), tby))
        key = (int(np.float32(r["depths"][i]).view(np.uint32)) << 32) | int(i)
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                lists[ty * tbx + tx].append(key)
    counts = np.array([len(l) for l in lists])
    assert np.array_equal(counts, r["tile_counts"])
    assert counts.sum() == r["tiles_hit"][r["radii"] > 0].sum() == r["num_isect"]
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert np.array_equal(r["tile_bins"][:, 0], starts)
    ids = np.concatenate([[k & 0xFFFFFFFF for k in sorted(l)] for l in lists if l])
    assert np.array_equal(ids, r["ids"])


def test_k_max_rule():
    L = O.lib()
    assert L.oracle_k_max(48 * 50000, 256) == 110          # 256x256 view, 50k Gaussians
    assert L.oracle_k_max(48 * 1500000, 6500) == 1         # >= 400 tiles: monolithic
    assert L.oracle_k_max(48 * 100, 300) == 2              # floor of 2 chunks
    assert L.oracle_k_max(100, 10) == 1                    # capped by capacity / 512 chunks


# ------------------------------------------------------------------- Adam and SGLD
def test_adam_step_is_textbook():
    g = scene.synth_gaussians(0, 31, 800, 1)
    c2w = scene.synth_cameras(32, 1, 3.0, 0.0)
    W = H = 64
    cam = scene.prepare_camera(c2w[0], 64.0, 64.0, 32, 32, W, H)
    gt = np.random.default_rng(0).integers(0, 256, (H, W, 3)).astype(np.uint8)
    st = O.HostState(g, dtype=np.float64)
    p0 = {k: v.copy() for k, v in st.view().items()}
    hp = scene.step_params(3, scene.LR_INIT, W, H, 3, 1, 1, sgld=False)
    st.m["means"][:] = 0.01
    st.v["means"][:] = 2e-4
    m0, v0 = st.m["means"].copy(), st.v["means"].copy()
    o = O.train_step(st, cam, gt, hp, want_ids=False)
    gr = o["v_mean"]
    b1, b2, eps = 0.9, 0.999, 1e-8
    m = b1 * m0 + (1 - b1) * gr
    v = b2 * v0 + (1 - b2) * gr * gr
    lr = np.float32(scene.LR_INIT[0])
    mhat = m / (1 - b1 ** 3)
    vhat = v / (1 - b2 ** 3)
    want = p0["means"] - lr * mhat / (np.sqrt(vhat) + eps * 1.0)
    # the reference divides by (sqrt(v)/sqrt(bc2) + eps), i.e. eps is not bias-corrected
    want_ref = p0["means"] - lr / (1 - b1 ** 3) * m / (np.sqrt(v) / np.sqrt(1 - b2 ** 3) + eps)
    assert np.abs(st.view()["means"] - want_ref).max() < 1e-9
    assert np.abs(st.view()["means"] - want).max() < 1e-6


def test_sgld_noise_is_standard_normal():
    L = O.lib()
    n = 100000
    z = np.zeros((n, 3), np.float32)
    L.oracle_sgld_noise_f32(n, 7, _p(z))
    flat = z.reshape(-1).astype(np.float64)
    assert abs(flat.mean()) < 0.01
    assert abs(flat.var() - 1) < 0.01
    assert stats.kstest(flat, "norm").pvalue > 1e-3
    # independent components
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.01


# -------------------------------------------------------------------- MCMC sampler
def test_relocation_sampler_chi_square():
    rng = np.random.default_rng(0)
    n_alive, n_dead = 150, 30000
    N = n_alive + n_dead
    g = scene.synth_gaussians(0, 41, N, 0)
    g["opac"][:n_alive] = rng.normal(0, 1.5, n_alive)
    g["opac"][n_alive:] = -8.0  # sigmoid < 0.005: dead
    g["means"][:n_alive] = rng.uniform(-1, 1, (n_alive, 3))
    g["means"][n_alive:] = 0.0
    st = O.HostState(g, capacity=N)  # no room to grow
    alive_means = st.p["means"][:n_alive].copy()
    alive_op = st.p["opac"][:n_alive].copy()
    info = np.zeros(6, np.int32)
    O.lib().oracle_mcmc_step(st.struct(), N, 3.0, 0, _p(info))
    assert info[1] == n_dead and info[2] == 0
    moved = st.p["means"][n_alive:N]
    src = np.array([np.flatnonzero((alive_means == mm).all(1))[0] for mm in moved])
    counts = np.bincount(src, minlength=n_alive)
    p = 1 / (1 + np.exp(-alive_op.astype(np.float64)))
    exp = p / p.sum() * n_dead
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, n_alive - 1) > 1e-3


def _ssim_fp64_scipy(img, gt):
    """Replicate-padded 11-tap Gaussian SSIM (sigma 1.5) in float64 with scipy."""
    x = np.arange(11) - 5.0
    w = np.exp(-(x * x) / 4.5)
    w /= w.sum()
    r = img.astype(np.float64)
    g = gt.astype(np.float64) / 255.0

    def blur(a):
        return ndimage.correlate1d(ndimage.correlate1d(a, w, axis=1, mode="nearest"), w, axis=0, mode="nearest")
    tot = 0.0
    for c in range(3):
        a, b = r[..., c], g[..., c]
        m1, m2 = blur(a), blur(b)
        s11, s22, s12 = blur(a * a) - m1 * m1, blur(b * b) - m2 * m2, blur(a * b) - m1 * m2
        c1, c2 = 1e-4, 9e-4
        tot += (((2 * m1 * m2 + c1) * (2 * s12 + c2)) / ((m1 * m1 + m2 * m2 + c1) * (s11 + s22 + c2))).sum()
    return tot


def test_eval_metrics_pins():
    """Evaluation metrics (model.cpp:34-49, ssim.cpp:61-123): SSIM(x, x) = 1 exactly, PSNR
    of a constant offset in closed form, and SSIM against an FP64 scipy formulation with
    replicate ('nearest') padding."""
    rng = np.random.default_rng(3)
    H, W = 37, 53
    gt = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    same = (gt.astype(np.float32) * np.float32(1.0 / 255.0)).astype(np.float32)
    s = O.eval_sums(same, gt)
    assert s[0] == 0.0 and s[1] == 0.0
    assert s[2] == H * W * 3  # every SSIM value is exactly 1
    gt_mid = np.full((H, W, 3), 100, np.uint8)
    off = (gt_mid.astype(np.float32) * np.float32(1.0 / 255.0) + np.float32(0.05)).astype(np.float32)
    m = O.eval_metrics(off, gt_mid)
    d = float(np.float32(off[0, 0, 0]) - np.float32(100 * np.float32(1.0 / 255.0)))
    assert abs(m["psnr"] - 10.0 * np.log10(1.0 / (d * d))) < 1e-4
    assert abs(m["l1"] - abs(d)) < 1e-7
    img = np.clip(same + rng.normal(0, 0.08, same.shape), 0, 1).astype(np.float32)
    s = O.eval_sums(img, gt)
    ref = _ssim_fp64_scipy(img, gt)
    assert abs(s[2] - ref) / abs(ref) < 2e-5, (s[2], ref)
    assert abs(s[0] - ((img.astype(np.float64) - gt / 255.0) ** 2).sum()) / s[0] < 1e-6


def test_batch_step_restatement_pins():
    """The oracle's multi-view batched step: a one-view batch is the plain step, and a batch
    of the same view twice (each weighted 1/2, an exact halving) is the plain step too, bit
    for bit; a two-view batch differs from either single view."""
    g = scene.synth_gaussians(0, 4, 600, 3)
    hidden = scene.synth_gaussians(0, 5, 400, 3)
    c2w = scene.synth_cameras(9, 2, 3.0, 0.0)
    W, H = 64, 48
    cams = [scene.prepare_camera(c, 64.0, 64.0, W / 2.0, H / 2.0, W, H) for c in c2w]
    gts = [scene.quantize(O.render(O.HostState(hidden), c, 3, scene.DEFAULT_BG, want_ids=False)["img"])
           for c in cams]
    hp = scene.step_params(1, scene.LR_INIT, W, H, 1, 3, 3)
    plain = O.HostState(g)
    O.train_step(plain, cams[0], gts[0], hp, want_ids=False)
    one = O.HostState(g)
    O.train_step_batch(one, [cams[0]], [gts[0]], [hp])
    hp2 = scene.step_params(1, scene.LR_INIT, W, H, 1, 3, 3)
    hp2.loss_inv_n = float(np.float32(1.0) / np.float32(W * H * 3 * 2))
    twice = O.HostState(g)
    O.train_step_batch(twice, [cams[0], cams[0]], [gts[0], gts[0]], [hp2, hp2])
    both = O.HostState(g)
    O.train_step_batch(both, cams, gts, [hp2, hp2])
    for k in O.GROUP_SHAPES:
        assert np.array_equal(one.p[k], plain.p[k]), k
        assert np.array_equal(one.m[k], plain.m[k]) and np.array_equal(one.v[k], plain.v[k]), k
        assert np.array_equal(twice.p[k], plain.p[k]), k
    assert not np.array_equal(both.p["means"], plain.p["means"])

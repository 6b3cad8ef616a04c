"""Pins of three oracle functions that the reference's own tests leave open, each against
what the mathematics fixes rather than a re-typed copy of the oracle's formula:

- sgld_apply (reference sgld_noise_kernel, core/metal/msplat_metal.metal:3892-3967):
  closed forms at the identity and at a 90-degree turn, the opacity weight's midpoint, the
  displacement cap and covariance under a rotation of the world frame;
- reg_grad (mcmc_reg_grad_kernel, msplat_metal.metal:3987-4026): FP64 central differences
  of the regulariser objective (mean over N opacities, over 3N scales, times the duty);
- ssim_loss_counted (ssim_fused_v_fwd_h_bwd_kernel's loss bookkeeping,
  msplat_metal.metal:3598-3680): with the render equal to the ground truth every counted
  pixel's loss is 0, so the loss sum is w times the number of border pixels whose terms
  are never added; that number is derived here by modular arithmetic over the
  threadgroup's 16 x 26 derivative sweep, and by hand for two sizes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

D = C.c_double


def _p(a):
    return a.ctypes.data


def _lib():
    L = O.lib()
    vp = C.c_void_p
    L.oracle_sgld_apply_f64.argtypes = [vp, vp, vp, D, vp, D, D, D]
    L.oracle_sgld_apply_f32.argtypes = [vp, vp, vp, C.c_float, vp, C.c_float, C.c_float, C.c_float]
    L.oracle_reg_grad_f64.argtypes = [C.c_int, vp, vp, vp, D, D, D, vp, vp]
    return L


def sgld_move(ls, q, opac_logit, n, noise_lr=5e5, xyz_lr=1.6e-4, max_disp=1e9):
    """The displacement sgld_apply adds to a mean at the origin (FP64 oracle)."""
    mean = np.zeros(3)
    ls, q, n = (np.ascontiguousarray(a, np.float64) for a in (ls, q, n))
    _lib().oracle_sgld_apply_f64(_p(mean), _p(ls), _p(q), opac_logit, _p(n), noise_lr, xyz_lr, max_disp)
    return mean


def quat_mul(a, b):
    """Hamilton product (w, x, y, z)."""
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


def rotation_of(q):
    """Rotation matrix of a unit quaternion via the Rodrigues formula (axis-angle), an
    independent route from the quaternion-product matrix the kernel writes."""
    q = np.asarray(q, np.float64) / np.linalg.norm(q)
    w, v = q[0], q[1:]
    s = np.linalg.norm(v)
    if s == 0:
        return np.eye(3)
    axis = v / s
    ang = 2 * np.arctan2(s, w)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


# noise_lr * xyz_lr * w(o) at o = -60: w(o) = sigmoid(100 ((1 - sigmoid(o)) - x0)) with the
# reference's float constant x0 = 0.995f, and sigmoid(-60) = 9e-27, so w = sigmoid(100 (1 - x0))
X0 = float(np.float32(0.995))
O_DEAD = -60.0
NS = 5e5 * 1.6e-4 / (1 + np.exp(-100 * (1 - X0)))


def test_sgld_identity_rotation_isotropic_scale():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.uniform(-6, 2)
        n = rng.standard_normal(3)
        d = sgld_move([a, a, a], [1, 0, 0, 0], O_DEAD, n)
        np.testing.assert_allclose(d, np.exp(2 * a) * NS * n, rtol=1e-13)
    # an isotropic Gaussian moves the same whatever its rotation
    for _ in range(20):
        a = rng.uniform(-6, 2)
        n = rng.standard_normal(3)
        d = sgld_move([a, a, a], rng.standard_normal(4), O_DEAD, n)
        np.testing.assert_allclose(d, np.exp(2 * a) * NS * n, rtol=1e-12)


def test_sgld_quarter_turn_maps_local_axes():
    # 90 degrees about z: local x -> world y, local y -> world -x; the world-x component of
    # the move is scaled by the local y variance and the world-y component by the local x
    # variance (a transposed rotation would swap them)
    s = np.array([-1.0, -2.5, -4.0])
    q = [np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)]
    n = np.array([0.3, -1.2, 0.7])
    d = sgld_move(s, q, O_DEAD, n)
    v = np.exp(2 * s)
    np.testing.assert_allclose(d, NS * np.array([v[1] * n[0], v[0] * n[1], v[2] * n[2]]), rtol=1e-12, atol=1e-18)


def test_sgld_opacity_weight():
    # w(o) = sigmoid(100 ((1 - sigmoid(o)) - x0)) is 1/2 where sigmoid(o) = 1 - x0
    o = np.log((1 - X0) / X0)
    n = np.array([1.0, -2.0, 0.5])
    d = sgld_move([-3, -3, -3], [1, 0, 0, 0], o, n)
    np.testing.assert_allclose(d, 0.5 * np.exp(-6) * 5e5 * 1.6e-4 * n, rtol=1e-10)
    # fully transparent: w = sigmoid(100 (1 - x0))
    d0 = sgld_move([-3, -3, -3], [1, 0, 0, 0], O_DEAD, n)
    np.testing.assert_allclose(d0, np.exp(-6) * NS * n, rtol=1e-12)
    # opaque (sigmoid(o) = 1/2): w = sigmoid(-49.5) = 3e-22, the mean stays put
    d1 = sgld_move([-3, -3, -3], [1, 0, 0, 0], 0.0, n)
    np.testing.assert_allclose(d1, np.exp(-6) * 5e5 * 1.6e-4 / (1 + np.exp(100 * (X0 - 0.5))) * n, rtol=1e-10)


def test_sgld_cap_keeps_direction():
    rng = np.random.default_rng(1)
    for _ in range(20):
        ls = rng.uniform(-1, 1, 3)
        q = rng.standard_normal(4)
        n = rng.standard_normal(3)
        free = sgld_move(ls, q, O_DEAD, n)
        cap = 0.25 * np.linalg.norm(free)
        d = sgld_move(ls, q, O_DEAD, n, max_disp=cap)
        assert abs(np.linalg.norm(d) - cap) <= 1e-12 * cap
        np.testing.assert_allclose(d / cap, free / np.linalg.norm(free), atol=1e-12)
        # below the cap nothing changes
        d2 = sgld_move(ls, q, O_DEAD, n, max_disp=4 * np.linalg.norm(free))
        np.testing.assert_allclose(d2, free, rtol=1e-14)


def test_sgld_frame_rotation_covariance():
    """Rotating the world by Q (q -> q_Q q, noise -> Q noise) rotates the move by Q:
    the move is R diag(s^2) R^T n, a covariance applied to the noise."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        ls = rng.uniform(-4, 0, 3)
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        n = rng.standard_normal(3)
        qQ = rng.standard_normal(4)
        qQ /= np.linalg.norm(qQ)
        Q = rotation_of(qQ)
        d = sgld_move(ls, q, O_DEAD, n)
        d_rot = sgld_move(ls, quat_mul(qQ, q), O_DEAD, Q @ n)
        np.testing.assert_allclose(d_rot, Q @ d, rtol=1e-10, atol=1e-12 * np.linalg.norm(d))
        # and the move is the covariance product, with the rotation built independently
        R = rotation_of(q)
        np.testing.assert_allclose(d, NS * R @ np.diag(np.exp(2 * ls)) @ R.T @ n, rtol=1e-10,
                                   atol=1e-12 * np.linalg.norm(d))


def test_sgld_fp32_within_float_rounding_of_fp64():
    L = _lib()
    rng = np.random.default_rng(3)
    for _ in range(50):
        ls = rng.uniform(-5, 0, 3)
        q = rng.standard_normal(4)
        n = rng.standard_normal(3)
        o = rng.uniform(-9, -4)
        d64 = sgld_move(ls, q, o, n)
        m32 = np.zeros(3, np.float32)
        ls32, q32, n32 = ls.astype(np.float32), q.astype(np.float32), n.astype(np.float32)
        L.oracle_sgld_apply_f32(_p(m32), _p(ls32), _p(q32), float(np.float32(o)), _p(n32), 5e5, 1.6e-4, 1e9)
        np.testing.assert_allclose(m32, d64, rtol=2e-5, atol=2e-6 * np.linalg.norm(d64))


# ------------------------------------------------------------------ regulariser gradient
def reg_objective(opac, ls, duty, op_reg, sc_reg):
    """duty-weighted op_reg * mean_i sigmoid(o_i) + sc_reg * mean_{i,k} exp(s_ik): the
    regulariser of 3DGS-MCMC whose gradient the reference adds before Adam."""
    N = opac.size
    sig = 1.0 / (1.0 + np.exp(-opac))
    return float(np.sum(duty * (op_reg * sig / N + sc_reg * np.exp(ls).sum(1) / (3 * N))))


def test_reg_grad_finite_differences():
    L = _lib()
    rng = np.random.default_rng(4)
    N = 37
    opac = rng.uniform(-4, 4, N)
    ls = rng.uniform(-6, 1, (N, 3))
    dc_sq = np.zeros((N, 3))
    seen = rng.random(N) < 0.5
    dc_sq[seen] = rng.uniform(1e-9, 1e-3, (seen.sum(), 3))
    op_reg, sc_reg, invis = 0.01, 0.02, 0.37
    duty = np.where(seen, invis, 1.0)
    v_o = rng.standard_normal(N)   # gradients are added to what is there
    v_s = rng.standard_normal((N, 3))
    v_o0, v_s0 = v_o.copy(), v_s.copy()
    L.oracle_reg_grad_f64(N, _p(opac), _p(ls), _p(dc_sq), op_reg, sc_reg, invis, _p(v_o), _p(v_s))
    # the objective is ~5e-3 summed over the rows: eps = 1e-4 keeps rounding in the difference
    # near 1e-9 of the derivative; the scale term carries the float constant 1/3 (3e-8)
    eps = 1e-4
    for i in range(N):
        a, b = opac.copy(), opac.copy()
        a[i] += eps
        b[i] -= eps
        fd = (reg_objective(a, ls, duty, op_reg, sc_reg) - reg_objective(b, ls, duty, op_reg, sc_reg)) / (2 * eps)
        assert abs((v_o[i] - v_o0[i]) - fd) <= 1e-6 * abs(fd) + 1e-15, (i, v_o[i] - v_o0[i], fd)
        for k in range(3):
            a, b = ls.copy(), ls.copy()
            a[i, k] += eps
            b[i, k] -= eps
            fd = (reg_objective(opac, a, duty, op_reg, sc_reg) - reg_objective(opac, b, duty, op_reg, sc_reg)) / (2 * eps)
            assert abs((v_s[i, k] - v_s0[i, k]) - fd) <= 1e-6 * abs(fd) + 1e-15, (i, k)


def test_reg_grad_clamped_scale_quirk():
    """Beyond |s| = 15 the reference adds exp(+-15) (the clamp bounds exp's input; its own
    derivative is not applied): the gradient stays exp(15) sc_reg / (3N)."""
    L = _lib()
    N = 4
    opac = np.zeros(N)
    ls = np.array([[16.0, -20.0, 0.0]] * N)
    dc_sq = np.zeros((N, 3))
    v_o, v_s = np.zeros(N), np.zeros((N, 3))
    L.oracle_reg_grad_f64(N, _p(opac), _p(ls), _p(dc_sq), 0.0, 0.3, 1.0, _p(v_o), _p(v_s))
    np.testing.assert_allclose(v_s[0], 0.3 / (3 * N) * np.exp([15.0, -15.0, 0.0]), rtol=1e-7)
    assert (v_o == 0).all()


# ---------------------------------------------------------------- SSIM loss bookkeeping
def uncounted_pixels(W, H):
    """Pixels whose loss terms ssim_fused_v_fwd_h_bwd never adds. In a 16 x 16
    threadgroup, pixel (lx, ly) is item i = 26 ly + lx + 5 of the 16 x 26 derivative sweep,
    whose terms thread i mod 256 accumulates; a thread adds them only when its own pixel
    (j mod 16, j div 16), j = i mod 256, lies in the image. Counted here tile by tile from
    the residues: j mod 16 = (10 ly + lx + 5) mod 16 since 26 = 10 (mod 16) and 256 = 0
    (mod 16)."""
    count = 0
    for ty in range((H + 15) // 16):
        for tx in range((W + 15) // 16):
            rw, rh = min(16, W - 16 * tx), min(16, H - 16 * ty)
            ly, lx = np.meshgrid(np.arange(rh), np.arange(rw), indexing="ij")
            col = (10 * ly + lx + 5) % 16
            row = ((26 * ly + lx + 5) % 256) // 16
            count += int(np.sum((col >= rw) | (row >= rh)))
    return count


def test_uncounted_pixels_by_hand():
    # W = 17, H = 16: the cut tile has one column (lx = 0); (10 ly + 5) mod 16 is odd, never 0,
    # so all 16 of its pixels are uncounted. W = 16, H = 17: the cut tile has one row
    # (ly = 0); items lx + 5 >= 16, i.e. lx = 11..15, belong to threads of row 1: 5 pixels.
    assert uncounted_pixels(17, 16) == 16
    assert uncounted_pixels(16, 17) == 5
    assert uncounted_pixels(64, 48) == 0


@pytest.mark.parametrize("W,H", [(17, 16), (16, 17), (53, 37), (96, 80), (61, 64), (160, 144), (100, 75)])
def test_ssim_loss_count_quirk(W, H):
    L = O.lib()
    rng = np.random.default_rng(W * 1000 + H)
    gt = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    # the render equal to the ground truth as the loss reads it (uint8 times 1/255 in float)
    rend32 = gt.astype(np.float32) * np.float32(1.0 / 255.0)
    rend = np.ascontiguousarray(rend32, np.float64)
    w = 0.2
    loss = L.oracle_ssim_loss_f64(W, H, _p(rend), _p(gt), w, 1.0 / (W * H * 3), None)
    want = w * uncounted_pixels(W, H)
    assert abs(loss - want) <= 1e-9 * W * H, (loss, want)
    rf = np.ascontiguousarray(rend32)
    loss32 = L.oracle_ssim_loss_f32(W, H, _p(rf), _p(gt), np.float32(w), np.float32(1.0 / (W * H * 3)), None)
    assert abs(loss32 - want) <= 1e-5 * W * H, (loss32, want)

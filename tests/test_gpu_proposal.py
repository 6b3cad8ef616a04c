"""The learned relocation proposal inside the trainer: its features against a host
recomputation from the gathered statistics (within one bf16 rounding step), its network
outputs against the oracle's network, the learned relocation draw and placement against the
oracle's sampler fed the same outputs (sources, copies and counters bit-exact, Eq. 9 rows
1e-6), and an online-training run whose counters and held-out judge behave as specified."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from msplat_amd import GaussianTrainer, TrainingConfig, scene
from parity_util import assert_exact, assert_rel, small_scene
from test_gpu_trainer import GROUPS, _dataset, _host_state, _mcmc_scene, _state

pytestmark = pytest.mark.gpu


def _features_ref(st, stats, views):
    n = st["opac"].size
    o = st["opac"].astype(np.float32)
    sc = st["scales"].reshape(n, 3)
    q = st["quats"].reshape(n, 4)
    dc = st["f_dc"].reshape(n, 3)
    vis, gsum, maxr, dsum, dsq = stats.reshape(5, -1)[:, :n]
    with np.errstate(all="ignore"):
        f = np.zeros((n, 16), np.float32)
        f[:, 0] = np.float32(1) / (np.float32(1) + np.exp(-o))
        f[:, 1:4] = sc
        qn = np.sqrt((q * q).sum(1))
        f[:, 4:8] = q * np.where(qn > 0, np.float32(1) / qn, 0)[:, None]
        iv = np.where(vis > 0, np.float32(1) / vis, 0).astype(np.float32)
        gn = gsum * iv
        f[:, 8] = np.log10(gn + np.float32(1e-10)) * np.float32(0.2) + np.float32(1)
        f[:, 9] = np.log1p(maxr) * np.float32(0.25)
        f[:, 10] = vis * np.float32(1.0 / views)
        dm = dsum * iv
        dv = np.maximum(dsq * iv - dm * dm, 0)
        f[:, 11] = dm * np.float32(0.2)
        f[:, 12] = np.sqrt(dv) * np.float32(0.2)
        f[:, 13:16] = dc
    f[~np.isfinite(f)] = 0
    return f


def _bf16_to_f32(h):
    return (h.astype(np.uint32) << 16).view(np.float32)


def test_learned_relocation_matches_oracle():
    W, H = 96, 80
    g = _mcmc_scene()
    n = g["means"].shape[0]
    hidden = scene.synth_gaussians(0, 99, 500, 3)
    ds, _ = _dataset(g, hidden, W, H, n_cams=2)
    cap = (4 * n) // 3 + 2
    tr = GaussianTrainer(ds, TrainingConfig(iterations=10000, cap_max=cap))
    tr.set_option("learned_relocation", 1)
    th = O.mlp_init(3)
    th[5248:5504] *= 4.0                       # logits and offsets well away from zero
    tr.proposal_params(th)
    tr.set_option("relocation_mode", 2)
    tr.step()
    tr.step()
    stats = tr.proposal_read("stats")
    before = _state(tr)
    host = _host_state(before, 3, cap=cap)
    info = tr.run_mcmc(500)
    feats = tr.proposal_read("features", np.uint16).reshape(-1, 16)[:n]
    out = np.ascontiguousarray(tr.proposal_read("out").reshape(-1, 4)[:n])
    # features: one bf16 rounding step of the host recomputation
    ref = _features_ref(before, stats, views=2)
    got = _bf16_to_f32(feats)
    bad = np.abs(got - ref) > np.abs(ref) * 2.0 ** -7 + 1e-30
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], ref[bad][:5])
    # network outputs: the oracle's network on the same features
    ref_out = O.mlp_forward(th, feats)
    assert float(np.abs(out - ref_out).max()) <= 1e-2 * max(1.0, float(np.abs(ref_out).max()))
    # the learned draw and placement: the oracle's sampler fed the same outputs
    info_o = (C.c_int32 * 6)()
    O.lib().oracle_mcmc_step_learned(host.struct(), cap, 3.0, 0, out.ctypes.data, info_o)
    want = dict(zip(("num_active", "dead", "grown", "nonfinite", "culled", "relocation_count"), list(info_o)))
    assert info == want
    n2 = info["num_active"]
    after = _state(tr)
    for k in GROUPS:
        got_p = after[k].reshape(n2, -1)
        want_p = host.p[k][:n2].reshape(n2, -1)
        if k in ("scales", "opac"):
            assert_rel(k, got_p, want_p, 1e-6)
        elif k == "means":
            # copies move by exp(max log-scale) * offset; the log-scales went through Eq. 9 (1e-6)
            assert_rel(k, got_p, want_p, 1e-5)
        else:
            assert_exact(k, got_p, want_p)
    p = tr.proposal_info()
    assert p["learned_events"] == 1 and p["events"] == 1


@pytest.mark.parametrize("batched", [False, True], ids=["single_view", "batched"])
def test_proposal_trains_online_and_is_judged(batched):
    W, H = 96, 80
    g, hidden, _ = small_scene(n=3000, n_hidden=1500, W=W, H=H)
    ds, _ = _dataset(g, hidden, W, H, n_cams=8, test_every=4)
    steps = 1250
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1500, cap_max=6000))
    tr.set_option("learned_relocation", 1)
    for _ in range(steps):
        tr.step_batch(2) if batched else tr.step()
    p = tr.proposal_info()
    print({k: round(v, 5) for k, v in p.items()})
    events = len([s for s in range(500, steps + 1, 100) if s < int(np.float32(1500) * np.float32(0.833))])
    assert p["events"] == events
    assert p["learned_events"] + p["plain_events"] == events
    assert 1 <= p["trained"] <= events - 1        # events after the first train on the outcomes
    assert p["samples"] > 0 and p["learned_events"] >= 1
    assert np.isfinite(p["val_bce"]) and np.isfinite(p["base_bce"]) and p["base_bce"] > 0
    assert np.isfinite(p["heldout_loss"]) and p["heldout_loss"] > 0

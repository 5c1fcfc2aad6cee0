"""The learned boundary end to end at a small size: training data from
wide-sponge shots, GPU training, then one shot on a region with the thin
learned layer, with no absorbing boundary and with the 40-point sponge,
against the 2x-wide FP64 run.  The trained layer must be stable (also over
a 4000-step run) and must reflect markedly less than the bare thin grid."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_layer_beats_the_bare_thin_grid(gpu):
    import learned
    print("\n[lb-pipeline] training data", flush=True)
    lb = gpu.LearnedBoundary(seed=3)
    counts = learned.training_data(lb, n_shots=12, region=32, margin=8, steps=300, collect_every=5, keep=0.05,
                                   capacity=400000)
    assert np.all(counts > 1000)
    print("[lb-pipeline] training", flush=True)
    l0, _ = lb.eval(with_grad=False)
    learned.train(lb, steps=1500, batch_per_face=2048, lr=2e-3, lr_final=2e-4)
    l1, _ = lb.eval(with_grad=False)
    assert l1 < 0.2 * l0
    print("[lb-pipeline] comparison", flush=True)
    res = learned.compare(lb, n=48, steps=500, f0=12.0)
    print(res, flush=True)
    assert res["learned"]["finite"] and res["none"]["finite"] and res["sponge40"]["finite"]
    assert res["learned"]["trace_rel_l2"] < 0.5 * res["none"]["trace_rel_l2"]
    # a long run: the layer (its output clamped to the largest missing term a
    # field no larger beyond the edge than inside can give) keeps the field
    # below the bare reflecting grid's amplitude over 4000 steps
    print("[lb-pipeline] 4000-step rollout", flush=True)
    ro = learned.rollout(lb, n=48, steps=4000, every=500, f0=12.0)
    print(ro, flush=True)
    assert np.all(np.isfinite(ro["learned"]))
    assert max(ro["learned"][2:]) <= max(ro["none"][2:])
    assert ro["learned"][-1] < ro["learned"][0]

"""GPU: the learned detector in front of the plain one, end to end.

Scenes of 1-40 bursts (CW, 2/4-FSK, OFDM-like, chirp) through the fused
chain; labels from the plain detector on the noise-free copies; the network
trained on the GPU by its own kernels; the proposal threshold picked on
validation scenes; then, on held-out scenes, the propose-then-confirm
decision against one plain decision per row: it keeps the plain method's
recall (98% of it, with a margin for this small set) with better precision,
in a fraction of its time.
"""
import json

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_learned_detector_recall_and_speed(dev):
    from revenant_amd import proposals

    res = proposals.run(dev, n_train=24, n_val=3, n_test=3, steps=1500, seed=3)
    print(json.dumps(res))
    assert res["loss_last"] < res["loss_first"]
    assert res["events"] >= 10
    assert res["recall_learned"] >= 0.9 * res["recall_plain"], res
    assert res["precision_learned"] > res["precision_plain"], res
    assert res["speedup"] > 5.0, res

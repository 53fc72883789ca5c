"""GPU: the learned detector in front of the plain one, end to end.

Scenes of 1-40 bursts (CW, 2/4-FSK, OFDM-like, chirp) through the fused
chain; labels from the plain detector on the noise-free copies; the network
trained on the GPU by its own kernels; the proposal threshold picked on
validation scenes; then, on held-out scenes, the propose-then-confirm
decision against one plain decision per row: it keeps the plain method's
recall with better precision, in a fraction of its time. The run is the
benchmark's own (96 training scenes, 6000 steps, 12 held-out scenes, seed 1),
and training is deterministic, so the bound holds run to run.
"""
import json

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_learned_detector_recall_and_speed(dev):
    from revenant_amd import proposals

    res = proposals.run(dev)
    print(json.dumps(res))
    assert res["loss_last"] < res["loss_first"]
    assert res["events"] >= 100
    assert res["recall_learned"] >= 0.98 * res["recall_plain"], res
    assert res["precision_learned"] > res["precision_plain"], res
    assert res["speedup"] > 5.0, res

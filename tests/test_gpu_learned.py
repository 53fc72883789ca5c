"""GPU: the learned proposal network's kernels against the CPU oracle.

Forward logits, loss and every layer's weight and bias gradients of the
matrix-core implementation equal the float64 oracle (oracle/cnn.py), which
rounds to bf16 at the same points; the remaining difference is the fp32
accumulation on the GPU (tolerances below). Then the network learns a toy
segmentation from its own gradients.
"""
import numpy as np
import pytest

from helpers import rng, to_dev
from oracle import cnn

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


@pytest.mark.parametrize("h,w,batch", [(24, 128, 2), (8, 64, 1), (17, 70, 3)])
def test_net_matches_oracle(dev, h, w, batch):
    from revenant_amd.learned import Net, split_params, to_bf16

    g = rng(0x4C45524E00000000 + h * 1000 + w)
    tiles = g.standard_normal((batch, h, w)).astype(np.float32)
    labels = (g.random((batch, h, w)) < 0.2).astype(np.float32)
    net = Net(dev, h, w, batch, seed=7)
    flat = net.params()
    # non-zero biases so the bias path is exercised
    for wl, bl in split_params(flat):
        bl[:] = g.standard_normal(bl.size).astype(np.float32) * 0.1
    net.set_params(flat)
    tb = to_bf16(dev, to_dev(tiles))
    loss = net.train_step(tb, to_dev(labels), lr=0.0, pos_weight=3.0)
    logits = net.forward(tb).cpu().numpy()
    grads = split_params(net.grads())
    layers = [(wl.astype(np.float64), bl.astype(np.float64)) for wl, bl in split_params(flat)]
    ref_loss, ref_z, ref_grads = cnn.loss_and_grads(layers, tiles.astype(np.float64), labels,
                                                    3.0)
    assert np.max(np.abs(logits - ref_z)) <= 1e-3 * max(1.0, np.max(np.abs(ref_z)))
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    for i, ((gw, gb), (rw, rb)) in enumerate(zip(grads, ref_grads)):
        assert _rel(gw, rw) < 2e-2, f"layer {i} weight gradient {_rel(gw, rw)}"
        assert _rel(gb, rb) < 2e-2, f"layer {i} bias gradient {_rel(gb, rb)}"
    net.close()


def test_net_learns_boxes(dev):
    """Bright rectangles on noise: after 150 Adam steps the network segments
    them (loss falls fourfold, IoU of the thresholded output > 0.8)."""
    from revenant_amd.learned import Net, to_bf16

    g = rng(0x4C45524E00000100)
    h = w = 64
    batch = 8

    def make(n):
        x = g.standard_normal((n, h, w)).astype(np.float32)
        y = np.zeros((n, h, w), dtype=np.float32)
        for i in range(n):
            for _ in range(int(g.integers(1, 4))):
                r0, c0 = int(g.integers(0, h - 8)), int(g.integers(0, w - 8))
                hh, ww = int(g.integers(4, 20)), int(g.integers(2, 16))
                x[i, r0:r0 + hh, c0:c0 + ww] += 2.5
                y[i, r0:r0 + hh, c0:c0 + ww] = 1.0
        return x, y

    net = Net(dev, h, w, batch, seed=3)
    first = None
    for step in range(150):
        x, y = make(batch)
        loss = net.train_step(to_bf16(dev, to_dev(x)), to_dev(y), lr=3e-3, pos_weight=2.0)
        first = loss if first is None else first
    x, y = make(batch)
    z = net.forward(to_bf16(dev, to_dev(x))).cpu().numpy()
    pred = z > 0
    iou = (pred & (y > 0)).sum() / max(1, (pred | (y > 0)).sum())
    assert loss < first / 4, (first, loss)
    assert iou > 0.8, iou
    net.close()


def test_training_is_deterministic(dev):
    """One seed, one data set: two trainings end with bit-identical weights
    (the kernels reduce in a fixed order and the batches come from a seeded
    generator), and a different seed does not."""
    from revenant_amd.learned import Net, to_bf16

    g = rng(0x4C45524E44455431)
    tiles = to_bf16(dev, to_dev(g.standard_normal((6, 32, 64)).astype(np.float32)))
    labels = to_dev((g.random((6, 32, 64)) < 0.2).astype(np.float32))

    def train(seed):
        net = Net(dev, 32, 64, 4, seed=seed)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        for step in range(60):
            idx = torch.randint(0, 6, (4,), generator=gen, device="cuda")
            net.train_step(tiles[idx].contiguous(), labels[idx].contiguous(), 2e-3, 3.0,
                           want_loss=step % 20 == 0)
        p = net.params()
        net.close()
        return p

    a, b, c = train(5), train(5), train(6)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(a, c)

"""The learned detector in front of the plain one: dataset, training and the
propose-then-confirm decision.

Spectrogram: the fused chain's frame partials (|X|^2 averaged over 16 blocks,
alpha = 1/16) of a 4096-channel, 2x-oversampled grid at 122.88 MS/s, one row
per 16 blocks (267 us), channels in frequency order. Tiles are 256 rows x 256
channels of log-power relative to the scene's noise floor (units of 10 dB),
bf16.

Labels: the plain detector (rvx_detect) run on every row of the noise-free
copy of the scene plus the noisy copy's floor level, growth stopping where
the signal falls below that floor (edge_floor_fraction = 1), windows up to 64
channels: the bins of each row that a plain decision reports once noise is out
of the way. Events whose mean power is more than 10 dB under the floor (filter
skirts, splatter) are dropped: no noisy row can show them.

Plain method: one plain decision per noisy row. Learned method: the network
proposes pixels; the plain detector confirms them with ONE decision per block
of 256 rows, on the spectrum averaged over the rows each channel was proposed
in (all rows where none was), searching only the proposed channels. A
proposal survives where a confirmed band covers it.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, scenes
from .detect import Detector, config
from .dsp import FORMAT_CS16, GridParams, build_twiddles, design_prototype
from .learned import Net, to_bf16
from .ops import Device

RATE = 122_880_000
M, TAPS, D = 4096, 16, 2048
FRAME = 16                      # blocks per spectrogram row
ROWS = 256
TILE = 256
FIRST_BLOCK = (M * TAPS - 1 + D - 1) // D
SAMPLES = (FIRST_BLOCK + ROWS * FRAME) * D + 1
SPAN_S = ROWS * FRAME * D / RATE
NOISE_SIGMA = 1000.0 / 32768.0  # per component, about -30 dBFS
BIN_HZ = RATE / M


class Spectrogram:
    """Rows of frame partials from the fused chain for one scene's samples."""

    def __init__(self, dev: Device):
        self.dev = dev
        grid = GridParams(M, TAPS, D)
        self.taps = torch.from_numpy(design_prototype(grid, beta=8.0).taps).cuda()
        self.tw = torch.from_numpy(build_twiddles(M).view(np.float32)).cuda()
        self.blocks = ROWS * FRAME
        self.ring = torch.empty(2 * M * self.blocks, dtype=torch.float32, device="cuda")
        self.partial = torch.empty(ROWS * M, dtype=torch.float32, device="cuda")
        self.desc = _lib.ChainDesc(FORMAT_CS16, M, TAPS, D, 0, SAMPLES, FIRST_BLOCK, self.blocks,
                                   self.blocks, FRAME, np.float32(1.0 / 16.0))

    def rows(self, x: torch.Tensor) -> torch.Tensor:
        """[ROWS, M] float32, channels in frequency order (lowest first)."""
        codes = scenes.to_cs16(x)
        self.dev.chain(self.desc, codes, self.taps, self.tw, self.ring, self.partial)
        return torch.roll(self.partial.view(ROWS, M), M // 2, dims=1).clone()


@dataclass
class SceneData:
    tiles: torch.Tensor     # [M // TILE, TILE rows, TILE] bf16 bit patterns (int16)
    labels: torch.Tensor    # [ROWS, M] float32 0/1
    noisy: torch.Tensor     # [ROWS, M] float32 linear power
    floor: float


VISIBLE = 0.1  # an event's mean signal power must reach -10 dB of the floor


def label_config():
    return config(M, BIN_HZ, 6.0, edge_floor_fraction=1.0, max_sum_bins=64)


def plain_config():
    return config(M, BIN_HZ, 6.0)


def make_scene(dev: Device, spec: Spectrogram, labeler: Detector, seed: int) -> SceneData:
    bursts = scenes.draw_scene(seed, RATE, SPAN_S)
    noisy = spec.rows(scenes.render(bursts, RATE, SAMPLES, NOISE_SIGMA, True, seed))
    clean = spec.rows(scenes.render(bursts, RATE, SAMPLES, NOISE_SIGMA, False, seed))
    floor = float(noisy.median())
    labels = torch.zeros(ROWS, M, dtype=torch.float32)
    base = (clean.double() + floor)
    for r in range(ROWS):
        bands, _ = labeler.decide(base[r].contiguous())
        for s, w, _ in bands:
            labels[r, s:s + w] = 1.0
    # Keep the events a noisy row could show at all: on a noise-free copy the
    # ladder also finds splatter and filter skirts tens of dB under the noise.
    lab, n = events(labels.numpy())
    if n:
        power = (clean / floor).cpu().numpy().ravel()
        size = np.bincount(lab.ravel(), minlength=n + 1)
        total = np.bincount(lab.ravel(), weights=power, minlength=n + 1)
        keep = total >= VISIBLE * np.maximum(size, 1)
        keep[0] = False
        labels = torch.from_numpy(keep[lab].astype(np.float32))
    logp = torch.log10(torch.clamp(noisy / floor, min=1e-6))
    tiles = logp.view(ROWS, M // TILE, TILE).permute(1, 0, 2).contiguous()
    return SceneData(to_bf16(dev, tiles), labels.cuda(), noisy, floor)


def tile_labels(s: SceneData) -> torch.Tensor:
    return s.labels.view(ROWS, M // TILE, TILE).permute(1, 0, 2).contiguous()


def train(dev: Device, data: list[SceneData], steps: int, batch: int = 8, lr: float = 1e-3,
          seed: int = 1, pos_weight: float = 4.0) -> tuple[Net, list[float]]:
    net = Net(dev, TILE, TILE, max(batch, M // TILE), seed=seed)
    tiles = torch.cat([s.tiles for s in data])
    labels = torch.cat([tile_labels(s) for s in data])
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    losses = []
    for step in range(steps):
        idx = torch.randint(0, tiles.shape[0], (batch,), generator=g, device="cuda")
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps))  # cosine decay
        loss = net.train_step(tiles[idx].contiguous(), labels[idx].contiguous(), float(rate),
                              pos_weight, want_loss=(step % 25 == 0 or step == steps - 1))
        if loss is not None:
            losses.append(loss)
    return net, losses


def propose(net: Net, s: SceneData, threshold: float = 0.0) -> torch.Tensor:
    """[ROWS, M] bool proposal mask: logits above `threshold`."""
    z = net.forward(s.tiles)  # [M // TILE, ROWS, TILE]
    return (z > threshold).permute(1, 0, 2).reshape(ROWS, M)


def confirm(det: Detector, s: SceneData, prop: torch.Tensor, dilate: int = 2) -> torch.Tensor:
    """One plain decision over the proposal-gated average, searching only the
    proposed channels; returns the confirmed proposal mask."""
    cols = prop.any(dim=0)
    count = prop.sum(dim=0)
    gated = torch.where(cols, (s.noisy * prop).sum(dim=0) / count.clamp(min=1),
                        s.noisy.mean(dim=0))
    allow = cols.clone()
    for k in range(1, dilate + 1):
        allow[k:] |= cols[:-k]
        allow[:-k] |= cols[k:]
    prefix = torch.zeros(M + 1, dtype=torch.int32, device="cuda")
    prefix[1:] = torch.cumsum(allow.int(), 0).int()
    bands, _ = det.decide(gated.double().contiguous(), prefix)
    ok = np.zeros(M, dtype=bool)
    for b0, w, _ in bands:
        ok[b0:b0 + w] = True
    return prop & torch.from_numpy(ok).to(prop.device)


def plain_rows(det: Detector, s: SceneData) -> torch.Tensor:
    """One plain decision per noisy row."""
    out = np.zeros((ROWS, M), dtype=bool)
    rows = s.noisy.double()
    for r in range(ROWS):
        bands, _ = det.decide(rows[r].contiguous())
        for b0, w, _ in bands:
            out[r, b0:b0 + w] = True
    return torch.from_numpy(out).cuda()


def plain_block(det: Detector, s: SceneData) -> torch.Tensor:
    """The reference's cadence (core/detect/detector.h DetectorConfig: a
    decision every 0.1 s on an exponential average over 1 s): a 256-row
    block spans 68 ms, less than one averaging time, so the plain method
    makes one decision per block on the block's mean spectrum; every row of a
    reported band counts as found."""
    bands, _ = det.decide(s.noisy.mean(dim=0).double().contiguous())
    ok = np.zeros(M, dtype=bool)
    for b0, w, _ in bands:
        ok[b0:b0 + w] = True
    return torch.from_numpy(np.broadcast_to(ok, (ROWS, M)).copy()).cuda()


def events(labels: np.ndarray):
    """Bursts: the label mask closed by one row in time and two channels in
    frequency (a burst's rows and the gaps in its own spectrum are one event),
    each event keeping only its labelled pixels."""
    from scipy import ndimage

    on = labels > 0.5
    grown = ndimage.binary_dilation(on, structure=np.ones((3, 5), dtype=bool))
    lab, n = ndimage.label(grown)
    return np.where(on, lab, 0), n


def recall(mask: np.ndarray, lab: np.ndarray, n: int, cover: float = 0.25) -> float:
    """Fraction of events found. A reported region counts only where at least
    half of it lies on labelled signal (two pixels of slack), so scattered
    false alarms earn nothing; an event is found when such regions cover a
    quarter of its labelled pixels."""
    from scipy import ndimage

    if n == 0:
        return 1.0
    comp, nc = ndimage.label(mask)
    near = ndimage.binary_dilation(lab > 0, iterations=2)
    if nc:
        size = np.bincount(comp.ravel(), minlength=nc + 1)
        inside = np.bincount(comp.ravel(), weights=near.ravel().astype(np.float64),
                             minlength=nc + 1)
        good = inside >= 0.5 * size
        good[0] = False
        credited = good[comp]
    else:
        credited = np.zeros_like(mask, dtype=bool)
    sizes = np.bincount(lab.ravel(), minlength=n + 1)
    hit = np.bincount(lab.ravel(), weights=credited.ravel().astype(np.float64), minlength=n + 1)
    return float(np.mean(hit[1:] >= cover * sizes[1:]))


def precision(mask: np.ndarray, labels: np.ndarray) -> float:
    """Fraction of reported pixels within two pixels of a labelled one."""
    from scipy import ndimage

    if not mask.any():
        return 1.0
    near = ndimage.binary_dilation(labels > 0.5, iterations=2)
    return float((mask & near).sum() / mask.sum())


def evaluate(dev: Device, net: Net, data: list[SceneData], reps: int = 3,
             threshold: float = 0.0, dilate: int = 2) -> dict:
    """Recall / precision of both methods over the scenes' events, and the time
    each takes per 256-row block of the 4096-channel spectrogram."""
    plain_det = Detector(dev, plain_config())
    r_plain, r_learned, p_plain, p_learned, n_events = [], [], [], [], 0
    r_raw, p_raw = [], []
    t_plain, t_learned = [], []
    r_block, p_block, t_block = [], [], []
    for s in data:
        lab, n = events(s.labels.cpu().numpy())
        n_events += n
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pm = plain_rows(plain_det, s)
        torch.cuda.synchronize()
        t_plain.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bm = plain_block(plain_det, s)
        torch.cuda.synchronize()
        t_block.append(time.perf_counter() - t0)
        bm_np = bm.cpu().numpy()
        r_block.append(recall(bm_np, lab, n) * n)
        p_block.append(precision(bm_np, s.labels.cpu().numpy()))
        best = float("inf")
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lm = confirm(plain_det, s, propose(net, s, threshold), dilate)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        t_learned.append(best)
        raw = propose(net, s, threshold).cpu().numpy()
        pm_np, lm_np = pm.cpu().numpy(), lm.cpu().numpy()
        r_raw.append(recall(raw, lab, n) * n)
        p_raw.append(precision(raw, s.labels.cpu().numpy()))
        r_plain.append(recall(pm_np, lab, n) * n)
        r_learned.append(recall(lm_np, lab, n) * n)
        lbl = s.labels.cpu().numpy()
        p_plain.append(precision(pm_np, lbl))
        p_learned.append(precision(lm_np, lbl))
    plain_det.close()
    rp = sum(r_plain) / max(1, n_events)
    rl = sum(r_learned) / max(1, n_events)
    ms_plain = 1e3 * float(np.mean(t_plain))
    ms_learned = 1e3 * float(np.mean(t_learned))
    return {"events": n_events, "recall_plain": round(rp, 4), "recall_learned": round(rl, 4),
            "recall_ratio": round(rl / rp, 4) if rp else None,
            "recall_proposals": round(sum(r_raw) / max(1, n_events), 4),
            "precision_proposals": round(float(np.mean(p_raw)), 4),
            "precision_plain": round(float(np.mean(p_plain)), 4),
            "precision_learned": round(float(np.mean(p_learned)), 4),
            "ms_plain_per_block": round(ms_plain, 3), "ms_learned_per_block": round(ms_learned, 3),
            "speedup": round(ms_plain / ms_learned, 2),
            "plain_at_reference_cadence": {
                "what": "one plain decision per 256-row block on its mean spectrum",
                "recall": round(sum(r_block) / max(1, n_events), 4),
                "precision": round(float(np.mean(p_block)), 4),
                "ms_per_block": round(1e3 * float(np.mean(t_block)), 3)}}


def build(dev: Device, seeds) -> list[SceneData]:
    spec = Spectrogram(dev)
    labeler = Detector(dev, label_config())
    out = [make_scene(dev, spec, labeler, s) for s in seeds]
    labeler.close()
    return out


THRESHOLDS = (1.0, 0.0, -0.5, -1.0, -1.25, -1.5, -1.75, -2.0, -2.25, -2.5)


def run(dev: Device, n_train: int = 96, n_val: int = 8, n_test: int = 12, steps: int = 6000,
        seed: int = 1) -> dict:
    """Build scenes, train on the GPU, pick the proposal threshold on the
    validation scenes (the most precise one whose recall there reaches the
    plain method's), and report the held-out scenes."""
    t0 = time.perf_counter()
    train_s = build(dev, range(10_000 * seed, 10_000 * seed + n_train))
    val = build(dev, range(20_000 * seed, 20_000 * seed + n_val))
    test = build(dev, range(30_000 * seed, 30_000 * seed + n_test))
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, losses = train(dev, train_s, steps=steps, seed=seed)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    det = Detector(dev, plain_config())
    ev = [events(s.labels.cpu().numpy()) for s in val]
    nev = max(1, sum(n for _, n in ev))
    plain = [plain_rows(det, s).cpu().numpy() for s in val]
    rp = sum(recall(m, l, n) * n for m, (l, n) in zip(plain, ev)) / nev
    pick = THRESHOLDS[-1]
    for thr in THRESHOLDS:  # most precise first
        masks = [confirm(det, s, propose(net, s, thr)).cpu().numpy() for s in val]
        rl = sum(recall(m, l, n) * n for m, (l, n) in zip(masks, ev)) / nev
        if rl >= rp:
            pick = thr
            break
    det.close()
    res = evaluate(dev, net, test, threshold=pick)
    res.update({"threshold_logit": pick, "train_scenes": n_train, "train_steps": steps,
                "train_s": round(t_train, 2), "data_s": round(t_data, 2),
                "loss_first": round(losses[0], 4), "loss_last": round(losses[-1], 4),
                "params": net.count})
    net.close()
    return res

"""CPU oracle of the learned proposal network (TEST INFRASTRUCTURE: only
tests/ import it, to check the GPU kernels' forward pass and gradients).

float64 arithmetic with the GPU path's quantisation points restated: the
tiles, the weights, every ReLU output and every back-propagated gradient are
rounded to bf16 (round to nearest even) where the GPU kernels round them;
everything else is exact-ish float64, so the GPU's fp32 accumulation is the
only difference left (tolerances in the tests state it).
"""
from __future__ import annotations

import numpy as np


def bf16(x) -> np.ndarray:
    """Round to bf16 (nearest even) and back, through float32."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def conv3x3(x: np.ndarray, w: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    """'same' 3x3 convolution, x [N,H,W,Cin], w [Cout,3,3,Cin] -> [N,H,W,Cout]."""
    n, h, wd, cin = x.shape
    xp = np.zeros((n, h + 2, wd + 2, cin))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((n, h, wd, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            y += xp[:, ky:ky + h, kx:kx + wd, :] @ w[:, ky, kx, :].T
    return y if b is None else y + b


def forward(layers, tiles: np.ndarray):
    """layers: [(W, b)]; tiles [N,H,W]. Returns (activations, logits)."""
    a = bf16(tiles)[..., None]
    acts = [a]
    for i, (w, b) in enumerate(layers):
        z = conv3x3(a, bf16(w), b.astype(np.float64))
        if i == len(layers) - 1:
            return acts, z[..., 0]
        a = bf16(np.maximum(z, 0.0))
        acts.append(a)


def loss_and_grads(layers, tiles, labels, pos_weight):
    acts, z = forward(layers, tiles)
    count = z.size
    wgt = np.where(labels > 0.5, pos_weight, 1.0)
    loss = float(np.sum(wgt * (np.maximum(z, 0) - z * labels + np.log1p(np.exp(-np.abs(z)))))
                 / count)
    sig = 1.0 / (1.0 + np.exp(-z))
    dy = bf16(wgt * (sig - labels) / count)[..., None]
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        w, b = layers[i]
        a = acts[i]
        n, h, wd, cin = a.shape
        ap = np.zeros((n, h + 2, wd + 2, cin))
        ap[:, 1:-1, 1:-1] = a
        gw = np.zeros(w.shape)
        for ky in range(3):
            for kx in range(3):
                patch = ap[:, ky:ky + h, kx:kx + wd, :].reshape(-1, cin)
                gw[:, ky, kx, :] = dy.reshape(-1, dy.shape[-1]).T @ patch
        grads[i] = (gw, dy.sum(axis=(0, 1, 2)))
        if i == 0:
            break
        # data gradient: the transposed convolution, masked by the ReLU
        wf = bf16(w)[:, ::-1, ::-1, :].transpose(3, 1, 2, 0)  # [cin, 3, 3, cout]
        da = conv3x3(dy, wf, None)
        dy = bf16(np.where(a > 0, da, 0.0))
    return loss, z, grads

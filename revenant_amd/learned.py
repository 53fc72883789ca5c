"""The learned proposal network (rvx_net_*): forward, training and parameter
access for spectrogram tiles; proposals are confirmed by the plain detector."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .ops import Device, _p

LAYERS = [(1, 32), (32, 32), (32, 32), (32, 32), (32, 1)]  # (cin, cout), 3x3 each


def to_bf16(dev: Device, x: torch.Tensor) -> torch.Tensor:
    """fp32 CUDA tensor -> bf16 bit patterns (int16 tensor), round to nearest even."""
    x = x.contiguous().float()
    out = torch.empty(x.shape, dtype=torch.int16, device=x.device)
    _lib.check(dev.lib.rvx_to_bf16(dev.ctx, _p(x), x.numel(), _p(out)))
    return out


class Net:
    def __init__(self, dev: Device, height: int, width: int, max_batch: int, seed: int = 1):
        self.dev = dev
        self.h, self.w, self.max_batch = height, width, max_batch
        h = C.c_void_p()
        _lib.check(dev.lib.rvx_net_create(dev.ctx, height, width, max_batch, seed, C.byref(h)))
        self.handle = h
        self.count = dev.lib.rvx_net_param_count(h)

    def params(self) -> np.ndarray:
        a = np.empty(self.count, dtype=np.float32)
        _lib.check(self.dev.lib.rvx_net_params(self.handle, a.ctypes.data_as(_lib.fptr),
                                               a.size, 0))
        return a

    def set_params(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32)
        _lib.check(self.dev.lib.rvx_net_params(self.handle, a.ctypes.data_as(_lib.fptr),
                                               a.size, 1))

    def grads(self) -> np.ndarray:
        a = np.empty(self.count, dtype=np.float32)
        _lib.check(self.dev.lib.rvx_net_grads(self.handle, a.ctypes.data_as(_lib.fptr), a.size))
        return a

    def forward(self, tiles_bf16: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        b = tiles_bf16.shape[0]
        if out is None:
            out = torch.empty((b, self.h, self.w), dtype=torch.float32, device="cuda")
        _lib.check(self.dev.lib.rvx_net_forward(self.handle, _p(tiles_bf16), b, _p(out)))
        return out

    def train_step(self, tiles_bf16: torch.Tensor, labels: torch.Tensor, lr: float,
                   pos_weight: float = 1.0, want_loss: bool = True) -> float | None:
        b = tiles_bf16.shape[0]
        loss = C.c_float()
        _lib.check(self.dev.lib.rvx_net_train_step(self.handle, _p(tiles_bf16), _p(labels), b,
                                                   lr, pos_weight,
                                                   C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def close(self) -> None:
        if self.handle:
            self.dev.lib.rvx_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def split_params(flat: np.ndarray):
    """[(W[cout][3][3][cin], b[cout])] views of the flat parameter vector."""
    out, off = [], 0
    for cin, cout in LAYERS:
        w = flat[off:off + cout * 9 * cin].reshape(cout, 3, 3, cin)
        off += cout * 9 * cin
        b = flat[off:off + cout]
        off += cout
        out.append((w, b))
    return out

#!/usr/bin/env python3
"""Throughput of the plain receive chain on one MI355X.

Workload (the wideband configuration): 2^27 complex int16 IQ samples at
122.88 MS/s holding 400 random carriers (CW/AM/NFM/WFM, -80..-10 dBFS) over
AWGN; a 4096-channel, 2x-oversampled polyphase channelizer (16 taps per
branch, Kaiser beta 8 prototype, 4096-point radix-2 transform); |X|^2 averaged
with alpha = 1/16 over 4-block frames. One step = one fused launch over the
whole resident input (convert -> branch filter -> transform -> correction ->
channel ring + averaged spectrum) plus the frame recurrence.

Prints one JSON line. The timed region brackets exactly K steps with device
synchronisation; the kernel's own average duration is measured with HIP
events on the stream it is launched on. The CPU baseline times the oracle
chain (a C++ restatement of the reference's CPU twins) on a bounded sample of
the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "input MS/s thru PFB+256 VRX; % HBM peak; learned-detect speedup @ equal recall"
HBM_PEAK_GBS = 8000.0

RATE = 122_880_000
SAMPLES = 1 << 27
M, TAPS, D = 4096, 16, 2048
FRAME = 4        # blocks per averaging frame (one workgroup's chunk)
ROW_FRAMES = 4   # frames per output row of the averaged spectrum (16 blocks)
ALPHA = 1.0 / 16.0
SEED = 0x5743454E45000002


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 1 GiB device to device three times: the "
                         "known-size kernel that calibrates FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the chain shards by time block with a halo; see DESIGN.md)", file=sys.stderr)
        return 2

    import torch

    from revenant_amd import _lib, synth
    from revenant_amd.dsp import (FORMAT_CS16, GridParams, build_twiddles, design_prototype,
                                  ema_frame_decay, ema_params)
    from revenant_amd.ops import Device

    dev = Device(0)
    grid = GridParams(M, TAPS, D)
    proto = design_prototype(grid, beta=8.0)
    taps = torch.from_numpy(proto.taps).cuda()
    tw_np = build_twiddles(M)
    tw = torch.from_numpy(tw_np.view(np.float32)).cuda()

    print(f"[bench] rendering {SAMPLES} cs16 samples (400 carriers + AWGN) on the GPU",
          file=sys.stderr, flush=True)
    scene = synth.wideband_scene(RATE, 400, SEED)
    packed = torch.empty(4 * SAMPLES, dtype=torch.uint8, device="cuda")
    synth.render_device(dev, scene, RATE, 0, SAMPLES, noise_dbfs=-60.0, seed=SEED,
                        fmt=FORMAT_CS16, out=packed)
    torch.cuda.synchronize()

    N = grid.prototype_length
    first_block = (N - 1 + D - 1) // D               # first block with full history
    blocks = (SAMPLES - 1) // D - first_block + 1
    blocks -= blocks % (FRAME * ROW_FRAMES)
    ring_blocks = 1 << (blocks - 1).bit_length()
    frames = blocks // FRAME
    alpha = np.float32(ALPHA)
    decay = ema_frame_decay(alpha, FRAME)
    ring = torch.empty(2 * M * ring_blocks, dtype=torch.float32, device="cuda")
    partial = torch.empty(frames * M, dtype=torch.float32, device="cuda")
    state = torch.zeros(M, dtype=torch.float32, device="cuda")
    rows = torch.empty((frames // ROW_FRAMES) * M, dtype=torch.float32, device="cuda")
    desc = _lib.ChainDesc(FORMAT_CS16, M, TAPS, D, 0, SAMPLES, first_block, blocks, ring_blocks,
                          FRAME, alpha)
    ema = ema_params(M, ring_blocks, first_block, FRAME, frames, alpha, ROW_FRAMES)

    def step():
        dev.chain(desc, packed, taps, tw, ring, partial)
        dev.ema_combine(ema, partial, state, rows)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Kernel-only timing of the dominant launch (HIP events on its stream).
    dev.timer_begin()
    for _ in range(args.steps):
        dev.chain(desc, packed, taps, tw, ring, partial)
    chain_ms = dev.timer_end() / args.steps

    # Whole-step timing, bracketed by device synchronisation.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps

    in_samples = blocks * D
    value = in_samples / (ms_per_step * 1e-3) / 1e6
    # Algorithmic bytes per launch: each input sample read once (4 B cs16,
    # the unique window incl. the filter's history), the channel ring written
    # once (8 B per channel sample), the frame partials written once.
    unique_in = (blocks - 1) * D + N
    algo_bytes = 4 * unique_in + 8 * M * blocks + 4 * M * frames
    achieved = algo_bytes / (chain_ms * 1e-3) / 1e9

    # Parity spot check of this very run: two frames against the oracle. The
    # oracle sees the input from sample start*D on, i.e. block indices shifted
    # by `start`; start is a multiple of the frame (and of M/D), so the phase
    # correction and the frame alignment are unchanged by the shift.
    parity = "skipped"
    try:
        from oracle import oracle
        check = 2 * FRAME
        start = blocks // 2 - (blocks // 2) % FRAME
        hi = (first_block + start + check - 1) * D + 1
        sub = packed[start * D * 4:hi * 4].cpu().numpy()
        cap = 1 << (check - 1).bit_length()
        od = oracle.OChain(FORMAT_CS16, M, TAPS, D, first_block, check, cap, FRAME, alpha, decay)
        ring_cpu, part_cpu = oracle.chain(od, sub, proto.taps, tw_np, threads=4)
        ring_cpu = ring_cpu.reshape(M, cap, 2)
        g_ring = ring.view(M, ring_blocks, 2).cpu().numpy()
        ok = True
        for b in range(check):
            m = first_block + start + b
            col = ring_cpu[:, (first_block + b) & (cap - 1)]
            ok &= np.array_equal(g_ring[:, m & (ring_blocks - 1)].view(np.uint32),
                                 col.view(np.uint32))
        gp = partial.view(frames, M).cpu().numpy()[start // FRAME:start // FRAME + 2]
        ok &= np.array_equal(gp.view(np.uint32), part_cpu.reshape(2, M).view(np.uint32))
        parity = f"bit-exact vs oracle on {check} blocks" if ok else "MISMATCH vs oracle"
    except Exception as exc:  # the oracle is optional on the box
        parity = f"skipped ({type(exc).__name__}: {exc})"

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(packed, proto.taps, tw_np, first_block, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "MS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain chain: 2^27 cs16 @ 122.88 MS/s, 400 carriers + AWGN -> "
                        "4096-ch PFB (D=2048, 16 taps/branch, Kaiser beta 8) -> channel ring "
                        "+ |X|^2 EMA alpha=1/16 (4-block frames); VRX stage not in this line",
            "input_samples": int(in_samples),
            "blocks": int(blocks),
            "channels": M,
            "decimation": D,
            "taps_per_branch": TAPS,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
            "kernel": "chain_kernel<M=4096, V=8, cs16, 16 taps>",
            "kernel_ms": round(chain_ms, 4),
            "algorithmic_bytes": int(algo_bytes),
        },
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(line), flush=True)
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    return 0


def traffic_from_profiles():
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated against a copy kernel), as committed under profiles/."""
    p = ROOT / "profiles" / "chain_c2_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(packed, taps, tw, first_block, seconds):
    """The oracle chain (convert -> branch -> transform -> correction -> frame
    average, the reference twins' operation order) on this host's cores, over
    a bounded sample of the same workload, repeated to about `seconds`."""
    try:
        from oracle import oracle
    except Exception as exc:
        return {"value": None, "unit": "MS/s", "cores": 0, "kind": "port",
                "sample": f"oracle unavailable: {exc}"}
    from revenant_amd.dsp import FORMAT_CS16, ema_frame_decay

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    alpha = np.float32(ALPHA)
    decay = ema_frame_decay(alpha, FRAME)
    blocks = 8192 - 8192 % (FRAME * threads)
    hi = (first_block + blocks) * D + 1
    sub = packed[: hi * 4].cpu().numpy()
    cap = 1 << (blocks - 1).bit_length()
    od = oracle.OChain(FORMAT_CS16, M, TAPS, D, first_block, blocks, cap, FRAME, alpha, decay)
    ring = np.zeros(2 * M * cap, dtype=np.float32)
    partial = np.zeros((blocks // FRAME) * M, dtype=np.float32)
    passes, total = 0, 0.0
    while total < seconds or passes == 0:
        t0 = time.perf_counter()
        oracle.chain(od, sub, taps, tw, threads=threads, ring=ring, partial=partial)
        total += time.perf_counter() - t0
        passes += 1
    rate = passes * blocks * D / total / 1e6
    return {"value": round(rate, 2), "unit": "MS/s", "cores": threads, "kind": "port",
            "sample": f"{passes} passes over {blocks} blocks ({blocks * D} input samples) of "
                      f"the same workload, {total:.1f} s on {threads} threads"}


if __name__ == "__main__":
    sys.exit(main())

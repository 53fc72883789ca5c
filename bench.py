#!/usr/bin/env python3
"""Throughput of the receive chain with 256 receivers on one MI355X.

Workload (the batching configuration, on the wideband stream): 2^27 complex
int16 IQ samples at 122.88 MS/s holding 400 random carriers (CW/AM/NFM/WFM,
-80..-10 dBFS) over AWGN, resident in HBM. One step, captured once as a HIP
graph and replayed:
  1. the 4096-channel, 2x-oversampled polyphase channelizer (16 taps per
     branch, Kaiser beta 8, 4096-point radix-2 transform) fused with the
     conversion and the |X|^2 average (alpha = 1/16, 4-block frames): one
     launch writing the full channel ring and the frame partials;
  2. the averaged spectrum's frame recurrence;
  3. the 512-channel grid the 200 kHz WFM receivers need (a 200 kHz receiver
     does not fit a 60 kHz channel, core/engine/vrx_place.cpp:187-197),
     storing only the channels those receivers read;
  4. 256 receivers in one persistent launch: 64 NFM 12.5 kHz, 64 AM 10 kHz,
     32 WFM 200 kHz, 64 USB 2.7 kHz, 32 CW 500 Hz at seeded uniform offsets,
     each resampled to 48 kHz and demodulated to FP32 audio over the whole
     step (the WFM receivers as the planner makes them by default: FM stereo,
     L and R, with the 75 us de-emphasis folded into their audio filter);
     tiles picked by the autotuner before capture.
`value` is input MS/s through all four. `--workload plain` times steps 1-2
alone (the plain chain).

A second leg (not in the step; `--no-conditioning` skips it) times the stages
around the path on the same input: the front-end moments over the resident
codes and the channelizer with the front-end correction fused in, the four
impulse-excision passes and the comb canceller's two kernels on a 2^24-sample
block of the cf32 ring, and the carrier loops of 256 SAM/DSB receivers over a
second of 48 kHz audio each.

Prints one JSON line. The timed region brackets exactly K steps with device
synchronisation; each kernel's average duration is measured with HIP events
on the stream it is launched on. The CPU baseline times the oracle (a C++
restatement of the reference's CPU twins) over a bounded sample of the same
workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "input MS/s thru PFB+256 VRX; % HBM peak; learned-detect speedup @ equal recall"
HBM_PEAK_GBS = 8000.0

RATE = 122_880_000
SAMPLES = 1 << 27
M, TAPS, D = 4096, 16, 2048
WM, WD = 512, 256  # the wide receivers' grid
FRAME = 4        # blocks per averaging frame (one workgroup's chunk)
ROW_FRAMES = 4   # frames per output row of the averaged spectrum (16 blocks)
ALPHA = 1.0 / 16.0
SEED = 0x5743454E45000002
RX_SEED = 0x5652584241544348
AUDIO_RATE = 48000


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["vrx", "plain"], default="vrx")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default tiles, no autotuning")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-detector leg (scenes, training, recall and speed)")
    ap.add_argument("--no-conditioning", action="store_true",
                    help="skip the ring-conditioning and carrier-loop leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 1 GiB device to device three times: the "
                         "known-size kernel that calibrates FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Grid:
    """One channelizer over the resident input: descriptor, tables, ring."""

    def __init__(self, torch, channels, decimation, frame, rows=None):
        from revenant_amd import _lib
        from revenant_amd.dsp import (FORMAT_CS16, GridParams, build_twiddles,
                                      design_prototype, prescale_taps)

        self.params = GridParams(channels, TAPS, decimation)
        self.proto = design_prototype(self.params, beta=8.0)
        self.taps = torch.from_numpy(self.proto.taps).cuda()
        # the chain reads cs16 against taps carrying the conversion's 2^-15
        # (bit-identical products; rvx_prescale_taps checks exactness)
        self.taps_scaled = torch.from_numpy(prescale_taps(self.proto.taps)).cuda()
        self.tw_np = build_twiddles(channels)
        self.tw = torch.from_numpy(self.tw_np.view(np.float32)).cuda()
        N = self.params.prototype_length
        self.first_block = (N - 1 + decimation - 1) // decimation  # first with full history
        blocks = (SAMPLES - 1) // decimation - self.first_block + 1
        if frame:
            blocks -= blocks % (frame * ROW_FRAMES)
        self.blocks = blocks
        self.ring_blocks = 1 << (blocks - 1).bit_length()
        # rows padded off a power of two: one block's writes then spread over
        # the HBM channels instead of landing on a few
        self.pitch = self.ring_blocks + _lib.RING_PITCH_PAD
        self.unique_in = (blocks - 1) * decimation + N
        self.rows = rows
        n_rows = channels if rows is None else len(rows)
        self.ring = torch.empty(2 * n_rows * self.pitch, dtype=torch.float32, device="cuda")
        self.desc = _lib.ChainDesc(FORMAT_CS16, channels, TAPS, decimation, 0, SAMPLES,
                                   self.first_block, blocks, self.ring_blocks, frame,
                                   np.float32(ALPHA))
        self.desc.taps_scaled = 1
        self.desc.ring_pitch = self.pitch
        self.row_table = None
        if rows is not None:
            table = np.full(channels, 0xFFFFFFFF, dtype=np.uint32)
            for r, k in enumerate(rows):
                table[k] = r
            self.row_table = torch.from_numpy(table.view(np.int32)).cuda()
            self.desc.ring_rows = len(rows)
            self.desc.channel_rows = self.row_table.data_ptr()
        self.frames = blocks // frame if frame else 0
        # algorithmic bytes per launch: the unique input window read once, each
        # stored channel sample written once, the frame partials written once
        self.algo_bytes = 4 * self.unique_in + 8 * n_rows * blocks + 4 * channels * self.frames


def build_receivers(main: Grid):
    """The 256-receiver mix, placed and planned; returns (receivers, wide channels)."""
    from revenant_amd.dsp import GridParams
    from revenant_amd.receivers import standard_mix

    return standard_mix(RATE, main.params, GridParams(WM, TAPS, WD), RX_SEED, AUDIO_RATE)


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the chain shards by time block with a halo, the receivers by receiver)",
              file=sys.stderr)
        return 2

    import torch

    from revenant_amd import synth
    from revenant_amd.dsp import FORMAT_CS16, build_nco_table, ema_frame_decay, ema_params
    from revenant_amd.ops import Device
    from revenant_amd.receivers import Receiver, ReceiverBatch

    dev = Device(0)
    side = torch.cuda.Stream()  # default priority; graphs capture a non-null stream
    dev.bind_stream(side)
    vrx = args.workload == "vrx"

    with torch.cuda.stream(side):
        log(f"rendering {SAMPLES} cs16 samples (400 carriers + AWGN) on the GPU")
        scene = synth.wideband_scene(RATE, 400, SEED)
        packed = torch.empty(4 * SAMPLES, dtype=torch.uint8, device="cuda")
        synth.render_device(dev, scene, RATE, 0, SAMPLES, noise_dbfs=-60.0, seed=SEED,
                            fmt=FORMAT_CS16, out=packed)

        main_grid = Grid(torch, M, D, FRAME)
        alpha = np.float32(ALPHA)
        frames = main_grid.frames
        partial = torch.empty(frames * M, dtype=torch.float32, device="cuda")
        state = torch.zeros(M, dtype=torch.float32, device="cuda")
        rows_out = torch.empty((frames // ROW_FRAMES) * M, dtype=torch.float32, device="cuda")
        ema = ema_params(M, main_grid.ring_blocks, main_grid.first_block, FRAME, frames, alpha,
                         ROW_FRAMES)

        launches = {}
        launches["chain_m4096"] = lambda: dev.chain(main_grid.desc, packed,
                                                    main_grid.taps_scaled, main_grid.tw,
                                                    main_grid.ring, partial)
        launches["ema_combine"] = lambda: dev.ema_combine(ema, partial, state, rows_out)

        batch = wide_grid = None
        tiles = None
        audio_total = 0
        if vrx:
            plans, wide_set = build_receivers(main_grid)
            wide_grid = Grid(torch, WM, WD, 0, rows=wide_set)
            nco = torch.from_numpy(build_nco_table(16).view(np.float32)).cuda()
            receivers, starts, counts = [], [], []
            for name, p in plans:
                if name == "wfm":
                    g, row = wide_grid, wide_set.index(p.plan.placement.channel)
                else:
                    g, row = main_grid, p.plan.placement.channel
                rx = Receiver(p, row * g.pitch, g.ring_blocks - 1, g.ring)
                a0, n = rx.audio_range(g.first_block, g.first_block + g.blocks - 1)
                receivers.append(rx)
                starts.append(a0)
                counts.append(n)
            sizes = [n * r.plan.demod.channels for n, r in zip(counts, receivers)]
            audio_buf = torch.empty(sum(sizes), dtype=torch.float32, device="cuda")
            offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
            audio = [audio_buf[offs[i]:offs[i + 1]] for i in range(len(receivers))]
            audio_total = int(sum(counts))
            batch = ReceiverBatch(dev, receivers, nco)
            launches["chain_m512_wfm"] = lambda: dev.chain(wide_grid.desc, packed,
                                                           wide_grid.taps_scaled, wide_grid.tw,
                                                           wide_grid.ring)
            # the receivers read channel rings that the chains fill
            launches["chain_m512_wfm"]()
            launches["chain_m4096"]()
            batch.prepare(starts, counts, audio)
            if not args.no_tune:
                dev.timer_begin()
                batch.launch()
                before = dev.timer_end()
                tiles = batch.tune(starts, counts, audio)
                dev.timer_begin()
                batch.launch()
                after = dev.timer_end()
                chosen = sorted({(p.plan.demod.mode, t) for (n, p), t in zip(plans, tiles)})
                log(f"tuned receiver tiles (mode, tile) {chosen}: {before:.3f} -> {after:.3f} ms")
            launches["vrx_batch"] = batch.launch
        side.synchronize()

        # One step, captured once.
        log("capturing the step as a HIP graph")
        order = ["chain_m4096", "ema_combine"] + (["chain_m512_wfm", "vrx_batch"] if vrx else [])
        for name in order:  # eager warm launch of every kernel before capture
            launches[name]()
        side.synchronize()
        aux = torch.cuda.Stream() if vrx else None  # default priority, as every stream here
        dev.graph_begin()
        launches["chain_m4096"]()
        if vrx:
            # The 512-channel chain reads only the input: it runs on a second
            # captured stream beside the averaged spectrum's combine (a
            # latency-bound recurrence over 64 waves), and the receivers wait
            # for both. (Beside the 4096-channel chain instead, the two chains
            # contend and the step is slower: 4.77 vs 4.64 ms.)
            fork = torch.cuda.Event()
            fork.record(side)
            aux.wait_event(fork)
            dev.bind_stream(aux)
            launches["chain_m512_wfm"]()
            join = torch.cuda.Event()
            join.record(aux)
            dev.bind_stream(side)
        launches["ema_combine"]()
        if vrx:
            side.wait_event(join)
            launches["vrx_batch"]()
        graph = dev.graph_end()

        for _ in range(args.warmup):
            graph.launch()
        side.synchronize()

        # Kernel-only timing of each launch (HIP events on the launch stream).
        kernel_ms = {}
        for name in order:
            dev.timer_begin()
            for _ in range(args.steps):
                launches[name]()
            kernel_ms[name] = dev.timer_end() / args.steps

        # The matrix-core forms of the same 4096-channel launch (not in the
        # step: their outputs are not bit-exact), timed the same way.
        from revenant_amd import _lib as L
        mfma_ms = {}
        for label, prec in (("bf16x3", L.PREC_BF16X3),):
            mdesc = _lib_copy(main_grid.desc)
            mdesc.taps_scaled = 0

            def run(prec=prec, mdesc=mdesc):
                dev.chain_mfma(mdesc, prec, packed, main_grid.taps, main_grid.tw,
                               main_grid.ring, partial)
            run()
            dev.timer_begin()
            for _ in range(args.steps):
                run()
            mfma_ms[label] = dev.timer_end() / args.steps
        launches["chain_m4096"]()  # leave the exact ring behind for the parity check

        # The receivers recorded the reference's way (fine + demod launches per
        # receiver segment through a fine ring), captured as one graph so the
        # host's launch rate does not count; same audio bits.
        unbatched_ms = None
        if vrx:
            fine_ring = torch.empty(2 << 17, dtype=torch.float32, device="cuda")
            dev.graph_begin()
            batch.unbatched(fine_ring)
            ug = dev.graph_end()
            ug.launch()
            dev.timer_begin()
            for _ in range(args.steps):
                ug.launch()
            unbatched_ms = dev.timer_end() / args.steps
            ug.close()
            batch.launch()  # the batch's own audio for the parity check

        # Whole-step timing, bracketed by device synchronisation.
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.launch()
        torch.cuda.synchronize()
        ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
        graph.close()

        engine = None
        if vrx:
            log("the same step through the streaming engine (rvx_engine_*)")
            engine = engine_leg(dev, torch, packed, plans, wide_set, tiles, starts, audio,
                                args.steps, args.warmup, ms_per_step)

    in_samples = main_grid.blocks * D
    value = in_samples / (ms_per_step * 1e-3) / 1e6

    algo = {"chain_m4096": main_grid.algo_bytes,
            "ema_combine": 4 * M * frames + 4 * M * (frames // ROW_FRAMES)}
    if vrx:
        algo["chain_m512_wfm"] = wide_grid.algo_bytes
        # each receiver reads its channel span once and writes its audio once
        rx_bytes = 0
        for rx, n in zip(batch.receivers, counts):
            p = rx.plan
            fine = (n - 1) * p.demod.decimation + p.history + 1
            chan = fine * p.plan.channel_rate // p.plan.demod_rate + p.fine.taps
            rx_bytes += 8 * chan + 4 * n * p.demod.channels
        algo["vrx_batch"] = rx_bytes
    stages = {k: {"ms": round(kernel_ms[k], 4),
                  "GB/s": round(algo[k] / (kernel_ms[k] * 1e-3) / 1e9, 1)} for k in order}
    dominant = max(order, key=lambda k: kernel_ms[k])
    achieved = algo[dominant] / (kernel_ms[dominant] * 1e-3) / 1e9
    valu = valu_roof(kernel_ms, main_grid, wide_grid if vrx else None)

    with torch.cuda.stream(side):
        precision = precision_in_run(dev, torch)
    parity = spot_check(torch, packed, main_grid, partial, frames, alpha,
                        batch if vrx else None, starts if vrx else None,
                        counts if vrx else None, audio if vrx else None)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(packed, main_grid, wide_grid, batch, args.cpu_seconds)

    learned = None
    if not args.no_learned:
        log("learned detector: scenes, training on the GPU, held-out evaluation")
        from revenant_amd import proposals
        with torch.cuda.stream(side):
            learned = proposals.run(dev)
            if not args.no_cpu:
                learned.update(plain_twin_per_block(dev, proposals))
                learned["speedup_vs_cpu_twin"] = round(
                    learned["ms_plain_cpu_twin_per_block"] / learned["ms_learned_per_block"], 1)

    log("timing the plain detector's decision at 65,536 bins")
    with torch.cuda.stream(side):
        decision = plain_decision(dev, torch, cpu=not args.no_cpu)

    conditioning = None
    if not args.no_conditioning:
        log("timing the ring-conditioning stages and the carrier loops")
        conditioning = conditioning_leg(dev, torch, packed, main_grid, partial, kernel_ms,
                                        args.steps, cpu=not args.no_cpu)

    host = None
    if not args.no_conditioning:
        log("timing the host stages (noise profile, comb detection and bank)")
        host = host_leg()

    workload = ("PFB + 256 VRX: 2^27 cs16 @ 122.88 MS/s (400 carriers + AWGN) -> 4096-ch PFB "
                "(D=2048, 16 taps/branch, Kaiser beta 8) -> channel ring + |X|^2 EMA "
                "alpha=1/16; 512-ch PFB (D=256) for the WFM receivers; 256 VRX "
                "(64 NFM 12.5k, 64 AM 10k, 32 WFM 200k stereo + 75us de-emphasis, 64 USB 2.7k, "
                "32 CW 500) -> 48 kHz "
                "FP32 audio, one persistent launch; the step is one HIP graph") if vrx else (
                "plain chain: 2^27 cs16 @ 122.88 MS/s, 400 carriers + AWGN -> 4096-ch PFB "
                "(D=2048, 16 taps/branch, Kaiser beta 8) -> channel ring + |X|^2 EMA "
                "alpha=1/16 (4-block frames)")
    traffic = traffic_from_profiles(dominant)
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
            "workload": workload,
            "input_samples": int(in_samples),
            "channels": M,
            "decimation": D,
            "taps_per_branch": TAPS,
            "receivers": len(batch.receivers) if vrx else 0,
            "audio_samples": audio_total,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": dominant,
            "kernel_ms": round(kernel_ms[dominant], 4),
            "algorithmic_bytes": int(algo[dominant]),
            "valu": valu.get(dominant),
        },
        "valu_roofline": valu,
        "stages": stages,
        "batched_receivers": None if not vrx else {
            "descriptors": batch.n_desc, "work_items": batch.n_work,
            "batch_ms": round(kernel_ms["vrx_batch"], 4),
            "per_receiver_launches_ms": round(unbatched_ms, 4),
            "launches_replaced": 2 * batch.n_desc,
            "speedup": round(unbatched_ms / kernel_ms["vrx_batch"], 2)},
        "mixed_precision": {
            "chain_m4096_ms": {"fp32_exact": round(kernel_ms["chain_m4096"], 4),
                               **{k: round(v, 4) for k, v in mfma_ms.items()}},
            "precision_vs_fp64": precision,
        },
        "engine_step": engine,
        "learned_detector": learned,
        "plain_decision": decision,
        "conditioning": conditioning,
        "host_stages": host,
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
    dev.bind_stream()
    return 0


def host_leg() -> dict:
    """The stages that run on the host as in the reference, each timed on one
    core: the adaptive noise profile over one receiver's 48 kHz audio (per
    second of audio), comb detection on a 16,384-bin averaged spectrum, and
    the comb bank's plan and observe per block for a 41-line comb."""
    from revenant_amd.comb import CombBank, detect_combs
    from revenant_amd.noise_profile import NoiseProfile, NoiseProfileSettings

    out = {}
    g = np.random.default_rng(0x484F5354)
    rate, seconds = 48_000, 10
    t = np.arange(rate * seconds) / rate
    audio = (0.05 * np.sin(2 * np.pi * 120 * t) + 0.1 * np.sin(2 * np.pi * 1500 * t)
             + g.normal(0.0, 0.01, t.size)).astype(np.float32)
    prof = NoiseProfile()
    cfg = NoiseProfileSettings(enabled=True, strength_db=12.0)
    t0 = time.perf_counter()
    for i in range(0, audio.size, 960):  # 20 ms blocks, quiet for the first second
        prof.process(audio[i:i + 960], rate, i < rate, cfg)
    ms = (time.perf_counter() - t0) * 1e3 / seconds
    out["noise_profile"] = {"ms_per_audio_second": round(ms, 3),
                            "realtime_factor": round(1e3 / ms, 1), "cores": 1,
                            "what": "one receiver, 48 kHz mono, 1024-point STFT pair per 512 samples"}

    crate, n = 1_152_000, 1 << 20
    tt = np.arange(n) / crate
    x = g.normal(0.0, 0.01, n) + 1j * g.normal(0.0, 0.01, n)
    for k in range(-20, 21):
        x = x + 0.004 * np.exp(2j * np.pi * ((1234.0 + 15_000.0 * k) * tt + g.uniform()))
    frames = x[: (n // 16384) * 16384].reshape(-1, 16384) * np.hanning(16384)
    power = np.fft.fftshift(np.mean(np.abs(np.fft.fft(frames, axis=1)) ** 2, axis=0))
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        found = detect_combs(power, crate)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["comb_detect"] = {"ms_per_spectrum": round(ms, 3), "bins": 16384, "combs": len(found),
                          "cores": 1}

    bank = CombBank(crate)
    bank.detected(found)
    block, blocks, lines = 65536, 200, 0
    t0 = time.perf_counter()
    for b in range(blocks):
        ln, nl, _ = bank.plan(b * block, block)
        lines = nl
        if nl:
            proj = (g.normal(size=(nl, 16)) + 1j * g.normal(size=(nl, 16))).astype(np.complex64)
            bank.observe(b * block, block, proj)
    ms = (time.perf_counter() - t0) * 1e3 / blocks
    out["comb_bank"] = {"ms_per_block": round(ms, 4), "lines": lines, "block_samples": block,
                        "realtime_factor_1152ksps": round(block / crate * 1e3 / ms, 1),
                        "cores": 1}
    return out


def conditioning_leg(dev, torch, packed, g: Grid, partial, kernel_ms, steps,
                     cpu: bool) -> dict:
    """The stages either side of the path, each timed with HIP events on its
    stream over `steps` launches: front-end moments (HBM-bound, 4 B read per
    sample), the chain with the correction fused in against the plain chain,
    the excision passes and the comb kernels on a 2^24-sample cf32 block, and
    256 carrier loops (a serial recurrence per receiver: latency-bound). With
    `cpu`, the oracle is timed beside each on a bounded sample of the same
    input on one host thread (the CPU path each stage replaces)."""
    import ctypes as C

    from revenant_amd import _lib as L
    from revenant_amd.dsp import build_nco_table, comb_step, comb_table
    from revenant_amd.excision import (ExcisionCursor, design_excision, excision_threshold,
                                       size_excision_rings)
    from revenant_amd.front_end import moments_chunks

    def timed(fn):
        fn()
        dev.timer_begin()
        for _ in range(steps):
            fn()
        return dev.timer_end() / steps

    out = {}
    # front end: moments over the whole resident input, and the corrected chain
    sums = torch.empty(moments_chunks(SAMPLES) * 5, dtype=torch.float32, device="cuda")
    ms = timed(lambda: dev.iq_moments_codes(2, packed, 0, SAMPLES, sums))
    moments_bytes = 4 * SAMPLES + 4 * sums.numel()
    corr = L.IqCorrectParams(0, 0, 0, 0.0213, -0.0097, -0.0524078, 0.9536746)
    before = L.IqCorrectParams(0, 0, 0, 0.0, 0.0, 0.0, 1.0)
    cdesc = _lib_copy(g.desc)
    cdesc.taps_scaled = 0
    cdesc.correction = C.cast(C.pointer(corr), C.c_void_p)
    cdesc.correction_before = C.cast(C.pointer(before), C.c_void_p)
    cdesc.correction_from = SAMPLES // 2
    chain_corr = timed(lambda: dev.chain(cdesc, packed, g.taps, g.tw, g.ring, partial))
    out["front_end"] = {
        "moments_ms": round(ms, 4),
        "moments_roofline": {"bound": "hbm", "achieved_gbs": round(moments_bytes / (ms * 1e-3) / 1e9, 1),
                             "frac": round(moments_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "bytes": moments_bytes},
        "moments_gsps": round(SAMPLES / (ms * 1e-3) / 1e9, 2),
        "chain_m4096_corrected_ms": round(chain_corr, 4),
        "chain_m4096_plain_ms": round(kernel_ms["chain_m4096"], 4),
    }
    if cpu:
        from oracle import oracle
        n_cpu = 1 << 22
        words = packed[:4 * n_cpu].cpu().numpy()
        ring_h = np.zeros(2 * n_cpu, np.float32)
        oracle.convert(2, words, L.ConvertParams(n_cpu - 1, 0, 0, n_cpu), ring_h)
        t0 = time.perf_counter()
        oracle.iq_moments(n_cpu - 1, 0, n_cpu, ring_h)
        t_m = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.iq_correct(n_cpu - 1, 0, n_cpu, 0.0213, -0.0097, -0.0524078, 0.9536746, ring_h)
        t_c = time.perf_counter() - t0
        out["front_end"]["cpu_oracle"] = {
            "moments_msps": round(n_cpu / t_m / 1e6, 1), "correct_msps": round(n_cpu / t_c / 1e6, 1),
            "cores": 1, "sample": f"{n_cpu} samples of the resident input (converted)"}

    # a cf32 block with noise and impulses for excision and comb
    n_block = 1 << 24
    ring_n = 1 << 26
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    ring = (torch.randn(2 * ring_n, generator=gen, device="cuda") * 0.01).contiguous()
    hits = torch.randint(0, ring_n, (4096,), generator=gen, device="cuda")
    ring.view(-1, 2)[hits] += 1.5
    cfg = design_excision(RATE)
    rings = size_excision_rings(cfg, n_block, 1, ring_n - 1)
    power = torch.zeros(rings.seg_mask + 1, dtype=torch.float32, device="cuda")
    level = torch.zeros_like(power)
    flags = torch.zeros(rings.flag_mask + 1, dtype=torch.int32, device="cuda")
    removed = torch.zeros(2 * (rings.removed_mask + 1), dtype=torch.float32, device="cuda")
    counters = torch.zeros(4, dtype=torch.int32, device="cuda")
    cursor = ExcisionCursor(cfg, ring_n - 1, rings, excision_threshold(15.0))
    p0 = cursor.plan(0, n_block)  # fills the side rings; the second block is steady state
    dev.excise_segment(cfg, p0.segment, ring, power)
    dev.excise_reference(cfg, p0.reference, power, level)
    dev.excise_flag(cfg, p0.flag, ring, level, flags)
    plan = cursor.plan(n_block, n_block)
    copy = ring.clone()  # the apply pass cleans in place: time it on a copy each launch
    seg_ms = timed(lambda: dev.excise_segment(cfg, plan.segment, ring, power))
    ref_ms = timed(lambda: dev.excise_reference(cfg, plan.reference, power, level))
    flag_ms = timed(lambda: dev.excise_flag(cfg, plan.flag, ring, level, flags))
    apply_ms = timed(lambda: dev.excise_apply(cfg, plan.apply, copy, flags, copy, removed,
                                              counters))
    cursor.close()
    total_ms = seg_ms + ref_ms + flag_ms + apply_ms
    # algorithmic bytes: segment reads 8 B/sample; flag reads 8 B and writes
    # 4 B; apply reads 8 B + 4 B of flags and writes 8 + 8 B (cleaned, removed)
    algo = 8 * n_block + 12 * n_block + 28 * n_block
    out["excision"] = {
        "block_samples": n_block,
        "segment_ms": round(seg_ms, 4), "reference_ms": round(ref_ms, 4),
        "flag_ms": round(flag_ms, 4), "apply_ms": round(apply_ms, 4),
        "msps": round(n_block / (total_ms * 1e-3) / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved_gbs": round(algo / (total_ms * 1e-3) / 1e9, 1),
                     "frac": round(algo / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "bytes": algo},
        "counters": [int(v) for v in counters.cpu().numpy()],
    }
    if cpu:
        from oracle import oracle
        n_cpu = 1 << 21
        c_ring = copy[:2 * (1 << 22)].cpu().numpy().copy()
        c_cfg = design_excision(RATE)
        c_rings = size_excision_rings(c_cfg, n_cpu, 1, (1 << 22) - 1)
        c_cur = ExcisionCursor(c_cfg, (1 << 22) - 1, c_rings, excision_threshold(15.0))
        c_pow = np.zeros(c_rings.seg_mask + 1, np.float32)
        c_lev = np.zeros_like(c_pow)
        c_flags = np.zeros(c_rings.flag_mask + 1, np.uint32)
        c_rem = np.zeros(2 * (c_rings.removed_mask + 1), np.float32)
        c_cnt = np.zeros(4, np.uint32)
        t_total = 0.0
        for k in range(2):
            pl = c_cur.plan(k * n_cpu, n_cpu)
            t0 = time.perf_counter()
            for which, fn, args_ in ((0, oracle.excise_segment, (pl.segment, c_ring, c_pow)),
                                     (1, oracle.excise_reference, (pl.reference, c_pow, c_lev)),
                                     (2, oracle.excise_flag, (pl.flag, c_ring, c_lev, c_flags))):
                if args_[0].count:
                    cc = L.ExcisionConfig.from_buffer_copy(bytes(c_cfg))
                    cc.pass_ = which
                    fn(cc, *args_)
            if pl.apply.count:
                cc = L.ExcisionConfig.from_buffer_copy(bytes(c_cfg))
                cc.pass_ = 3
                oracle.excise_apply(cc, pl.apply, c_ring, c_flags, c_ring, c_rem, c_cnt)
            if k == 1:
                t_total = time.perf_counter() - t0
        c_cur.close()
        out["excision"]["cpu_oracle"] = {"msps": round(n_cpu / t_total / 1e6, 1), "cores": 1,
                                         "sample": f"the second of two {n_cpu}-sample blocks"}

    # comb: 64 lines over the block, 16 sub-blocks
    table = torch.from_numpy(comb_table().view(np.float32)).cuda()
    lines = (L.CombLine * 64)()
    for k in range(64):
        lines[k].phase0 = (k * 2654435761) & 0xFFFFFFFF
        lines[k].step = comb_step(1234.5 + (k - 32) * 45_678.25, RATE)
        lines[k].amplitude[0], lines[k].amplitude[1] = 0.01, -0.005
    lines_d = torch.frombuffer(bytearray(bytes(lines)), dtype=torch.uint8).cuda()
    cp = L.CombParams(ring_n - 1, 0, n_block, 64, 16, rings.removed_mask, 0, 0)
    proj = torch.zeros(2 * 64 * 16, dtype=torch.float32, device="cuda")
    proj_ms = timed(lambda: dev.comb_project(cp, ring, lines_d, 64, table, proj))
    sub_ms = timed(lambda: dev.comb_subtract(cp, ring, lines_d, 64, table, copy, removed))
    sub_bytes = 8 * n_block * 3  # read x, write cleaned and removed
    out["comb"] = {"lines": 64, "block_samples": n_block, "project_ms": round(proj_ms, 4),
                   "subtract_ms": round(sub_ms, 4),
                   "subtract_msps": round(n_block / (sub_ms * 1e-3) / 1e6, 1),
                   "subtract_roofline": {
                       "bound": "hbm", "achieved_gbs": round(sub_bytes / (sub_ms * 1e-3) / 1e9, 1),
                       "frac": round(sub_bytes / (sub_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "bytes": sub_bytes}}
    if cpu:
        from oracle import oracle
        n_cpu = 1 << 18
        x = ring[:2 * n_cpu].cpu().numpy().copy()
        tab = comb_table()
        cpp = L.CombParams(n_cpu - 1, 0, n_cpu, 64, 16, n_cpu - 1, 0, 0)
        t0 = time.perf_counter()
        oracle.comb_subtract(cpp, x, lines, tab, n_cpu)
        t_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.comb_project(cpp, x, lines, tab)
        t_p = time.perf_counter() - t0
        out["comb"]["cpu_oracle"] = {"subtract_msps": round(n_cpu / t_s / 1e6, 1),
                                     "project_msps": round(n_cpu / t_p / 1e6, 1), "cores": 1,
                                     "sample": f"{n_cpu} samples, 64 lines"}
    del ring, copy

    # carrier loops: 256 receivers (SAM and DSB alternating), one second of
    # 48 kHz audio each, one lane per receiver
    n_rx, count, cap = 256, 48_000, 1 << 16
    rings_c = (torch.randn(n_rx, 2 * cap, generator=gen, device="cuda") * 0.3).contiguous()
    states = torch.zeros(n_rx, 8, dtype=torch.int32, device="cuda")
    nco = torch.from_numpy(build_nco_table(16).view(np.float32)).cuda()
    from revenant_amd.dsp import DEMOD_DSB, DEMOD_SAM, GridParams, VrxRequest, plan_vrx
    plans = [plan_vrx(GridParams(64, 17, 32), 2_400_000, VrxRequest(196_500, m, bw))
             for m, bw in ((DEMOD_SAM, 10_000), (DEMOD_DSB, 6_000))]
    items = []
    for k in range(n_rx):
        pl = plans[k % 2]
        it = L.VrxCarrierItem(pl.plan.carrier, pl.carrier_block(cap - 1, 0, count),
                              rings_c[k].data_ptr(), cap, states[k].data_ptr(), 0)
        L.check(L.lib().rvx_vrx_carrier_item_check(C.byref(it), 1 << 16))
        items.append(it)
    items_d = torch.frombuffer(bytearray(b"".join(bytes(i) for i in items)),
                               dtype=torch.uint8).cuda()
    car_ms = timed(lambda: dev.vrx_carrier_batch(items_d, n_rx, nco))
    out["carrier"] = {"receivers": n_rx, "samples_per_receiver": count,
                      "ms": round(car_ms, 4),
                      "ns_per_sample_serial": round(car_ms * 1e6 / count, 1),
                      "aggregate_msps": round(n_rx * count / (car_ms * 1e-3) / 1e6, 1),
                      "realtime_factor_48k": round(1000.0 / car_ms, 1),
                      "bound": "latency (one serial recurrence per receiver)"}
    if cpu:
        from oracle import oracle
        n_cpu_rx = 8
        host = rings_c[:n_cpu_rx].cpu().numpy().copy()
        st = np.zeros((n_cpu_rx, 8), np.uint32)
        nco_h = build_nco_table(16)
        t0 = time.perf_counter()
        for k in range(n_cpu_rx):
            pl = plans[k % 2]
            oracle.vrx_carrier(pl.plan.carrier, pl.carrier_block(cap - 1, 0, count), nco_h,
                               host[k], st[k])
        t_c = time.perf_counter() - t0
        out["carrier"]["cpu_oracle"] = {
            "ns_per_sample_serial": round(t_c * 1e9 / (n_cpu_rx * count), 1),
            "msps_per_core": round(n_cpu_rx * count / t_c / 1e6, 1), "cores": 1,
            "sample": f"{n_cpu_rx} receivers x {count} samples"}
    del rings_c
    out["noise"] = noise_leg(dev, torch, gen, timed, cpu)
    out["engine_conditioning"] = engine_conditioning_leg(dev, torch, packed)
    return out


def engine_conditioning_leg(dev, torch, packed, push: int = 1 << 24, pushes: int = 6) -> dict:
    """The engine's conditioned path against its plain one on the 4096-channel
    grid (averaged spectrum on, no receivers): pushes of 2^24 resident codes,
    plain (the chain reads the codes), with impulse excision (each push
    converted into the sample ring, the four passes, the cleaned samples into
    the cf32 window, the chain's cf32 form), and with 64 comb lines projected
    and subtracted ahead of the excision. ms per push, the first push untimed."""
    from revenant_amd import _lib as L
    from revenant_amd.dsp import FORMAT_CS16, GridParams, comb_step, design_prototype
    from revenant_amd.engine import Engine
    from revenant_amd.excision import design_excision, excision_threshold

    grid = GridParams(M, TAPS, D)
    proto = design_prototype(grid, beta=8.0).taps
    codes = packed.view(torch.int32)
    lines = (L.CombLine * 64)()
    for k in range(64):
        lines[k].phase0 = (0x9E3779B9 * (k + 1)) % (1 << 32)
        lines[k].step = comb_step(1_000_000.0 * (k + 1) - 31_000_000.0, RATE)
        lines[k].amplitude[0] = np.float32(1e-3)
    out = {"push_samples": push}
    for name in ("plain", "excision", "comb64_excision"):
        eng = Engine(dev, grid, proto, FORMAT_CS16, max_push=push, ring_blocks=1 << 14,
                     frame_blocks=FRAME, row_frames=ROW_FRAMES, alpha=ALPHA)
        if name == "comb64_excision":
            eng.set_comb(lines)
        if name != "plain":
            eng.set_excision(design_excision(RATE), excision_threshold(15.0))
        eng.push(codes[:push])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(1, pushes + 1):
            at = (k * push) % (SAMPLES - push)
            eng.push(codes[at:at + push])
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / pushes
        out[name] = {"ms_per_push": round(ms, 3), "MS_per_s": round(push / (ms * 1e-3) / 1e6, 1)}
        eng.close()
    return out


def noise_leg(dev, torch, gen, timed, cpu: bool) -> dict:
    """The per-receiver noise stages: the impulse blanker's two passes over a
    2^24-sample channel span (HBM-bound: detect reads 8 B and writes a 4 B
    flag per sample, apply reads 8 B and a flag and writes 8 B), and the line
    kernel (two notches and the line enhancer) and spectral noise reduction
    (frame 512) for 256 receivers, 10 ms of 48 kS/s audio each, one workgroup
    per receiver (serial recurrences: latency-bound)."""
    import ctypes as C

    from revenant_amd import _lib as L
    from revenant_amd.dsp import (LINE_ALE, blanker_threshold, design_blanker, design_notch,
                                  line_notch_flags, line_state_size, set_notch,
                                  spectral_params, spectral_state_size, spectral_tables)
    out = {}
    # blanker over one long span of a channel ring
    n, ring = 1 << 24, 1 << 25
    chan = (torch.randn(2 * ring, generator=gen, device="cuda") * 0.01).contiguous()
    hits = torch.randint(0, ring, (8192,), generator=gen, device="cuda")
    chan.view(-1, 2)[hits] += 0.5
    flags = torch.zeros(ring, dtype=torch.float32, device="cuda")
    blanked = torch.zeros(2 * ring, dtype=torch.float32, device="cuda")
    cfg = design_blanker(144_000)
    bp = L.BlankerParams(0, ring - 1, 1000, ring - 1, 1000, n, blanker_threshold(12.0))
    ap = L.BlankerParams(0, ring - 1, 1000 + cfg.hang, ring - 1, 1000 + cfg.hang,
                         n - cfg.hang - cfg.lead, bp.threshold)
    det_ms = timed(lambda: dev.blank_detect(cfg, bp, chan, flags))
    app_ms = timed(lambda: dev.blank_apply(cfg, ap, chan, flags, blanked))
    det_bytes, app_bytes = 12 * n, 20 * n
    out["blanker"] = {
        "samples": n, "window": cfg.window,
        "detect_ms": round(det_ms, 4), "apply_ms": round(app_ms, 4),
        "msps": round(n / ((det_ms + app_ms) * 1e-3) / 1e6, 1),
        "detect_roofline": {"bound": "hbm",
                            "achieved_gbs": round(det_bytes / (det_ms * 1e-3) / 1e9, 1),
                            "frac": round(det_bytes / (det_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "bytes": det_bytes},
        "apply_roofline": {"bound": "hbm",
                           "achieved_gbs": round(app_bytes / (app_ms * 1e-3) / 1e9, 1),
                           "frac": round(app_bytes / (app_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "bytes": app_bytes},
        "flagged": int((flags[1000:1000 + n] != 0).sum().item())}
    if cpu:
        from oracle import oracle
        n_cpu = 1 << 21
        h_chan = chan[:2 * (1 << 22)].cpu().numpy().reshape(-1, 2).copy()
        h_flags = np.zeros(1 << 22, np.float32)
        h_out = np.zeros((1 << 22, 2), np.float32)
        c_bp = L.BlankerParams(0, (1 << 22) - 1, 1000, (1 << 22) - 1, 1000, n_cpu, bp.threshold)
        t0 = time.perf_counter()
        oracle.blank_detect(cfg, c_bp, h_chan, h_flags)
        oracle.blank_apply(cfg, c_bp, h_chan, h_flags, h_out)
        t_b = time.perf_counter() - t0
        out["blanker"]["cpu_oracle"] = {"msps": round(n_cpu / t_b / 1e6, 1), "cores": 1,
                                        "sample": f"{n_cpu} samples, both passes"}
    del chan, flags, blanked

    # line kernel and spectral stage: 256 receivers x 480 samples
    n_rx, count = 256, 480
    audio = (torch.randn(n_rx, count, generator=gen, device="cuda") * 0.1).contiguous()
    lc = L.LineConfig()
    lp = L.LineParams(flags=line_notch_flags(2) | LINE_ALE, stride=7, delay=240, mu=1.0e-4,
                      leak=float(np.float32(1.0 - 3.0e-6)), eps=1.0e-10, count=count)
    set_notch(lp, 0, design_notch(1000.0, 40.0, 100.0, 48_000))
    set_notch(lp, 1, design_notch(2200.0, 30.0, 80.0, 48_000))
    l_state = torch.zeros(n_rx, line_state_size(lc), dtype=torch.float32, device="cuda")
    items = []
    for k in range(n_rx):
        it = L.NoiseLineItem(lp, audio[k].data_ptr(), l_state[k].data_ptr())
        L.check(L.lib().rvx_noise_line_item_check(C.byref(lc), C.byref(it), count,
                                                  l_state.shape[1]))
        items.append(it)
    l_items = torch.frombuffer(bytearray(b"".join(bytes(i) for i in items)),
                               dtype=torch.uint8).cuda()
    line_ms = timed(lambda: dev.noise_line_batch(lc, l_items, n_rx))
    frame = 512
    tables = torch.from_numpy(spectral_tables(frame)).cuda()
    sp = spectral_params(0.5)
    sp.count = count
    s_state = torch.zeros(n_rx, spectral_state_size(frame), dtype=torch.float32, device="cuda")
    items = []
    for k in range(n_rx):
        it = L.NoiseSpectralItem(sp, (C.c_uint32 * 2)(), audio[k].data_ptr(),
                                 s_state[k].data_ptr())
        L.check(L.lib().rvx_noise_spectral_item_check(frame, C.byref(it), count,
                                                      s_state.shape[1]))
        items.append(it)
    s_items = torch.frombuffer(bytearray(b"".join(bytes(i) for i in items)),
                               dtype=torch.uint8).cuda()
    spec_ms = timed(lambda: dev.noise_spectral_batch(frame, tables, s_items, n_rx))
    block_ms = 1000.0 * count / 48_000
    out["line"] = {"receivers": n_rx, "samples_per_receiver": count, "ms": round(line_ms, 4),
                   "ns_per_sample_serial": round(line_ms * 1e6 / count, 1),
                   "realtime_factor_48k": round(block_ms / line_ms, 1),
                   "bound": "latency (one serial recurrence per receiver)"}
    out["spectral"] = {"receivers": n_rx, "samples_per_receiver": count, "frame": frame,
                       "ms": round(spec_ms, 4), "realtime_factor_48k": round(block_ms / spec_ms, 1),
                       "bound": "latency (direct DFT per hop, one workgroup per receiver)"}
    if cpu:
        from oracle import oracle
        n_cpu_rx = 16
        host = audio[:n_cpu_rx].cpu().numpy().copy()
        st = np.zeros((n_cpu_rx, line_state_size(lc)), np.float32)
        t0 = time.perf_counter()
        for k in range(n_cpu_rx):
            oracle.noise_line(lc, lp, host[k], st[k])
        t_l = time.perf_counter() - t0
        host = audio[:n_cpu_rx].cpu().numpy().copy()
        st = np.zeros((n_cpu_rx, spectral_state_size(frame)), np.float32)
        tab = spectral_tables(frame)
        t0 = time.perf_counter()
        for k in range(n_cpu_rx):
            oracle.noise_spectral(frame, sp, tab, host[k], st[k])
        t_s = time.perf_counter() - t0
        sample = f"{n_cpu_rx} receivers x {count} samples"
        out["line"]["cpu_oracle"] = {"ms_for_256": round(t_l / n_cpu_rx * n_rx * 1e3, 2),
                                     "cores": 1, "sample": sample}
        out["spectral"]["cpu_oracle"] = {"ms_for_256": round(t_s / n_cpu_rx * n_rx * 1e3, 2),
                                         "cores": 1, "sample": sample}
    return out


def spot_check(torch, packed, g: Grid, partial, frames, alpha, batch, starts, counts, audio):
    """Parity of this very run against the oracle: two frames of the channel
    ring and averaged spectrum, and the first audio of one receiver of each
    kind recomputed from the GPU's own channel rings."""
    try:
        from oracle import oracle
        from revenant_amd.dsp import FORMAT_CS16, ema_frame_decay
    except Exception as exc:  # the oracle is optional on the box
        return f"skipped ({type(exc).__name__}: {exc})"
    # The oracle sees the input from sample start*D on, i.e. block indices
    # shifted by `start`; start is a multiple of the frame (and of M/D), so the
    # phase correction and the frame alignment are unchanged by the shift.
    check = 2 * FRAME
    start = g.blocks // 2 - (g.blocks // 2) % FRAME
    hi = (g.first_block + start + check - 1) * D + 1
    sub = packed[start * D * 4:hi * 4].cpu().numpy()
    cap = 1 << (check - 1).bit_length()
    od = oracle.OChain(FORMAT_CS16, M, TAPS, D, g.first_block, check, cap, FRAME, alpha,
                       ema_frame_decay(alpha, FRAME))
    ring_cpu, part_cpu = oracle.chain(od, sub, g.proto.taps, g.tw_np, threads=4)
    ring_cpu = ring_cpu.reshape(M, cap, 2)
    g_ring = g.ring.view(M, g.pitch, 2)
    ok = True
    for b in range(check):
        m = g.first_block + start + b
        col = ring_cpu[:, (g.first_block + b) & (cap - 1)]
        ok &= np.array_equal(g_ring[:, m & (g.ring_blocks - 1)].cpu().numpy().view(np.uint32),
                             col.view(np.uint32))
    gp = partial.view(frames, M)[start // FRAME:start // FRAME + 2].cpu().numpy()
    ok &= np.array_equal(gp.view(np.uint32), part_cpu.reshape(2, M).view(np.uint32))
    what = f"chain bit-exact vs oracle on {check} blocks" if ok else "chain MISMATCH vs oracle"
    if batch is None:
        return what
    from revenant_amd.dsp import build_nco_table
    nco = build_nco_table(16)
    seen, rx_ok, n_rx = set(), True, 0
    for v, rx in enumerate(batch.receivers):
        kind = (rx.plan.demod.mode, rx.plan.demod.decimation)
        if kind in seen:
            continue
        seen.add(kind)
        n = min(600, counts[v])
        p = rx.plan
        chan = rx.ring.view(-1, 2)[rx.channel_base:rx.channel_base + rx.channel_mask + 1]
        chan = chan.cpu().numpy()
        r, hist = p.demod.decimation, p.history
        fine_count = (n - 1) * r + hist + 1
        out_cap = 1 << (fine_count - 1).bit_length()
        fp, _ = p.fine_block(0, rx.channel_mask, out_cap - 1, starts[v] * r - hist, fine_count)
        fine = oracle.vrx_fine(p.fine, fp, chan.ravel(), p.fine_taps, nco, out_cap)
        dp = p.demod_block(out_cap - 1, starts[v], n)
        ref = oracle.vrx_demod(p.demod, dp, fine, p.weights)
        got = audio[v][:n * p.demod.channels].cpu().numpy()
        rx_ok &= np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        n_rx += 1
    return what + (f"; audio bit-exact vs oracle for {n_rx} receiver kinds" if rx_ok else
                   "; audio MISMATCH vs oracle")


def engine_leg(dev, torch, packed, plans, wide_set, tiles, starts, audio, steps: int,
               warmup: int, graph_ms: float) -> dict:
    """The step through the streaming engine: one push of the 2^27 resident
    codes per step into an engine with the 4096-channel grid (averaged
    spectrum on) and the 512-channel grid added beside it for the 200 kHz
    receivers, all 256 receivers placed on it (the tuned tiles). The first
    push covers exactly the blocks and audio the graph step covers, so its
    audio is compared with the graph step's bit for bit; the pushes after it
    continue the stream (the same codes again) and are timed, the host's
    planning and launches included."""
    from revenant_amd import _lib
    from revenant_amd.dsp import FORMAT_CS16, GridParams, design_prototype
    from revenant_amd.engine import Engine

    main = GridParams(M, TAPS, D)
    wide = GridParams(WM, TAPS, WD)
    eng = Engine(dev, main, design_prototype(main, beta=8.0).taps, FORMAT_CS16,
                 max_push=SAMPLES, ring_blocks=1 << 17, frame_blocks=FRAME,
                 row_frames=ROW_FRAMES, alpha=ALPHA, ring_pitch=(1 << 17) + _lib.RING_PITCH_PAD)
    gid = eng.add_grid(wide, design_prototype(wide, beta=8.0).taps, ring_blocks=1 << 20,
                       max_rows=len(wide_set), ring_pitch=(1 << 20) + _lib.RING_PITCH_PAD)
    ids = [eng.add(p, audio_capacity=1 << 17, tile=(tiles[i] if tiles else 0),
                   grid=gid if name == "wfm" else 0) for i, (name, p) in enumerate(plans)]
    codes = packed.view(torch.int32)
    eng.push(codes)
    same = True
    for rid, a0, want in zip(ids, starts, audio):
        first, got = eng.read_audio(rid)
        ref = want.cpu().numpy()
        same &= first == a0 and np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    for _ in range(warmup):
        eng.push(codes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.push(codes)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = eng.stats()
    eng.close()
    return {"ms_per_step": round(ms, 4), "MS_per_s": round(SAMPLES / (ms * 1e-3) / 1e6, 1),
            "vs_graph_step": round(graph_ms / ms, 3), "pushes": 1 + warmup + steps,
            "blocks": int(st.blocks), "grid_blocks": int(st.grid_blocks),
            "first_push_audio_bit_exact_vs_graph_step": bool(same)}


def plain_decision(dev, torch, cpu: bool, bins: int = 65536, reps: int = 40) -> dict:
    """One plain-detector decision (noise floor, ladder, greedy growth) over
    an averaged spectrum of the reference's shipped size, 65,536 bins: the
    device decision (rvx_detect, median of `reps` calls, each waiting for its
    result) beside the CPU twin (the oracle, one thread) on the same input,
    and the bands of both compared."""
    from revenant_amd.detect import Detector, config

    g = np.random.default_rng(SEED)
    avg = g.gamma(64.0, 1.0 / 64.0, bins) * 1e-6  # a 64-look average near 1e-6
    for _ in range(bins // 4000):
        w = int(g.integers(1, bins // 30))
        s0 = int(g.integers(0, bins - w))
        avg[s0:s0 + w] += avg[s0:s0 + w].mean() * float(g.uniform(0.05, 30.0))
    bin_hz = RATE / M / 16
    det = Detector(dev, config(bins, bin_hz))
    x = torch.from_numpy(avg).cuda()
    for _ in range(5):
        det.decide(x)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        bands, peaks = det.decide(x)
        times.append((time.perf_counter() - t0) * 1e3)
    det.close()
    out = {"bins": bins, "bin_hz": bin_hz, "peaks": peaks, "bands": len(bands),
           "ms_gpu_decision": round(float(np.median(times)), 4),
           "ms_reference_published": 7.4}
    if cpu:
        try:
            from oracle import oracle
        except Exception as exc:
            out["cpu_twin"] = f"oracle unavailable: {exc}"
            return out
        cfg = oracle.detect_config(bins, bin_hz)
        ct = []
        for _ in range(5):
            t0 = time.perf_counter()
            ob, _, _, _ = oracle.detect(cfg, avg)
            ct.append((time.perf_counter() - t0) * 1e3)
        out["ms_cpu_twin"] = round(float(np.median(ct)), 3)
        out["cpu_twin_cores"] = 1
        out["bands_match_cpu_twin"] = [b[:2] for b in bands] == [b[:2] for b in ob]
        out["speedup_vs_cpu_twin"] = round(out["ms_cpu_twin"] / out["ms_gpu_decision"], 1)
    return out


def plain_twin_per_block(dev, proposals) -> dict:
    """The plain detector's CPU twin (the oracle, as the reference runs it on
    the host) deciding every row of one 256-row block: ms per block."""
    try:
        from oracle import oracle
    except Exception as exc:
        return {"plain_cpu_twin": f"oracle unavailable: {exc}"}
    scene = proposals.build(dev, [99_999])[0]
    rows = scene.noisy.double().cpu().numpy()
    cfg = oracle.detect_config(proposals.M, proposals.BIN_HZ)
    t0 = time.perf_counter()
    for r in range(rows.shape[0]):
        oracle.detect(cfg, rows[r])
    ms = (time.perf_counter() - t0) * 1e3
    return {"ms_plain_cpu_twin_per_block": round(ms, 2), "plain_cpu_twin_cores": 1}


def _lib_copy(desc):
    return type(desc).from_buffer_copy(desc)


# VALU lane-operations per second: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz. The
# bit-exact kernels run without contraction and without packed FP32, so each
# floating-point operation is one lane-operation.
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9


def chain_lane_ops(channels: int, decimation: int, taps: int) -> int:
    """The fewest lane-operations one block of the bit-exact chain can take:
    the branch sums (a multiply and an add per real part: 4 per tap and
    branch), each new input sample converted once (2), the radix-2 transform
    (5 N log2 N), the output correction (a complex multiply, 6 per channel),
    and the power and its average (3 + 3 per channel)."""
    n = channels
    return (4 * taps * n + 2 * decimation + 5 * n * (n.bit_length() - 1) + 6 * n + 6 * n)


def valu_roof(kernel_ms: dict, main: "Grid", wide: "Grid | None") -> dict:
    """Per chain kernel: its lane-operations per launch, the time they take at
    the VALU peak, and the fraction of that floor the measured time reaches."""
    out = {}
    grids = [("chain_m4096", main)] + ([("chain_m512_wfm", wide)] if wide is not None else [])
    for name, g in grids:
        p = g.params
        ops = chain_lane_ops(p.channels, p.decimation, p.taps_per_branch) * g.blocks
        if g.frames == 0:
            ops -= 6 * p.channels * g.blocks  # no power average on this grid
        floor = ops / VALU_PEAK_LANE_OPS * 1e3
        out[name] = {"lane_ops": int(ops), "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                     "floor_ms": round(floor, 4),
                     "frac": round(floor / kernel_ms[name], 4)}
    return out


def precision_in_run(dev, torch):
    """Error against float64 of each form of the 4096-channel chain, measured
    in this run after the timed region: the 100 dB dynamic-range multitone of
    revenant_amd.precision through the bit-exact fp32 chain and the
    matrix-core forms, each checked against the oracle's float64 channelizer
    of the same codes."""
    try:
        from oracle import oracle
    except Exception as exc:  # the oracle is optional on the box
        return {"skipped": f"{type(exc).__name__}: {exc}"}
    from revenant_amd import _lib as L
    from revenant_amd import precision as pr
    from revenant_amd.dsp import GridParams, design_prototype

    codes, first, _, _, _ = pr.multitone()
    taps = design_prototype(GridParams(pr.M, pr.L, pr.D), beta=8.0).taps
    ref = oracle.fp64_chain(codes, taps, pr.M, pr.D, first, pr.BLOCKS)
    return pr.run_forms(dev, torch, ref, (("plain_fp32", None), ("bf16x3", L.PREC_BF16X3)))


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated against a copy kernel), as committed under profiles/."""
    p = ROOT / "profiles" / f"{kernel}_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(packed, main: Grid, wide: Grid | None, batch, seconds):
    """The oracle over a bounded sample of the same workload on this host's
    cores: the 4096-channel chain (convert -> branch -> transform ->
    correction -> frame average), and with receivers the 512-channel chain and
    every receiver's fine stage + demodulator over the sample's channel
    samples, repeated to about `seconds`."""
    try:
        from oracle import oracle
    except Exception as exc:
        return {"value": None, "unit": "MS/s", "cores": 0, "kind": "port",
                "sample": f"oracle unavailable: {exc}"}
    from revenant_amd.dsp import FORMAT_CS16, build_nco_table, ema_frame_decay

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    alpha = np.float32(ALPHA)
    decay = ema_frame_decay(alpha, FRAME)
    blocks = 2048 - 2048 % (FRAME * threads)
    hi = (main.first_block + blocks) * D + 1
    sub = packed[: hi * 4].cpu().numpy()
    cap = 1 << (blocks - 1).bit_length()
    od = oracle.OChain(FORMAT_CS16, M, TAPS, D, main.first_block, blocks, cap, FRAME, alpha,
                       decay)
    ring = np.zeros(2 * M * cap, dtype=np.float32)
    partial = np.zeros((blocks // FRAME) * M, dtype=np.float32)
    wide_od = wide_ring = None
    if wide is not None:
        wblocks = (hi - 1) // WD - wide.first_block + 1
        wblocks -= wblocks % threads
        wcap = 1 << (wblocks - 1).bit_length()
        wide_od = oracle.OChain(FORMAT_CS16, WM, TAPS, WD, wide.first_block, wblocks, wcap, 0,
                                alpha, 0.0)
        wide_ring = np.zeros(2 * WM * wcap, dtype=np.float32)
    nco = build_nco_table(16)

    def receivers(nthreads):
        # every receiver over the channel samples of this sample
        jobs = []
        for rx in batch.receivers:
            p = rx.plan
            if rx.ring is main.ring:
                first, last, rows, rcap, src = main.first_block, main.first_block + blocks - 1, \
                    M, cap, ring
            else:
                first, last, rows, rcap, src = wide.first_block, \
                    wide.first_block + wide_od.block_count - 1, WM, wide_od.out_ring_blocks, \
                    wide_ring
            a0, n = rx.audio_range(first, last)
            if n == 0:
                continue
            k = p.plan.placement.channel
            jobs.append((p, src.reshape(rows, rcap, 2)[k], rcap, a0, n))

        def one(job):
            p, chan, rcap, a0, n = job
            r, hist = p.demod.decimation, p.history
            fine_count = (n - 1) * r + hist + 1
            out_cap = 1 << (fine_count - 1).bit_length()
            fp, _ = p.fine_block(0, rcap - 1, out_cap - 1, a0 * r - hist, fine_count)
            fine = oracle.vrx_fine(p.fine, fp, np.ascontiguousarray(chan).ravel(),
                                   p.fine_taps, nco, out_cap)
            dp = p.demod_block(out_cap - 1, a0, n)
            return oracle.vrx_demod(p.demod, dp, fine, p.weights).size

        with ThreadPoolExecutor(nthreads) as pool:
            return sum(pool.map(one, jobs))

    def timed_passes(nthreads, budget):
        passes, total = 0, 0.0
        while total < budget or passes == 0:
            t0 = time.perf_counter()
            oracle.chain(od, sub, main.proto.taps, main.tw_np, threads=nthreads, ring=ring,
                         partial=partial)
            if wide is not None:
                oracle.chain(wide_od, sub, wide.proto.taps, wide.tw_np, threads=nthreads,
                             ring=wide_ring)
                receivers(nthreads)
            total += time.perf_counter() - t0
            passes += 1
        return passes, total

    passes, total = timed_passes(threads, seconds)
    rate = passes * blocks * D / total / 1e6
    one_passes, one_total = timed_passes(1, seconds / 2)
    one_rate = one_passes * blocks * D / one_total / 1e6
    what = "4096-ch chain" + (", 512-ch chain and 256 receivers" if wide is not None else "")
    return {"value": round(rate, 2), "unit": "MS/s", "cores": threads, "kind": "port",
            "single_thread_value": round(one_rate, 3),
            "cpu_model": cpu_model(),
            "host_threads": len(os.sched_getaffinity(0)),
            "sample": f"{passes} passes over {blocks * D} input samples ({blocks} blocks) of "
                      f"the same workload ({what}), {total:.1f} s on {threads} threads; "
                      f"{one_passes} passes, {one_total:.1f} s on one thread"}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())

"""A closed loop around the comb bank for the tests: a drifting comb of
interference lines over white noise (tests/engine/test_engine_comb.cpp:1-120:
41 lines, k in [-20, 20], the spacing starting at 15 kHz and drifting
0.5 Hz/s, the offset at 1234 Hz drifting 0.2 Hz/s, lines about 20 dB over the
floor in a bin), the project and subtract passes restated in numpy (what the
kernels compute, comb_kernels_reference.cpp:100-180), and the line height a
spectrum shows (a line's bin over the frame's median, dB)."""
import numpy as np

RATE = 1_152_000
K_LO, K_HI = -20, 20
OFFSET_HZ, SPACING_HZ = 1_234.0, 15_000.0
OFFSET_DRIFT, SPACING_DRIFT = 0.2, 0.5      # Hz/s
NOISE_RMS, LINE_AMPLITUDE = 0.01, 0.004
SEED = 0x434F4D4243414E43


def offset_at(t):
    return OFFSET_HZ + OFFSET_DRIFT * t


def spacing_at(t):
    return SPACING_HZ + SPACING_DRIFT * t


def cycles_at(k, t):
    """Line k's phase at t in cycles: the integral of its frequency."""
    return OFFSET_HZ * t + 0.5 * OFFSET_DRIFT * t * t + k * (SPACING_HZ * t + 0.5 * SPACING_DRIFT * t * t)


def render(seconds: float, lines: bool = True) -> np.ndarray:
    g = np.random.default_rng(SEED)
    n = int(seconds * RATE)
    x = (g.normal(0.0, NOISE_RMS, n) + 1j * g.normal(0.0, NOISE_RMS, n)).astype(np.complex128)
    if lines:
        t = np.arange(n) / RATE
        for k in range(K_LO, K_HI + 1):
            cyc = cycles_at(k, t) + g.uniform()
            x += LINE_AMPLITUDE * np.exp(2j * np.pi * (cyc - np.floor(cyc)))
    return x.astype(np.complex64)


def tone_spur(seconds: float, freq_hz: float, seed: int = SEED + 1) -> np.ndarray:
    g = np.random.default_rng(seed)
    n = int(seconds * RATE)
    x = g.normal(0.0, NOISE_RMS, n) + 1j * g.normal(0.0, NOISE_RMS, n)
    cyc = freq_hz * np.arange(n) / RATE
    x += LINE_AMPLITUDE * np.exp(2j * np.pi * (cyc - np.floor(cyc)))
    return x.astype(np.complex64)


def spectrum(x: np.ndarray, nfft: int) -> np.ndarray:
    """Averaged Hann periodogram, natural order (DC at nfft / 2)."""
    frames = x[: len(x) // nfft * nfft].reshape(-1, nfft) * np.hanning(nfft).astype(np.float32)
    p = np.mean(np.abs(np.fft.fft(frames, axis=1)) ** 2, axis=0)
    return np.fft.fftshift(p).astype(np.float32)


def tones(lines, count: int, table: np.ndarray) -> np.ndarray:
    """[lines, count] complex64: table[(phase0 + n step) >> 20] per line."""
    n = np.arange(count, dtype=np.uint64)
    out = np.empty((len(lines), count), dtype=np.complex64)
    for i, ln in enumerate(lines):
        ph = (np.uint64(ln.phase0) + n * np.uint64(ln.step)) & np.uint64(0xFFFFFFFF)
        out[i] = table[(ph >> np.uint64(20)).astype(np.int64)]
    return out


def project(block: np.ndarray, tone: np.ndarray, subs: int) -> np.ndarray:
    """[lines, subs] sums of x * conj(tone) over equal sub-blocks."""
    prod = block[None, :] * np.conj(tone)
    return prod.reshape(tone.shape[0], subs, -1).sum(axis=2).astype(np.complex64)


def run(x: np.ndarray, bank, table, block: int = 65536, detect_every: float = 1.0,
        detect=None, subs: int = 16):
    """The loop: every `detect_every` seconds of input, detect on the last
    second's spectrum (when `detect` is given); per block: plan, project the
    raw block, subtract if asked, observe. Returns the cleaned signal."""
    out = x.copy()
    next_detect = int(detect_every * RATE)
    for start in range(0, len(x) - block + 1, block):
        if detect is not None and start >= next_detect:
            detect(x[max(0, start - RATE):start])
            next_detect += int(detect_every * RATE)
        lines, n, subtract = bank.plan(start, block)
        if n == 0:
            continue
        seg = x[start:start + block]
        tone = tones(lines[:n], block, table)
        bank.observe(start, block, project(seg, tone, subs))
        if subtract:
            amp = np.array([complex(ln.amplitude[0], ln.amplitude[1]) for ln in lines[:n]],
                           dtype=np.complex64)
            out[start:start + block] = seg - amp @ tone
    return out


def line_height_db(x: np.ndarray, t0: float, nfft: int = 4096, t1: float | None = None,
                   ks=None) -> float:
    """Mean over the comb's lines (those of `ks`, all when None) of its bin
    over the median bin, dB, over the signal from t0 to t1 (the end when None;
    the lines placed where they are at mid-span)."""
    seg = x[int(t0 * RATE):None if t1 is None else int(t1 * RATE)]
    p = spectrum(seg, nfft)
    med = np.median(p)
    mid = t0 + 0.5 * len(seg) / RATE
    heights = []
    for k in (range(K_LO, K_HI + 1) if ks is None else ks):
        f = offset_at(mid) + k * spacing_at(mid)
        b = int(round(f / (RATE / nfft))) + nfft // 2
        if 0 <= b < nfft:
            heights.append(10 * np.log10(p[b] / med))
    return float(np.mean(heights))


def render_moved(seconds: float, jump_at: float, jump_hz: float) -> np.ndarray:
    """The drifting comb over noise, every line shifted by jump_hz from
    jump_at on (a comb that moves while a pass holds it, as a switching
    supply's comb does when its load changes)."""
    g = np.random.default_rng(SEED + 2)
    n = int(seconds * RATE)
    x = (g.normal(0.0, NOISE_RMS, n) + 1j * g.normal(0.0, NOISE_RMS, n)).astype(np.complex128)
    t = np.arange(n) / RATE
    shift = np.where(t >= jump_at, jump_hz * (t - jump_at), 0.0)
    for k in range(K_LO, K_HI + 1):
        cyc = cycles_at(k, t) + shift + g.uniform()
        x += LINE_AMPLITUDE * np.exp(2j * np.pi * (cyc - np.floor(cyc)))
    return x.astype(np.complex64)


def render_rippled(seconds: float, ripple_hz: float, depth: float, seed: int = SEED + 3) -> np.ndarray:
    """The drifting comb over noise with every line amplitude-modulated by a
    ripple (1 + depth cos(2 pi ripple t + theta_k)): sidebands depth / 2 of
    the carrier at +- ripple_hz, as a rectifier's ripple puts them."""
    g = np.random.default_rng(seed)
    n = int(seconds * RATE)
    x = (g.normal(0.0, NOISE_RMS, n) + 1j * g.normal(0.0, NOISE_RMS, n)).astype(np.complex128)
    t = np.arange(n) / RATE
    for k in range(K_LO, K_HI + 1):
        cyc = cycles_at(k, t) + g.uniform()
        env = 1.0 + depth * np.cos(2 * np.pi * ripple_hz * t + 2 * np.pi * g.uniform())
        x += LINE_AMPLITUDE * env * np.exp(2j * np.pi * (cyc - np.floor(cyc)))
    return x.astype(np.complex64)


def render_curved(seconds: float, curvature_hz: float, seed: int = SEED + 4) -> np.ndarray:
    """The drifting comb over noise with line k moved by k^2 curvature_hz: a
    comb whose spacing changes across it (c + k d + k^2 e)."""
    g = np.random.default_rng(seed)
    n = int(seconds * RATE)
    x = (g.normal(0.0, NOISE_RMS, n) + 1j * g.normal(0.0, NOISE_RMS, n)).astype(np.complex128)
    t = np.arange(n) / RATE
    for k in range(K_LO, K_HI + 1):
        cyc = cycles_at(k, t) + k * k * curvature_hz * t + g.uniform()
        x += LINE_AMPLITUDE * np.exp(2j * np.pi * (cyc - np.floor(cyc)))
    return x.astype(np.complex64)

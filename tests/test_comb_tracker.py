"""The comb bank in a closed loop on the CPU (rvx_comb_bank_*; the tracker of
core/dsp/comb_tracker.cpp and the bank of comb_bank.cpp): the drifting comb of
tests/engine/test_engine_comb.cpp detected, locked and taken out of the
spectrum, the kernels' projections restated in numpy. The same loop through
the engine on the GPU is tests/test_gpu_comb_cancel.py."""
import numpy as np
import pytest

import comb_loop as cl
from revenant_amd import _lib
from revenant_amd.comb import CombBank, detect_combs, spur
from revenant_amd.dsp import comb_table

SECONDS = 4.0
MEASURE_FROM = 3.0


@pytest.fixture(scope="module")
def capture():
    return cl.render(SECONDS)


def test_a_drifting_comb_is_detected_locked_and_taken_out(capture):
    table = comb_table()
    bank = CombBank(cl.RATE)
    found_any = []

    def detect(last_second):
        found = detect_combs(cl.spectrum(last_second, 16384), cl.RATE, bank.active())
        found_any.append(len(found))
        bank.detected(found)

    cleaned = cl.run(capture, bank, table, detect=detect)
    st = bank.status()
    # up to the end of the last whole block the loop ran
    end = (len(capture) // 65536) * 65536 / cl.RATE
    off = cl.line_height_db(capture, MEASURE_FROM, t1=end)
    on = cl.line_height_db(cleaned, MEASURE_FROM, t1=end)
    print(f"comb off {off:.1f} dB, on {on:.1f} dB; lock {st.lock} lines {st.lines_cancelled} "
          f"suppression {st.suppression_db:.1f} dB spacing {st.comb.spacing_hz:.3f} "
          f"offset {st.comb.offset_hz:.3f} drift {st.spacing_drift_hz_per_s:.3f} Hz/s "
          f"residual {st.residual_rad:.3f} rad; detections {found_any}")
    assert off > 10.0
    assert found_any and found_any[0] > 0
    assert st.lock == _lib.COMB_LOCKED
    assert st.lines_cancelled >= _lib.COMB_MIN_LINES
    spacing = st.comb.spacing_hz
    assert abs(spacing - cl.spacing_at(SECONDS)) < 5.0 or \
        abs(spacing - 2 * cl.spacing_at(SECONDS)) < 10.0
    assert on < off - 10.0 and on < 6.0
    assert st.suppression_db > 6.0


def test_a_named_spur_is_locked_and_cancelled():
    f = 23_456.7
    x = cl.tone_spur(3.0, f)
    table = comb_table()
    bank = CombBank(cl.RATE)
    bank.spurs([spur(f + 20.0)])   # named from a coarse detection, 20 Hz off
    cleaned = cl.run(x, bank, table)
    st = bank.status()
    assert st.spurs_cancelled == 1
    nfft = 4096
    b = int(round(f / (cl.RATE / nfft))) + nfft // 2
    end = len(x) // 65536 * 65536  # the last whole block the loop ran
    raw = cl.spectrum(x[2 * cl.RATE:end], nfft)
    out = cl.spectrum(cleaned[2 * cl.RATE:end], nfft)
    before = 10 * np.log10(raw[b] / np.median(raw))
    after = 10 * np.log10(out[b] / np.median(out))
    print(f"spur: {before:.1f} dB over the floor before, {after:.1f} dB after")
    assert before > 10.0 and after < 3.0
    bank.drop_spurs()
    assert bank.status().spurs_cancelled == 0


def test_noise_alone_starts_nothing_and_a_restart_forgets():
    x = cl.render(1.5, lines=False)
    table = comb_table()
    bank = CombBank(cl.RATE)
    found = detect_combs(cl.spectrum(x[: cl.RATE], 16384), cl.RATE)
    assert found == []
    bank.detected(found)
    lines, n, subtract = bank.plan(0, 65536)
    assert n == 0 and not subtract
    assert bank.status().lock == _lib.COMB_IDLE
    # a comb that is not there: acquisition never locks, and reset clears it
    e = _lib.CombEstimate()
    e.offset_hz, e.spacing_hz, e.k_lo, e.k_hi = 1000.0, 15_000.0, -10, 10
    bank.detected([e])
    assert bank.status().lock == _lib.COMB_ACQUIRING
    cleaned = cl.run(x, bank, table)
    assert bank.status().lock != _lib.COMB_LOCKED
    assert np.array_equal(cleaned, x)  # nothing subtracted while not locked
    bank.reset()
    assert bank.status().lock == _lib.COMB_IDLE


def test_observations_split_back_to_their_trackers():
    """Two combs in two trackers: one plan holds both trackers' lines in turn,
    and the observation of that block is split back by the counts."""
    bank = CombBank(cl.RATE)
    a, b = _lib.CombEstimate(), _lib.CombEstimate()
    a.offset_hz, a.spacing_hz, a.k_lo, a.k_hi = 1_000.0, 15_000.0, -10, 10
    b.offset_hz, b.spacing_hz, b.k_lo, b.k_hi = 3_333.0, 22_930.0, -5, 5
    bank.detected([a, b])
    assert bank.status().combs_active == 2
    lines, n, _ = bank.plan(0, 65536)
    assert n == 21 + 11
    steps = [ln.step for ln in lines[:n]]
    want = [_lib.lib().rvx_comb_step(a.offset_hz + k * a.spacing_hz, cl.RATE)
            for k in range(-10, 11)] + \
           [_lib.lib().rvx_comb_step(b.offset_hz + k * b.spacing_hz, cl.RATE)
            for k in range(-5, 6)]
    assert [abs(int(np.int32(np.uint32(s - w)))) <= 1 for s, w in zip(steps, want)] == [True] * n
    # a second detection of a held comb is dropped
    bank.detected([a])
    assert bank.status().combs_active == 2
    # projections of the wrong size are refused
    with pytest.raises(_lib.RvxError):
        bank.observe(0, 65536, np.zeros((n - 1, 16), np.complex64))


def test_a_withdrawn_spur_goes_and_is_held_off():
    """comb_cancel.h:614-624: a spur the detector withdraws goes at once, and a
    spur that went is not started again near its offset for three seconds of
    planned blocks (RVX_COMB_SPUR_HOLD_OFF_S); after that it may."""
    f = 23_456.7
    block = 65536
    x = cl.tone_spur(3.0, f)
    table = comb_table()
    bank = CombBank(cl.RATE)
    bank.spurs([spur(f)])
    cl.run(x, bank, table)
    assert bank.status().spurs_cancelled == 1
    last = (len(x) - block) // block * block  # the last block the loop planned
    bank.withdraw_spurs([f - 500.0])          # not this one
    assert bank.status().spurs_cancelled == 1
    bank.withdraw_spurs([f + 30.0])           # within 45 Hz of where it was detected
    assert bank.status().spurs_cancelled == 0
    t = last + block
    assert bank.plan(t, block)[1] == 0
    bank.spurs([spur(f + 10.0)])              # offered again at once: held off
    assert bank.plan(t + block, block)[1] == 0
    bank.spurs([spur(f + 500.0)])             # another carrier starts
    assert bank.plan(t + 2 * block, block)[1] == 1
    hold = int(_lib.COMB_SPUR_HOLD_OFF_S * cl.RATE)
    bank.plan(last + hold + block, block)     # the hold-off has run out
    bank.spurs([spur(f)])
    assert bank.plan(last + hold + 2 * block, block)[1] == 2


def test_spurs_dropped_in_bands_and_dropped_all_are_held_off():
    """drop_spurs_in (comb_cancel.h:626-634) lets go of the spurs detected in a
    band, edges included, and keeps the rest; drop_spurs lets go of all; each
    one dropped is held off like a withdrawn one."""
    block = 65536
    bank = CombBank(cl.RATE)
    bank.spurs([spur(10_000.0), spur(50_000.0)])
    assert bank.plan(0, block)[1] == 2
    bank.drop_spurs_in([(45_000.0, 55_000.0), (-80_000.0, -60_000.0)])
    assert bank.plan(block, block)[1] == 1
    bank.spurs([spur(50_010.0)])               # held off
    assert bank.plan(2 * block, block)[1] == 1
    bank.drop_spurs_in([(10_000.0, 10_000.0)])  # the band's edge holds it
    assert bank.plan(3 * block, block)[1] == 0
    bank.spurs([spur(80_000.0)])
    assert bank.plan(4 * block, block)[1] == 1
    bank.drop_spurs()
    assert bank.plan(5 * block, block)[1] == 0
    bank.spurs([spur(80_000.0), spur(10_000.0), spur(50_000.0)])  # all three held off
    assert bank.plan(6 * block, block)[1] == 0
    bank.reset()                               # a reset forgets the hold-offs too
    bank.spurs([spur(80_000.0)])
    assert bank.plan(7 * block, block)[1] == 1


def test_a_starting_comb_takes_its_lines_back_from_the_newest_spurs():
    """comb_bank.cpp detected(): combs have first claim on the 1024 line slots,
    so a comb needing 1022 lines starts beside four held spurs, and the two
    newest spurs give their lines back (and are held off); a comb of fewer
    than eight lines never starts, and the second of two combs whose needs do
    not both fit takes an even share."""
    block = 65536
    bank = CombBank(cl.RATE)
    bank.spurs([spur(f) for f in (100_000.0, 110_000.0, 120_000.0, 130_000.0)])
    assert bank.plan(0, block)[1] == 4
    e = _lib.CombEstimate()
    e.offset_hz, e.spacing_hz, e.k_lo, e.k_hi = 500.0, 1_000.0, -511, 510  # 1022 lines
    bank.detected([e])
    lines, n, _ = bank.plan(block, block)
    assert n == 1024
    # the comb's lines sit at 500 + 1000 k Hz, never on a spur's 10 kHz multiple
    freqs = [round(lines[i].step / 2**32 * cl.RATE) for i in range(n)]
    kept = sorted(f for f in freqs if f in (100_000, 110_000, 120_000, 130_000))
    assert kept == [100_000, 110_000]  # the two oldest spurs stayed
    bank.spurs([spur(130_000.0)])        # went a moment ago: held off
    assert bank.plan(2 * block, block)[1] == 1024

    small = CombBank(cl.RATE)
    e7 = _lib.CombEstimate()
    e7.offset_hz, e7.spacing_hz, e7.k_lo, e7.k_hi = 500.0, 1_000.0, 0, 6  # seven lines
    small.detected([e7])
    assert small.plan(0, block)[1] == 0

    two = CombBank(cl.RATE)
    a, b = _lib.CombEstimate(), _lib.CombEstimate()
    a.offset_hz, a.spacing_hz, a.k_lo, a.k_hi = 500.0, 1_000.0, -400, 399     # 800 lines
    b.offset_hz, b.spacing_hz, b.k_lo, b.k_hi = 700.0, 3_333.0, -100, 99      # 200 lines
    two.detected([a, b])
    assert two.plan(0, block)[1] == 1000                                      # both fit
    three = CombBank(cl.RATE)
    c = _lib.CombEstimate()
    c.offset_hz, c.spacing_hz, c.k_lo, c.k_hi = 900.0, 7_777.0, -300, 299     # 600 lines
    three.detected([a, c])
    # 800 + 600 do not fit: the first takes max(even 512, room 424) = 512, the
    # second what is left, 512
    assert three.plan(0, block)[1] == 1024


def test_a_pass_started_on_part_of_a_comb_grows_to_cover_it(capture):
    """comb_bank.cpp / comb_tracker.cpp growth: a pass started on the central
    21 of the comb's 41 lines, handed later detections of the whole comb
    (held, so they widen the range its pass may grow into), grows once locked
    by up to 16 lines a block and takes the outer lines out as well."""
    from revenant_amd.comb import detect_comb
    import ctypes as C
    table = comb_table()
    bank = CombBank(cl.RATE)
    whole = detect_comb(cl.spectrum(capture[: cl.RATE], 16384), cl.RATE)
    assert whole is not None and whole.k_hi - whole.k_lo + 1 >= 35
    part = _lib.CombEstimate()
    C.memmove(C.byref(part), C.byref(whole), C.sizeof(part))
    mid = (whole.k_lo + whole.k_hi) // 2
    part.k_lo, part.k_hi = mid - 10, mid + 10
    bank.detected([part])
    assert bank.plan(0, 65536)[1] == 21

    def detect(last_second):
        bank.detected([whole])

    cleaned = cl.run(capture, bank, table, detect=detect)
    st = bank.status()
    assert st.lock == _lib.COMB_LOCKED
    assert st.lines_cancelled >= 35, st.lines_cancelled
    end = (len(capture) // 65536) * 65536 / cl.RATE
    outer = [k for k in range(cl.K_LO, cl.K_HI + 1) if abs(k) > 10]
    off = cl.line_height_db(capture, MEASURE_FROM, t1=end, ks=outer)
    on = cl.line_height_db(cleaned, MEASURE_FROM, t1=end, ks=outer)
    print(f"outer lines: {off:.1f} dB over the floor before, {on:.1f} dB after; "
          f"{st.lines_cancelled} lines cancelled")
    assert off > 10.0 and on < off - 10.0 and on < 6.0


def test_a_recovering_pass_takes_its_moved_comb():
    """comb_bank.cpp detected() / comb_tracker.cpp recentre: the comb locked by
    two passes (its halves) jumps 2 kHz at 2.3 s; the passes lose it and
    recover; the next detection names the comb at its new place on the same
    spacing, overlapping none of the passes' lines, and the recovering passes
    take it (no pass starts beside them), relock and take its lines out."""
    x = cl.render_moved(6.0, 2.3, 2000.0)
    table = comb_table()
    bank = CombBank(cl.RATE)
    seen = []

    def detect(last_second):
        found = detect_combs(cl.spectrum(last_second, 16384), cl.RATE, bank.active())
        before = bank.status().combs_active
        bank.detected(found)
        seen.append((len(found), before, bank.status().combs_active,
                     [round(e.offset_hz) for e in found]))

    cleaned = cl.run(x, bank, table, detect=detect)
    print(seen)
    # the detection after the jump: the moved comb found, no pass started
    n_found, before, after, offsets = seen[2]
    assert n_found >= 1 and any(abs(o - 3234) < 100 or abs(o + 11766) < 100 for o in offsets)
    assert before == after == 2
    st = bank.status()
    assert st.lock == _lib.COMB_LOCKED
    end = (len(x) // 65536) * 65536 / cl.RATE
    seg = lambda s: s[int(5.0 * cl.RATE):int(end * cl.RATE)]
    p_raw, p_out = cl.spectrum(seg(x), 4096), cl.spectrum(seg(cleaned), 4096)
    heights = []
    for name, p in (("raw", p_raw), ("cleaned", p_out)):
        med = np.median(p)
        mid = 0.5 * (5.0 + end)
        h = [10 * np.log10(p[int(round((cl.offset_at(mid) + 2000.0 + k * cl.spacing_at(mid))
                                        / (cl.RATE / 4096))) + 2048] / med)
             for k in range(cl.K_LO, cl.K_HI + 1)]
        heights.append(float(np.mean(h)))
    print(f"moved comb over the floor, 5 s on: {heights[0]:.1f} dB raw, {heights[1]:.1f} dB cleaned")
    assert heights[0] > 10.0 and heights[1] < 6.0


def test_ripple_sidebands_are_tracked_and_taken_out():
    """comb_tracker.cpp sidebands: every line of the drifting comb amplitude-
    modulated at 500 Hz (depth 0.6: sidebands 10.5 dB under their carriers);
    the detection reports the ripple and one order, each pass holds its
    carriers with their sidebands (the other lines' leakage taken out of each
    line's block sum in closed form), steers the ripple by the sidebands' phase
    errors past the carriers' fit, and cancels carriers and sidebands alike."""
    ripple = 500.0
    x = cl.render_rippled(4.0, ripple, 0.6)
    table = comb_table()
    bank = CombBank(cl.RATE)
    found_first = []

    def detect(last_second):
        found = detect_combs(cl.spectrum(last_second, 16384), cl.RATE, bank.active())
        found_first.append(found)
        bank.detected(found)

    cleaned = cl.run(x, bank, table, detect=detect)
    first = found_first[0]
    assert first and all(abs(e.ripple_hz - ripple) < 5.0 and e.sideband_orders == 1 for e in first)
    st = bank.status()
    assert st.lock == _lib.COMB_LOCKED
    assert st.lines_cancelled >= 3 * 35, st.lines_cancelled
    assert abs(st.comb.ripple_hz - ripple) < 1.0 and st.comb.sideband_orders == 1
    end = (len(x) // 65536) * 65536 / cl.RATE

    def height(sig, sides):
        seg = sig[int(MEASURE_FROM * cl.RATE):int(end * cl.RATE)]
        p = cl.spectrum(seg, 16384)
        med, mid = np.median(p), 0.5 * (MEASURE_FROM + end)
        h = [10 * np.log10(p[int(round((cl.offset_at(mid) + k * cl.spacing_at(mid) + sgn * ripple)
                                        / (cl.RATE / 16384))) + 8192] / med)
             for k in range(cl.K_LO, cl.K_HI + 1) for sgn in sides]
        return float(np.mean(h))

    carriers = (height(x, (0,)), height(cleaned, (0,)))
    sidebands = (height(x, (-1, 1)), height(cleaned, (-1, 1)))
    print(f"carriers {carriers[0]:.1f} -> {carriers[1]:.1f} dB, sidebands {sidebands[0]:.1f} -> "
          f"{sidebands[1]:.1f} dB over the floor; {st.lines_cancelled} lines cancelled, ripple "
          f"{st.comb.ripple_hz:.3f} Hz")
    assert carriers[0] > 20.0 and carriers[1] < 6.0
    assert sidebands[0] > 12.0 and sidebands[1] < 6.0


def test_a_curved_comb_is_tracked_with_its_curvature():
    """comb_tracker.cpp choose_model: line k of the comb at c + k d + k^2 e with
    e = 0.05 Hz (20 Hz off a straight grid at the outer lines, 0.7 cycles a
    block: a straight pass never locks on it, residual 1.7 rad). The pass votes
    the curved model in on the carriers' frequency errors, tracks e, locks and
    takes every line out."""
    from revenant_amd.comb import detect_comb
    e = 0.05
    x = cl.render_curved(4.0, e)
    table = comb_table()
    bank = CombBank(cl.RATE)

    def detect(last_second):
        whole = detect_comb(cl.spectrum(last_second, 16384), cl.RATE)
        bank.detected([whole] if whole is not None else [])

    cleaned = cl.run(x, bank, table, detect=detect)
    st = bank.status()
    assert st.lock == _lib.COMB_LOCKED
    assert abs(st.comb.curvature_hz - e) < 0.002, st.comb.curvature_hz
    assert st.lines_cancelled >= 35
    end = (len(x) // 65536) * 65536 / cl.RATE

    def height(sig):
        seg = sig[int(MEASURE_FROM * cl.RATE):int(end * cl.RATE)]
        p = cl.spectrum(seg, 4096)
        med, mid = np.median(p), 0.5 * (MEASURE_FROM + end)
        return float(np.mean([10 * np.log10(p[int(round(
            (cl.offset_at(mid) + k * cl.spacing_at(mid) + k * k * e) / (cl.RATE / 4096))) + 2048]
            / med) for k in range(cl.K_LO, cl.K_HI + 1)]))

    before, after = height(x), height(cleaned)
    print(f"curved comb: {before:.1f} -> {after:.1f} dB over the floor, curvature "
          f"{st.comb.curvature_hz:.4f} Hz, {st.lines_cancelled} lines")
    assert before > 20.0 and after < 6.0

"""The spur rule at the end of every decision (Detector::mark_spurs,
core/detect/detector.cpp:687-780), the emitter memory whose key-ups it reads
(refine.h:261-295, detector.cpp:1885-1930) and the spur book carried across a
retune (SpurBook, :648-685), on the cases of the reference's
tests/detect/test_detector.cpp:2679-2920: a constant carrier probed as one is
a spur after five continuous seconds; keyed (two key-ups on its emitter),
fading, double-sideband and unprobed carriers are not; a carrier at the same
baseband offset after a retune is a spur at once, one at the same absolute
frequency or with its old frequency still occupied is not; a probe renaming a
spur withdraws it. Also the rule's edges (the 50-frame and 1.5 dB level test,
the 500 Hz / 3-bin width, the DC guard, held spells, a spur kept while
cancelled, the book's 256-entry eviction and 300 s memory, the emitter's 1 kHz
match) and the named spurs handed to the comb bank (graph.cpp:1615-1655). The
probe's family is recorded by hand, as the reference's tests record a probe,
so the rule is what is under test. Host only."""
import numpy as np
import pytest

from revenant_amd import _lib, events

BINS = 4096
RATE = 4_096_000
BW = RATE / BINS                      # 1 kHz bins, as the reference's scene
ZERO = -BINS / 2 * BW
CENTER = 100_000_000
SPUR_BIN = 700
DECISION = RATE // 10                 # a decision every 0.1 s of source time


def _candidate(first, last, center_hz, snr_db=20.0):
    c = _lib.Candidate()
    c.first_bin, c.last_bin = first, last
    c.center_hz, c.bandwidth_hz = int(center_hz), int((last - first + 1) * BW)
    c.snr_2500_db = snr_db
    return c


def _at(bin_, center=CENTER):
    return center + int(ZERO + bin_ * BW)


class Scene:
    """One tracker fed a decision every 0.1 s and one level frame per decision
    (the carriers' bins at `level_db`, the rest at -90 dB)."""

    def __init__(self, center=CENTER, seed=1):
        cfg = events.event_config(BW, ZERO, RATE, center, residual_decisions=0,
                                  centre_smoothing=1.0)
        self.t = events.Tracker(cfg, BINS)
        self.center = center
        self.k = 0
        self.rng = np.random.default_rng(seed)

    def step(self, bins, level_db=-70.0, jitter_db=0.3, width=1):
        cands = [_candidate(b, b + width - 1, _at(b + (width - 1) / 2, self.center))
                 for b in bins]
        frame = np.full(BINS, -90.0, np.float32)
        for b in bins:
            lv = level_db(self.k) if callable(level_db) else level_db
            frame[b:b + width] = lv + self.rng.normal(0.0, jitter_db, width)
        self.t.update(np.full(BINS, 2.0), np.ones(BINS), cands, 1_000 + self.k * DECISION, 0.1)
        self.t.measure_levels(frame)
        self.k += 1
        return self.t.tracks()

    def run(self, seconds, bins, **kw):
        for _ in range(int(round(seconds * 10))):
            tracks = self.step(bins, **kw)
        return tracks

    def track_at(self, bin_):
        want = _at(bin_, self.center)
        near = [tr for tr in self.t.tracks() if abs(tr.center_hz - want) < 3000]
        return near[0] if near else None


def _probed_scene(family=_lib.FAMILY_UNMODULATED, dsb=False, seconds=7.5, **kw):
    s = Scene()
    s.run(0.5, [SPUR_BIN], **kw)               # born (three hits) and live
    tr = s.track_at(SPUR_BIN)
    assert tr is not None and tr.state == 1
    s.t.record_probe(tr.id, family, dsb)
    s.run(seconds - 0.5, [SPUR_BIN], **kw)
    return s, s.track_at(SPUR_BIN)


def test_a_constant_carrier_is_a_spur_after_five_continuous_seconds():
    s = Scene()
    s.run(0.5, [SPUR_BIN])
    tr = s.track_at(SPUR_BIN)
    s.t.record_probe(tr.id, _lib.FAMILY_UNMODULATED)
    named_at = None
    for _ in range(70):
        s.step([SPUR_BIN])
        tr = s.track_at(SPUR_BIN)
        if tr.spur and named_at is None:
            named_at = s.k - 1                 # the decision that named it
    assert tr.spur and not tr.lo_spur
    # first seen at the first decision (k = 0, sample 1000), so five seconds
    # have passed from k = 50; but levels are measured only while live (from
    # the third hit, k = 2, one frame after each decision), so the decision
    # at k has k - 2 frames and the 50-frame test binds at k = 52
    assert tr.first_seen == 1_000
    assert named_at == 52
    book = s.t.spur_book()
    assert len(book) == 1
    assert book[0].baseband_hz == _at(SPUR_BIN) - CENTER and book[0].center_hz == CENTER


@pytest.mark.parametrize("keyups", [1, 2])
def test_keyups_are_counted_on_the_emitter_and_two_rule_a_carrier_out(keyups):
    """The emitter memory (refine.h:261-295, detector.cpp:1885-1930): the
    first probe remembers the track's emitter; a carrier that goes away until
    its track is dropped and comes back is a new track on the same emitter,
    and its coming live after more than a second of silence is a key-up. At
    most one key-up still lets a probed carrier be a spur; two do not."""
    s = Scene()
    s.run(0.5, [SPUR_BIN])
    tr = s.track_at(SPUR_BIN)
    assert tr.emitter == 0                      # no probe yet, no emitter
    s.t.record_probe(tr.id, _lib.FAMILY_UNMODULATED)
    emitter = s.track_at(SPUR_BIN).emitter
    assert emitter != 0
    for n in range(1, keyups + 1):
        s.run(6.0, [])                          # off: held, then dropped
        assert s.track_at(SPUR_BIN) is None
        s.run(0.5, [SPUR_BIN])                  # back: a new track joins the emitter
        tr = s.track_at(SPUR_BIN)
        assert tr.state == 1 and tr.emitter == emitter and tr.keyups == n
        assert tr.held_spells == 0
    s.t.record_probe(tr.id, _lib.FAMILY_UNMODULATED)
    s.run(7.0, [SPUR_BIN])
    tr = s.track_at(SPUR_BIN)
    assert tr.keyups == keyups
    assert tr.spur == (keyups <= 1)


def test_a_track_joins_an_emitter_within_its_tolerance():
    """EmitterBook::match: within max(1 kHz, half the narrower width)."""
    for offset_bins, joins in ((1, True), (2, False)):
        s = Scene()
        s.run(0.5, [SPUR_BIN])
        s.t.record_probe(s.track_at(SPUR_BIN).id, _lib.FAMILY_UNMODULATED)
        emitter = s.track_at(SPUR_BIN).emitter
        s.run(6.0, [])
        s.run(0.5, [SPUR_BIN + offset_bins])    # 1 kHz wide, 1 or 2 kHz away
        tr = s.track_at(SPUR_BIN + offset_bins)
        assert (tr.emitter == emitter) == joins
        assert tr.keyups == (1 if joins else 0)


@pytest.mark.parametrize("case", ["fading", "dsb", "other", "unprobed", "held"])
def test_what_is_not_a_spur(case):
    if case == "fading":                     # 6 dB over a 4 s period: sigma > 1.5 dB
        _, tr = _probed_scene(level_db=lambda k: -70.0 + 6.0 * np.sin(2 * np.pi * k / 40.0))
    elif case == "dsb":                        # an AM carrier, as steady as a spur
        _, tr = _probed_scene(dsb=True)
    elif case == "other":
        _, tr = _probed_scene(family=_lib.FAMILY_OTHER)
    elif case == "unprobed":
        s = Scene()
        s.run(7.5, [SPUR_BIN])
        tr = s.track_at(SPUR_BIN)
        assert s.t.spur_book() == []           # never passed the rule: not remembered
    else:                                      # a held spell breaks continuity
        s = Scene()
        s.run(0.5, [SPUR_BIN])
        tr = s.track_at(SPUR_BIN)
        s.t.record_probe(tr.id, _lib.FAMILY_UNMODULATED)
        s.run(1.0, [SPUR_BIN])
        s.step([])                             # one missed decision: held
        assert s.track_at(SPUR_BIN).state == 2
        s.run(7.0, [SPUR_BIN])
        tr = s.track_at(SPUR_BIN)
        assert tr.held_spells == 1
    assert tr is not None and tr.state == 1
    assert not tr.spur and not tr.lo_spur


def test_the_steady_level_needs_fifty_frames_under_one_and_a_half_db():
    # sigma just under and just over the limit: jitter of 1.2 dB passes, 2.5 fails
    _, tr = _probed_scene(jitter_db=1.2)
    sigma = np.sqrt(tr.level_m2 / (tr.level_frames - 1))
    assert sigma <= 1.5 and tr.spur
    _, tr = _probed_scene(jitter_db=2.5)
    sigma = np.sqrt(tr.level_m2 / (tr.level_frames - 1))
    assert sigma > 1.5 and not tr.spur


def test_width_and_dc_guard():
    # 3 bins of 1 kHz: 3 kHz <= max(500, 3 * 1000): a line; 4 bins is not
    _, tr = _probed_scene(width=3)
    assert tr.spur
    _, tr = _probed_scene(width=4)
    assert not tr.spur
    # a carrier within 500 Hz of the source centre is the front end's business
    s = Scene()
    dc_bin = BINS // 2                          # baseband 0 Hz
    s.run(0.5, [dc_bin])
    tr = s.track_at(dc_bin)
    s.t.record_probe(tr.id, _lib.FAMILY_UNMODULATED)
    s.run(7.0, [dc_bin])
    assert not s.track_at(dc_bin).spur


def test_a_renamed_spur_is_withdrawn_and_a_cancelled_one_is_kept():
    s, tr = _probed_scene()
    assert tr.spur
    where = tr.center_hz
    # while cancelled, its residue keys and fades; it stays named
    s.run(1.0, [SPUR_BIN], level_db=lambda k: -70.0 + 12.0 * (k % 2), jitter_db=0.0)
    tr = s.track_at(SPUR_BIN)
    assert tr.spur and s.t.withdrawn_spurs() == []
    # a probe naming another family takes it back, reported once
    s.t.record_probe(tr.id, _lib.FAMILY_OTHER)
    s.step([SPUR_BIN])
    assert s.t.withdrawn_spurs() == [where]
    assert not s.track_at(SPUR_BIN).spur
    s.step([SPUR_BIN])
    assert s.t.withdrawn_spurs() == []


RETUNE = 50_000
SHIFT = RETUNE // 1000                          # bins


def _book():
    s, tr = _probed_scene()
    assert tr.spur
    return s.t.spur_book()


def test_a_carrier_at_the_same_baseband_offset_after_a_retune_is_a_spur():
    book = _book()
    s = Scene(center=CENTER + RETUNE, seed=2)
    s.t.adopt_spur_book(book)
    s.run(5.5, [SPUR_BIN])                      # unprobed, but steady after 50 frames
    tr = s.track_at(SPUR_BIN)
    assert tr.lo_spur and tr.spur


def test_a_carrier_at_the_same_absolute_frequency_after_a_retune_is_a_station():
    book = _book()
    s = Scene(center=CENTER + RETUNE, seed=2)
    s.t.adopt_spur_book(book)
    s.run(5.5, [SPUR_BIN - SHIFT])
    tr = s.track_at(SPUR_BIN - SHIFT)
    assert tr.center_hz == _at(SPUR_BIN)        # where the old spur stood
    assert not tr.lo_spur and not tr.spur


def test_the_old_frequency_still_occupied_proves_nothing():
    book = _book()
    s = Scene(center=CENTER + RETUNE, seed=2)
    s.t.adopt_spur_book(book)
    s.run(5.5, [SPUR_BIN - SHIFT, SPUR_BIN])
    tr = s.track_at(SPUR_BIN)
    assert not tr.lo_spur and not tr.spur


def test_book_capacity_eviction_and_memory():
    full = [_lib.SpurSighting(baseband_hz=10_000 + 100 * i, center_hz=CENTER - 10_000_000,
                              last_seen=1_000 + i) for i in range(_lib.SPUR_BOOK_CAPACITY)]
    s = Scene()
    s.t.adopt_spur_book(full)
    s.run(0.5, [SPUR_BIN])
    tr = s.track_at(SPUR_BIN)
    s.t.record_probe(tr.id, _lib.FAMILY_UNMODULATED)
    s.run(5.5, [SPUR_BIN])
    book = s.t.spur_book()
    assert len(book) == _lib.SPUR_BOOK_CAPACITY
    # the longest unseen (last_seen 1000) went; the new sighting is in
    assert min(b.last_seen for b in book) == 1_001
    assert any(b.center_hz == CENTER and b.baseband_hz == _at(SPUR_BIN) - CENTER for b in book)
    # sightings unseen for more than 300 s of source time are forgotten
    t2 = events.Tracker(events.event_config(BW, ZERO, RATE, CENTER), BINS)
    t2.adopt_spur_book([_lib.SpurSighting(baseband_hz=5_000, center_hz=CENTER, last_seen=0)])
    t2.update(np.full(BINS, 2.0), np.ones(BINS), [], 300 * RATE, 0.1)
    assert len(t2.spur_book()) == 1             # exactly 300 s: kept
    t2.update(np.full(BINS, 2.0), np.ones(BINS), [], 300 * RATE + 1, 0.1)
    assert t2.spur_book() == []
    with pytest.raises(RuntimeError):
        t2.adopt_spur_book(full + full[:1])
    with pytest.raises(RuntimeError):
        t2.record_probe(12345, _lib.FAMILY_UNMODULATED)
    with pytest.raises(RuntimeError):
        t2.record_probe(12345, 7)


def test_named_spurs_reach_the_comb_bank_and_withdrawals_take_them_out():
    """The handoff (server.cpp:5661-5681, graph.cpp:1615-1655): a spur the
    rule names becomes a one-line spur tracker in the bank at its baseband
    offset; one inside a listened band does not; a withdrawal removes it."""
    from revenant_amd.comb import CombBank, take_spurs
    block = 65_536
    s, tr = _probed_scene()
    assert tr.spur
    offset = tr.center_hz - CENTER
    bank = CombBank(RATE)
    take_spurs(bank, s.t.tracks(), s.t.withdrawn_spurs(), CENTER,
               listened=[(offset - 1000.0, offset + 1000.0)])
    assert bank.plan(0, block)[1] == 0           # a receiver listens there
    take_spurs(bank, s.t.tracks(), s.t.withdrawn_spurs(), CENTER)
    assert bank.plan(block, block)[1] == 1       # one spur tracker, one line
    s.t.record_probe(tr.id, _lib.FAMILY_OTHER)
    s.step([SPUR_BIN])
    take_spurs(bank, s.t.tracks(), s.t.withdrawn_spurs(), CENTER)
    assert bank.plan(2 * block, block)[1] == 0


def test_a_cancelled_carrier_coming_back_is_not_a_keyup():
    """detector.cpp:1921: while the carrier at a centre is being cancelled,
    its going quiet is the cancellation's doing, so its return on the same
    emitter counts no key-up (the same return after no spur was named counts
    one, test_keyups_are_counted_on_the_emitter_and_two_rule_a_carrier_out)."""
    s, tr = _probed_scene()
    assert tr.spur
    emitter = tr.emitter
    s.run(6.0, [])                              # quiet (cancelled): dropped
    assert s.track_at(SPUR_BIN) is None
    s.run(0.5, [SPUR_BIN])
    tr = s.track_at(SPUR_BIN)
    assert tr.state == 1 and tr.emitter == emitter and tr.keyups == 0

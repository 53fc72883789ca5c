"""The delivery contract's host half (core/engine/spectrum_scale.cpp,
engine.h SpectrumGeometry), after tests/engine/test_spectrum_scale.cpp:81-560:
the first frame jumps; both ends track; a sudden signal is followed within a
few frames and a vanished one decays over thirty seconds (28 / 63 / 95 % of the
step at 10 / 30 / 90 s); one anomalous frame is rejected wherever it lands and
two consecutive are followed; a non-finite percentile is dropped; a uniform
frame opens to the minimum span about its level; a pinned end does not move;
refused configurations; and the spectrum geometry's exact rationals."""
import math

import numpy as np
import pytest

from revenant_amd.delivery import (SpectrumScale, SpectrumScaleConfig, spectrum_geometry)
from revenant_amd.dsp import GridParams

FRAME_S = 65536 / 2_400_000  # a frame every 27.3 ms


def settle(scale, low, high, seconds):
    ends = None
    for _ in range(int(seconds / FRAME_S)):
        ends = scale.update(low, high, FRAME_S)
    return ends


def test_the_first_frame_jumps():
    assert SpectrumScale().update(-92.0, -40.0, 0.0) == (-92.0, -40.0)


def test_both_ends_track_the_signal():
    s = SpectrumScale()
    s.update(-90.0, -30.0, 0.0)
    floor, ceiling = settle(s, -70.0, -10.0, 120.0)
    assert abs(floor + 70.0) < 0.5 and abs(ceiling + 10.0) < 0.5


def test_a_sudden_signal_is_followed_within_frames_and_decays_over_thirty_seconds():
    s = SpectrumScale()
    s.update(-95.0, -78.0, 0.0)
    settle(s, -95.0, -78.0, 60.0)
    ceilings = [s.update(-95.0, -38.0, FRAME_S)[1] for _ in range(4)]
    assert ceilings[0] < -77.0          # the first frame of a step is held back
    assert ceilings[2] > -38.0 - 0.12 * 40.0  # 88.7 % of the way after three
    s = SpectrumScale()
    s.update(-90.0, -20.0, 0.0)
    gone = []
    for secs in (10.0, 20.0, 60.0):
        gone.append((-20.0 - settle(s, -90.0, -60.0, secs)[1]) / 40.0)
    assert abs(gone[0] - 0.283) < 0.03 and abs(gone[1] - 0.632) < 0.03
    assert abs(gone[2] - 0.950) < 0.02


@pytest.mark.parametrize("at", [0, 1, 2])
def test_one_anomalous_frame_does_not_move_the_map(at):
    s = SpectrumScale()
    s.update(-95.0, -78.0, 0.0)
    settle(s, -95.0, -78.0, 10.0)
    for k in range(6):
        high = -10.0 if k == at else -78.0
        _, ceiling = s.update(-95.0, high, FRAME_S)
        assert ceiling < -77.5
    # the floor the same way, downward
    s2 = SpectrumScale()
    s2.update(-95.0, -60.0, 0.0)
    settle(s2, -95.0, -60.0, 10.0)
    for k in range(6):
        floor, _ = s2.update(-150.0 if k == at else -95.0, -60.0, FRAME_S)
        assert floor > -95.5


def test_two_consecutive_anomalous_frames_are_a_signal():
    s = SpectrumScale()
    s.update(-95.0, -78.0, 0.0)
    settle(s, -95.0, -78.0, 10.0)
    s.update(-95.0, -20.0, FRAME_S)
    _, ceiling = s.update(-95.0, -20.0, FRAME_S)
    assert ceiling > -50.0


def test_a_non_finite_percentile_is_dropped():
    s = SpectrumScale()
    s.update(-90.0, -30.0, 0.0)
    assert s.update(float("nan"), float("inf"), FRAME_S) == (-90.0, -30.0)


def test_a_uniform_frame_opens_about_its_level():
    floor, ceiling = settle(SpectrumScale(), -64.5, -64.5, 120.0)
    assert ceiling - floor >= 12.0 and floor < -64.5 < ceiling
    assert abs((-64.5 - floor) / (ceiling - floor) - 0.5) < 0.01


def test_a_pinned_end_does_not_move():
    s = SpectrumScale(SpectrumScaleConfig(pinned_floor_db=-110.0))
    s.update(-90.0, -30.0, 0.0)
    floor, ceiling = settle(s, -70.0, -10.0, 60.0)
    assert floor == -110.0 and abs(ceiling + 10.0) < 0.5
    s = SpectrumScale(SpectrumScaleConfig(pinned_ceiling_db=-50.0))
    floor, ceiling = settle(s, -55.0, -52.0, 60.0)
    assert ceiling == -50.0 and floor == -62.0   # the minimum span below the pinned end


def test_refused_configurations():
    for kw in ({"decay_seconds": 0.0}, {"attack_seconds": -1.0}, {"attack_seconds": 40.0},
               {"minimum_span_db": 0.0}, {"outlier_window_frames": 0},
               {"outlier_window_frames": 9}, {"pinned_floor_db": math.inf},
               {"pinned_floor_db": -10.0, "pinned_ceiling_db": -20.0}):
        with pytest.raises(ValueError):
            SpectrumScale(SpectrumScaleConfig(**kw))


def test_a_gap_is_crossed_in_one_step():
    s = SpectrumScale()
    s.update(-90.0, -20.0, 0.0)
    s.update(-90.0, -60.0, 600.0)
    _, ceiling = s.update(-90.0, -60.0, 600.0)
    assert abs(ceiling + 60.0) < 1e-3


def test_spectrum_geometry_is_exact():
    g = spectrum_geometry(GridParams(64, 16, 32), 2_400_000, 256)
    assert (g.transform, g.bins_per_channel, g.channels, g.bins) == (256, 128, 64, 8192)
    assert g.bin_width_hz() == 2_400_000 / (32 * 256)
    assert g.bin_zero_hz() == -2_400_000 * 65 / 128
    # the bins tile the span with no gap and nothing counted twice
    assert g.bin_hz(g.bins) - g.bin_hz(0) == pytest.approx(2_400_000)
    with pytest.raises(ValueError):
        spectrum_geometry(GridParams(64, 16, 64), 2_400_000, 256)
    with pytest.raises(ValueError):
        spectrum_geometry(GridParams(64, 16, 32), 0, 256)

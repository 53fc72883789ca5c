"""From a request to the noise kernels' parameters (noise_reference.cpp:
736-1000, notch_map.h), after tests/reference/test_noise.cpp:678-857 and
:978-1020: plan_noise cascades every enabled notch inside the passband in
order; NFM takes both notches with no stride, CW the automatic one with the
guard and a stride from the audio's reach, WFM neither; the ranges name the
notch at fault; on CW the planned guard keeps the pitch and takes a
heterodyne through the line kernel's twin; and a notch maps to its passband
offsets by mode."""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_CW, DEMOD_DSB, DEMOD_LSB, DEMOD_NFM, DEMOD_RAW,
                              DEMOD_SAM, DEMOD_USB, DEMOD_WFM, LINE_ALE, LINE_GUARD, LINE_NOTCH,
                              VrxRequest, design_notch, line_notch_sections, line_state_size,
                              plan_vrx)
from revenant_amd.noise import (NoiseRequest, NoiseRequestError, NotchSpec, auto_notch_offered,
                                notch_audio_hz, notch_offered, notch_offsets, notches_of,
                                plan_noise, validate_noise_request)
from tests.display_cases import GRID, RATE


def vrx(mode, bw=0, passband=None, pitch=0):
    return plan_vrx(GRID, RATE, VrxRequest(196_500, mode, bw, cw_pitch=pitch, passband=passband))


def test_plan_noise_cascades_every_enabled_notch_inside_the_passband():
    usb = NoiseRequest(demod=DEMOD_USB)
    notches = [NotchSpec(True, 1_000, 40.0, 100), NotchSpec(False, 1_500, 40.0, 100),
               NotchSpec(True, 2_200, 30.0, 80), NotchSpec(True, 5_000, 30.0, 80)]
    plan = vrx(DEMOD_USB, passband=(100, 2_700))
    assert plan.plan.output_rate == 48_000
    out = plan_noise(usb, plan, notches=notches)
    assert out.line and line_notch_sections(out.line_params.flags) == 2
    second = design_notch(2_200.0, 30.0, 80.0, 48_000)
    assert out.line_params.more[0].b1 == np.float32(second[1])
    assert out.line_params.more[0].a2 == np.float32(second[4])
    usb.notch_enabled, usb.notch_hz = True, 1_000
    single = plan_noise(usb, plan)
    assert single.line_params.flags == LINE_NOTCH
    assert single.line_params.b1 == out.line_params.b1
    assert notches_of(usb)[0] == NotchSpec(True, 1_000, 40.0, 100)
    assert not notches_of(usb)[1].enabled
    notches[2].depth_db = 90.0
    with pytest.raises(NoiseRequestError) as e:
        validate_noise_request(usb, notches)
    assert "notch 3" in str(e.value)
    notches[2].depth_db = 30.0
    notches[1] = NotchSpec(True, -1_500, 40.0, 100)  # the side USB discards
    with pytest.raises(NoiseRequestError):
        validate_noise_request(usb, notches)
    later = [NotchSpec(), NotchSpec(), NotchSpec(), NotchSpec(True)]
    with pytest.raises(NoiseRequestError):
        validate_noise_request(NoiseRequest(demod=DEMOD_WFM), later)


def test_nfm_takes_both_notches_cw_the_automatic_one_and_wfm_neither():
    assert notch_offered(DEMOD_NFM) and auto_notch_offered(DEMOD_NFM)
    assert notch_offered(DEMOD_CW) and auto_notch_offered(DEMOD_CW)
    assert not notch_offered(DEMOD_WFM) and not auto_notch_offered(DEMOD_WFM)
    nfm = NoiseRequest(demod=DEMOD_NFM, notch_enabled=True, notch_hz=-1_000,
                       auto_notch_enabled=True)
    validate_noise_request(nfm)
    assert notch_audio_hz(DEMOD_NFM, -1_000, 700) == 1_000.0
    out = plan_noise(nfm, vrx(DEMOD_NFM, 16_000))
    f = out.line_params.flags
    assert f & LINE_NOTCH and f & LINE_ALE and not f & LINE_GUARD
    assert out.line_params.stride == 1
    wfm = NoiseRequest(demod=DEMOD_WFM, auto_notch_enabled=True)
    with pytest.raises(NoiseRequestError):
        validate_noise_request(wfm)
    wfm = NoiseRequest(demod=DEMOD_WFM, notch_enabled=True)
    with pytest.raises(NoiseRequestError):
        validate_noise_request(wfm)
    cw = NoiseRequest(demod=DEMOD_CW, cw_pitch=700, auto_notch_enabled=True)
    out = plan_noise(cw, vrx(DEMOD_CW, passband=(-250, 250), pitch=700))
    assert out.line_params.flags & LINE_GUARD
    # the stride follows the audio's reach, the passband shifted up by the pitch:
    # 3000 Hz at a pitch of 2750, floor(48000 / 7500) = 6
    cw.cw_pitch = 2_750
    out = plan_noise(cw, vrx(DEMOD_CW, passband=(-250, 250), pitch=2_750))
    assert out.line_params.stride == 6
    cw.cw_pitch = 0
    with pytest.raises(NoiseRequestError) as e:
        validate_noise_request(cw)
    assert "automatic notch" in str(e.value)


def test_ranges():
    for kw in ({"nb_threshold_db": 2.0}, {"nb_threshold_db": 41.0}, {"nr_strength": 1.5},
               {"profile_strength_db": 21.0}, {"notch_depth_db": 2.0}, {"notch_width_hz": 5},
               {"demod": DEMOD_RAW, "nb_enabled": True}, {"demod": DEMOD_RAW, "nr_enabled": True}):
        with pytest.raises(NoiseRequestError):
            validate_noise_request(NoiseRequest(**kw))
    out = plan_noise(NoiseRequest(demod=DEMOD_AM, nb_enabled=True, nr_enabled=True,
                                  nr_strength=1.0), vrx(DEMOD_AM, 10_000))
    assert out.blank and out.blank_threshold == np.float32(10 ** 1.2)
    assert out.spectral and out.spectral_params.alpha == 2.0 and not out.line


def test_on_cw_the_planned_guard_keeps_the_pitch_and_removes_a_heterodyne():
    req = NoiseRequest(demod=DEMOD_CW, cw_pitch=700, notch_width_hz=100,
                       auto_notch_enabled=True)
    lp = plan_noise(req, vrx(DEMOD_CW, passband=(-250, 250), pitch=700)).line_params
    count = 96_000
    t = np.arange(count) / 48_000.0
    wanted = (0.3 * np.sin(2 * np.pi * 700.0 * t)).astype(np.float32)
    whistle = (0.3 * np.sin(2 * np.pi * 1300.0 * t)).astype(np.float32)

    def run(params):
        c = _lib.LineConfig()
        out = (wanted + whistle).copy()
        params.count = count
        oracle.noise_line(c, params, out, np.zeros(line_state_size(c), np.float32))
        return out

    def kept(out, comp):
        h = slice(count // 2, None)
        c = comp[h].astype(np.float64)
        return 20 * np.log10(abs((out[h] * c).sum() / (c * c).sum()))

    guarded = run(lp)
    assert kept(guarded, wanted) > -1.0 and kept(guarded, whistle) < -6.0
    bare = _lib.LineParams.from_buffer_copy(bytes(lp))
    bare.flags &= ~LINE_GUARD
    assert kept(run(bare), wanted) < -6.0


def test_a_notch_maps_to_its_passband_offsets_by_mode():
    assert notch_offsets(DEMOD_USB, 1_200.0).hz == [1_200.0]
    assert notch_offsets(DEMOD_LSB, 1_200.0).hz == [-1_200.0]
    for m in (DEMOD_AM, DEMOD_SAM, DEMOD_DSB):
        o = notch_offsets(m, 900.0)
        assert o.hz == [-900.0, 900.0] and not o.audio_only
    cw = notch_offsets(DEMOD_CW, 800.0, 700.0)
    assert cw.hz == [100.0] and notch_audio_hz(DEMOD_CW, 100, 700) == 800.0
    for m in (DEMOD_NFM, DEMOD_WFM):
        o = notch_offsets(m, 1_000.0)
        assert o.count == 0 and o.audio_only
    raw = notch_offsets(DEMOD_RAW, 1_000.0)
    assert raw.count == 0 and not raw.audio_only
    assert notch_offsets(DEMOD_USB, 0.0).count == 0
    assert notch_audio_hz(DEMOD_USB, 1_200, 0) == 1_200.0
    assert notch_audio_hz(DEMOD_LSB, -1_200, 0) == 1_200.0

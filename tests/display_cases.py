"""Receivers whose display taps the tests plan (after
tests/reference/test_vrx.cpp:1360-1383): the canonical 64-channel grid at
2.4 MS/s (75 kS/s channels), and each case's expected decimation."""
from revenant_amd.dsp import (DEMOD_AM, DEMOD_CW, DEMOD_NFM, DEMOD_USB, GridParams, VrxRequest,
                              plan_vrx)

RATE = 2_400_000
GRID = GridParams(64, 17, 32)


def display_cases():
    """(name, plan, expected R)."""
    def plan(mode, bw=0, pitch=0, passband=None):
        return plan_vrx(GRID, RATE, VrxRequest(196_500, mode, bw, cw_pitch=pitch,
                                               passband=passband))
    return [("nfm 16 kHz", plan(DEMOD_NFM, 16_000), 1),
            ("nfm 9 kHz", plan(DEMOD_NFM, 9_000), 2),
            ("nfm 4 kHz", plan(DEMOD_NFM, 4_000), 4),
            ("am 2 kHz", plan(DEMOD_AM, 2_000), 8),
            ("nfm 900 Hz", plan(DEMOD_NFM, 900), 20),
            ("cw 500 Hz at 700", plan(DEMOD_CW, 500, pitch=700), 8),
            ("usb 300 to 2700", plan(DEMOD_USB, passband=(300, 2_700)), 2)]

"""Pins the deterministic transcendentals (vrx_reference.cpp:227-306,
spectrum_reference.cpp:123) against libm on dense sweeps
(tests/reference/test_vrx.cpp:301, test_spectrum.cpp:419)."""
import numpy as np

from oracle import oracle


def test_log2():
    x = np.geomspace(1e-20, 1e30, 200001).astype(np.float32)
    x = np.concatenate([x, np.float32([1.0, 2.0, 0.5, 1.4142135, 1.4142137, 3.0e38])])
    got = oracle.det("log2", x).astype(np.float64)
    ref = np.log2(x.astype(np.float64))
    # the integer part is exact; the series costs a few 1e-7, the final sum rounds once
    bound = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 3e-7
    assert np.all(np.abs(got - ref) <= bound)
    # exact at powers of two
    p = np.float32(2.0) ** np.arange(-60, 60).astype(np.float32)
    assert np.array_equal(oracle.det("log2", p), np.arange(-60, 60).astype(np.float32))


def test_recip_sqrt_rsqrt():
    x = np.geomspace(1e-30, 1e30, 100001).astype(np.float32)
    xd = x.astype(np.float64)
    assert np.max(np.abs(oracle.det("recip", x) * xd - 1.0)) < 3e-7
    assert np.max(np.abs(oracle.det("rsqrt", x) * np.sqrt(xd) - 1.0)) < 3e-7
    assert np.max(np.abs(oracle.det("sqrt", x) / np.sqrt(xd) - 1.0)) < 3e-7
    assert oracle.det("sqrt", np.float32([0.0, -1.0, np.nan])).tolist() == [0.0, 0.0, 0.0]


def test_atan2_all_quadrants():
    g = np.random.default_rng(7)
    y = g.uniform(-3, 3, 200000).astype(np.float32)
    x = g.uniform(-3, 3, 200000).astype(np.float32)
    got = oracle.det("atan2", y, x).astype(np.float64)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 1e-6
    assert oracle.det("atan2", np.float32([0.0]), np.float32([0.0]))[0] == 0.0
    axes = oracle.det("atan2", np.float32([1, 0, -1, 0]), np.float32([0, 1, 0, -1]))
    assert np.allclose(axes, [np.pi / 2, 0, -np.pi / 2, np.pi], atol=1e-6)

"""Pins the oracle's render colours (FluidSimCPU::updateColors,
fluidSimCPU.cc:100-125, over getSpeedNormalzied, physicsWorld.cc:178-182)
by what the gradient fixes: the stop colours exactly at speed 0, at the two
breakpoints (|v| = 0.495 and 0.99 give normalized 0.33 and 0.66 up to float
rounding, checked against the stops to 1e-6), at and beyond the 1.5 clamp;
continuity across the breakpoints; monotone interpolation between
neighbouring stops; alpha = 1 throughout; NaN propagation; and the stop
defaults of fluidSimCPU.h:26-29."""
import numpy as np

import pyoracle as O

C = O.REFERENCE_COLOR_STOPS.reshape(4, 4)


def _col(speeds, axis=0, stops=None):
    v = np.zeros((len(speeds), 4), np.float32)
    v[:, axis] = np.asarray(speeds, np.float32)
    return O.speed_colors(v, stops)


def test_default_stops_are_the_reference_colours():
    np.testing.assert_array_equal(C[0], [0.0, 0.75, 1.0, 1.0])
    np.testing.assert_array_equal(C[1], [0.0, 1.0, 0.0, 1.0])
    np.testing.assert_array_equal(C[2], [1.0, 1.0, 0.0, 1.0])
    np.testing.assert_array_equal(C[3], [1.0, 0.0, 0.0, 1.0])


def test_end_stops_and_clamp():
    c = _col([0.0, -0.0, 1.5, 2.0, 1e30])
    np.testing.assert_array_equal(c[0], C[0])            # at rest: Color1
    np.testing.assert_array_equal(c[1], C[0])
    for k in (2, 3, 4):                                   # at or beyond the clamp: Color4
        np.testing.assert_allclose(c[k], C[3], atol=1e-6)
    # the sign of the velocity does not matter (length)
    np.testing.assert_array_equal(_col([-0.7])[0], _col([0.7])[0])
    np.testing.assert_array_equal(_col([0.7], axis=2)[0], _col([0.7])[0])


def test_breakpoints_and_continuity():
    eps = np.float32(1e-6)
    for b, stop in ((0.33, 1), (0.66, 2)):
        s = np.float32(b) * np.float32(1.5)
        c = _col([s - eps * 4, s, s + eps * 4])
        for k in range(3):
            np.testing.assert_allclose(c[k], C[stop], atol=2e-5)


def test_interpolation_is_between_neighbouring_stops():
    speeds = np.linspace(0.0, 1.5, 2001, dtype=np.float32)
    c = _col(speeds)
    n = speeds / np.float32(1.5)
    seg = np.where(n <= 0.33, 0, np.where(n <= 0.66, 1, 2))
    for k in range(len(speeds)):
        a, b = C[seg[k]], C[seg[k] + 1]
        lo, hi = np.minimum(a, b) - 1e-6, np.maximum(a, b) + 1e-6
        assert np.all(c[k] >= lo) and np.all(c[k] <= hi), (speeds[k], c[k])
    np.testing.assert_allclose(c[:, 3], 1.0, atol=1e-6)
    # each channel is monotone within a segment
    for s in range(3):
        m = seg == s
        d = np.diff(c[m], axis=0)
        for ch in range(4):
            assert np.all(d[:, ch] >= -1e-7) or np.all(d[:, ch] <= 1e-7)


def test_custom_stops_and_midpoints():
    stops = np.arange(16, dtype=np.float32)
    c = _col([0.0, np.float32(0.165) * np.float32(1.5)], stops=stops)
    np.testing.assert_array_equal(c[0], stops[0:4])
    np.testing.assert_allclose(c[1], 0.5 * (stops[0:4] + stops[4:8]), atol=1e-5)


def test_nan_velocity_gives_nan_colour():
    c = _col([np.nan])
    assert np.all(np.isnan(c[0]))

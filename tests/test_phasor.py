"""exp(i theta) of the batched Horner kernels (host build of the same routine).

Pinned by the mathematics: against the C library's correctly reduced sin/cos
over the argument range the kernels see (|2 pi w nu / c| up to ~1e6 rad),
the identity sin^2 + cos^2 = 1, exact special angles, symmetry, and the
quadrant boundaries where the reduction switches branches.
"""
import numpy as np
import pytest


def test_phasor_matches_libm(leap):
    rng = np.random.default_rng(5)
    for scale in (1.0, 10.0, 1e3, 3e4, 1e6):
        th = rng.uniform(-scale, scale, 20000)
        c, s = leap.phasor64(th)
        err = max(np.abs(c - np.cos(th)).max(), np.abs(s - np.sin(th)).max())
        assert err < 6e-16, (scale, err)
        assert np.abs(c * c + s * s - 1).max() < 6e-16


def test_phasor_special_angles(leap):
    th = np.array([0.0, -0.0, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi, -np.pi / 2, np.pi / 4, np.pi / 6])
    c, s = leap.phasor64(th)
    assert c[0] == 1.0 and s[0] == 0.0
    np.testing.assert_allclose(c, np.cos(th), atol=4e-16)
    np.testing.assert_allclose(s, np.sin(th), atol=4e-16)
    assert abs(s[8] - 0.5) < 2e-16 and abs(c[7] - s[7]) < 2e-16


@pytest.mark.parametrize("k", [1, 2, 3, 7, 1000, 123457])
def test_phasor_quadrant_boundaries(leap, k):
    # arguments on both sides of (k + 1/2) pi / 2, where rint switches quadrant
    mid = (k + 0.5) * np.pi / 2
    th = mid + np.array([-1e-9, -1e-15 * mid, 0.0, 1e-15 * mid, 1e-9])
    c, s = leap.phasor64(th)
    np.testing.assert_allclose(c, np.cos(th), atol=6e-16)
    np.testing.assert_allclose(s, np.sin(th), atol=6e-16)
    # odd / even symmetry
    c2, s2 = leap.phasor64(-th)
    np.testing.assert_array_equal(c2, c)
    np.testing.assert_array_equal(s2, -s)

"""Pins of the oracle's inversion tail and short-baseline filter by answers the
mathematics fixes (no GPU needed).

The inversion (algorithm/cpu/CpuLeapCalibrator.cc:159-178) is checked against
  * a three-antenna set with autocorrelations solved by hand (tests/planted.py),
    which tells row n of avgData from row I(n) (:171), the factor 2 pi on A cal1
    (:171), the sign of the cal1 term and the appended 0 (:174-178);
  * the closed form of a planted per-antenna phase screen on a connected array
    without autocorrelations or flags;
  * LAPACK's pseudo-inverse applied to dInt built from the planted phases (not
    from arg of avgData) on sets with autocorrelations, flags and wrapping
    phases.
Each test also checks that the plausible misreadings give a different answer,
so that a mistake of that kind cannot pass.

The short-baseline filter (ms/MeasurementSet.cc:205-228) is checked against a
hand-written UVW table whose expected flags are derived by hand.
"""
import numpy as np
import pytest

import planted as pl


def wrapped_max(a, b):
    return float(np.abs(pl.wrap(np.asarray(a) - np.asarray(b))).max())


# ---------------------------------------------------------------------------
# Three antennas with autocorrelations, solved by hand
# ---------------------------------------------------------------------------
def hand_ms():
    ms = pl.inversion_ms(3, autos=True)
    assert list(zip(ms["antenna1"], ms["antenna2"])) == [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    return ms


def test_hand_set_indices(oracle):
    md = oracle.metadata(hand_ms())
    assert md.reference_antenna == 2
    assert md.I.tolist() == [1, 2, 4]
    assert md.I1.tolist() == [2, 4]
    assert md.A.tolist() == [[1, -1, 0], [1, 0, -1], [0, 1, -1], [1, 0, 0]]
    assert md.A1.tolist() == [[1, 0, -1], [0, 1, -1], [0, 0, 1]]


def test_hand_set_inversion(oracle):
    md = oracle.metadata(hand_ms())
    avg = np.zeros((6, 4), dtype=np.complex128)
    avg[:, 0] = pl.HAND_AVG
    avg[:, 1:] = 7.0 - 3.0j  # only pol 0 reaches the calibration
    want = pl.hand_cal(pl.HAND_AVG)
    got = md.apply_inversion(avg)
    assert np.abs(got - want).max() < 1e-14
    # the hand numbers: cal1 = [0.1, -0.2, 0], A cal1 = [0.3, 0.1, -0.2],
    # dInt = [-0.6 pi, 0.3 - 0.2 pi, 0.1 + 0.4 pi]
    d = np.array([-0.6 * np.pi, 0.3 - 0.2 * np.pi, 0.1 + 0.4 * np.pi])
    by_hand = np.array([0.1, -0.2 + (-2 * d[0] - d[1] + d[2]) / 3, (-d[0] - 2 * d[1] - d[2]) / 3])
    assert np.abs(want - by_hand).max() < 1e-15
    # each misreading moves the answer by far more than the tolerance
    for wrong in (pl.hand_cal(pl.HAND_AVG, rows=(1, 2, 4)),     # avgData row I(n)
                  pl.hand_cal(pl.HAND_AVG, two_pi=1.0),         # no extra 2 pi
                  pl.hand_cal(pl.HAND_AVG, two_pi=-pl.TWO_PI),  # sign of the phase term
                  pl.hand_cal(pl.HAND_AVG, cal1_sign=-1.0)):    # Ad dphi - cal1
        assert wrapped_max(got, wrong) > 0.05


def test_hand_set_without_the_appended_zero_differs(oracle):
    """The constraint row's entry of dphi is 0 (:174); a solve that dropped the
    constraint row (least squares on the baseline rows only, minimum norm)
    gives another answer."""
    md = oracle.metadata(hand_ms())
    avg = np.zeros((6, 4), dtype=np.complex128)
    avg[:, 0] = pl.HAND_AVG
    got = md.apply_inversion(avg)
    p = np.angle(pl.HAND_AVG)
    cal1 = np.array([p[2], p[4], 0.0])
    acal1 = md.A[:3] @ cal1
    dint = pl.wrap(p[:3] - pl.TWO_PI * acal1)
    wrong = np.linalg.pinv(md.A[:3]) @ dint + cal1
    assert wrapped_max(got, wrong) > 0.05


# ---------------------------------------------------------------------------
# Planted phase screens
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ref_ant", [None, 0, 5])
def test_planted_closed_form(oracle, ref_ant):
    """No autocorrelations, no flags, baseline phases too small to wrap: cal is
    (2 - 2 pi)(phi - phi[r]) - (1 - 2 pi)(phi[0] - phi[r])."""
    n_ant = 12
    ms = pl.inversion_ms(n_ant, autos=False)
    rng = np.random.default_rng(11 + (ref_ant or 0))
    phi = rng.uniform(-0.25, 0.25, n_ant)
    avg = pl.planted_avg(ms, phi, rng)
    md = oracle.metadata(ms, ref_ant=ref_ant)
    r = md.reference_antenna
    assert r == (n_ant - 1 if ref_ant is None else ref_ant)
    assert np.array_equal(md.I, np.arange(ms["num_baselines"]))  # I(n) = n
    phi1 = np.append(np.angle(avg[md.I1, 0]), 0.0)
    assert np.abs(md.Ad1 @ phi1 - (phi - phi[r])).max() < 1e-12
    got = md.apply_inversion(avg)
    want = pl.closed_form_cal(phi, r)
    assert np.abs(got - want).max() < 1e-12
    # a mistake in the 2 pi factor or the cal1 sign does not satisfy the closed form
    assert np.abs(pl.planted_cal(ms, phi, r, md.A, md.I, two_pi=1.0) - want).max() > 1e-2
    assert np.abs(pl.planted_cal(ms, phi, r, md.A, md.I, cal1_sign=-1.0) - want).max() > 1e-2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_planted_with_autocorrelations_flags_and_wraps(oracle, seed):
    """Autocorrelations and flagged baselines (so I(n) != n), phases over the
    whole circle (so 2 pi A cal1 and A cal1 wrap differently): the oracle equals
    LAPACK's least squares of the planted dInt, and not any misreading."""
    n_ant = 10
    rng = np.random.default_rng(seed)
    base = pl.inversion_ms(n_ant, autos=True)
    nbl = base["num_baselines"]
    cross = np.nonzero(base["antenna1"] != base["antenna2"])[0]
    flagged = rng.choice(cross, 4, replace=False)
    ms = pl.inversion_ms(n_ant, autos=True, flagged=flagged)
    phi = rng.uniform(-np.pi, np.pi, n_ant)
    avg = pl.planted_avg(ms, phi, rng)
    for ref_ant in (None, 3):
        md = oracle.metadata(ms, ref_ant=ref_ant)
        r = md.reference_antenna
        assert not np.array_equal(md.I, np.arange(len(md.I)))
        got = md.apply_inversion(avg)
        want = pl.planted_cal(ms, phi, r, md.A, md.I)
        assert wrapped_max(got, want) < 1e-10
        wrongs = {
            "row I(n)": pl.planted_cal(ms, phi, r, md.A, md.I, delta_rows=md.I),
            "no 2 pi": pl.planted_cal(ms, phi, r, md.A, md.I, two_pi=1.0),
            "phase sign": pl.planted_cal(ms, phi, r, md.A, md.I, two_pi=-pl.TWO_PI),
            "minus cal1": pl.planted_cal(ms, phi, r, md.A, md.I, cal1_sign=-1.0),
        }
        for name, wrong in wrongs.items():
            assert wrapped_max(got, wrong) > 1e-3, name
        assert nbl == n_ant * (n_ant + 1) // 2


def test_planted_amplitudes_do_not_matter(oracle):
    """arg is scale invariant: the sum (not mean) of avgData (CpuLeapCalibrator.cc:212)
    and per-baseline gains leave cal unchanged."""
    ms = pl.inversion_ms(8, autos=True, flagged=(3,))
    rng = np.random.default_rng(5)
    phi = rng.uniform(-np.pi, np.pi, 8)
    avg = pl.planted_avg(ms, phi, rng)
    md = oracle.metadata(ms)
    a = md.apply_inversion(avg)
    b = md.apply_inversion(avg * rng.uniform(0.1, 50.0, (ms["num_baselines"], 1)))
    assert np.abs(a - b).max() < 1e-12


# ---------------------------------------------------------------------------
# Short baselines on a hand-written UVW table
# ---------------------------------------------------------------------------
def short_ms():
    """Three antennas without autocorrelations (baselines (0,1), (0,2), (1,2)),
    two timesteps.  v and w are 5 km on every row, so a filter on the true
    baseline length sqrt(u^2 + v^2) would flag nothing."""
    ms = pl.inversion_ms(3, autos=False)
    ms["num_rows"] = 6
    ms["antenna1"] = np.tile(ms["antenna1"], 2)
    ms["antenna2"] = np.tile(ms["antenna2"], 2)
    ms["time"] = np.repeat([0.0, 1.0], 3)
    ms["vis"] = np.zeros((6, 1, 4), dtype=np.complex128)
    u = np.array([30.0, 40.0, 500.0, 1200.0, 7.0, 9.0])
    ms["uvw"] = np.column_stack([u, np.full(6, 5000.0), np.full(6, 5000.0)])
    return ms


def test_short_baselines_hand_table(oracle):
    """The reference reads u only, into a (1, rows) matrix indexed as uv(i, 0)
    and uv(i, 1): the u of rows i and i + 1.  Lengths by hand: baseline 0
    sqrt(30^2 + 40^2) = 50, baseline 1 sqrt(40^2 + 500^2) = 501.60, baseline 2
    sqrt(500^2 + 1200^2) = 1300 (row 3 belongs to the second timestep)."""
    ms = short_ms()
    assert oracle.filtered_baselines(ms, 0.0).tolist() == [False, False, False]
    assert oracle.filtered_baselines(ms, 50.0).tolist() == [False, False, False]  # strict <
    assert oracle.filtered_baselines(ms, 50.01).tolist() == [True, False, False]
    assert oracle.filtered_baselines(ms, 600.0).tolist() == [True, True, False]
    assert oracle.filtered_baselines(ms, 1300.5).tolist() == [True, True, True]
    ms["flag"] = np.array([0, 0, 1], dtype=np.uint8)  # FLAG of the first epoch is OR-ed in
    assert oracle.filtered_baselines(ms, 600.0).tolist() == [True, True, True]
    assert oracle.filtered_baselines(ms, 0.0).tolist() == [False, False, True]


def test_short_baselines_feed_the_phase_matrix(oracle):
    ms = short_ms()
    md = oracle.metadata(ms, ref_ant=2, threshold=60.0)  # baseline 0 (0,1) is short
    assert md.I.tolist() == [1, 2]
    assert md.A.tolist() == [[1, 0, -1], [0, 1, -1], [1, 0, 0]]


def test_degeneracy_rule(oracle):
    """(Ad A).isApprox(I, 0.001) (MetaData.cc:129-136) fails exactly when the
    baseline graph has a component without the constraint antenna: Ad A is then
    the projection I - sum_S 1_S 1_S^T / |S|, at Frobenius distance sqrt(#S) from I."""
    base = pl.inversion_ms(6, autos=True)
    a1, a2 = base["antenna1"], base["antenna2"]
    touches3 = np.nonzero(((a1 == 3) | (a2 == 3)) & (a1 != a2))[0]
    b03 = int(np.nonzero((a1 == 0) & (a2 == 3))[0][0])
    assert oracle.metadata(base, ref_ant=0).degenerate() == (False, False)
    iso = oracle.metadata(pl.inversion_ms(6, autos=True, flagged=touches3), ref_ant=0)
    assert iso.degenerate() == (True, True)
    P = iso.Ad @ iso.A
    assert abs(np.linalg.norm(P - np.eye(6)) - 1.0) < 1e-12  # one null component: antenna 3
    assert oracle.metadata(pl.inversion_ms(6, autos=True, flagged=(b03,)), ref_ant=0).degenerate() == (False, True)

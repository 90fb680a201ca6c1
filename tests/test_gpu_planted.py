"""The HIP inversion and short-baseline filter against answers the mathematics
fixes (tests/planted.py), not only against the oracle: the hand-solved
three-antenna set with autocorrelations, the closed form of a planted phase
screen, and LAPACK's least squares of the planted dInt
(algorithm/cpu/CpuLeapCalibrator.cc:159-178, ms/MeasurementSet.cc:205-228).

Bars: FP64 routes within 1e-10 rad (wrapped), the split-3 x BF16 MFMA route
within 2e-6 rad; flags and phase-matrix indices exact.
"""
import numpy as np
import pytest

import planted as pl

pytestmark = pytest.mark.gpu

ROUTES = [(0, 0, 1e-10), (0, 1, 1e-10), (1, 1, 2e-6)]
ROUTE_IDS = ["fp64-gemv", "fp64-gemm", "bf16x3-mfma"]


def wrapped_max(a, b):
    return float(np.abs(pl.wrap(np.asarray(a) - np.asarray(b))).max())


def opts_for(leap, precision, batched):
    return leap.options(precision=precision,
                        strategy=leap.LEAP_STRATEGY_BATCHED if batched else leap.LEAP_STRATEGY_PER_DIRECTION)


@pytest.mark.parametrize("precision,batched,tol", ROUTES, ids=ROUTE_IDS)
def test_hand_set_inversion(leap, ctx, precision, batched, tol):
    ms = pl.inversion_ms(3, autos=True)
    md = ctx.metadata(leap.view_from_arrays(ms))
    got_md = md.copy()
    assert got_md["I"].tolist() == [1, 2, 4] and got_md["I1"].tolist() == [2, 4]
    avg = np.zeros((2, 6, 4), dtype=np.complex128)
    avg[0, :, 0] = pl.HAND_AVG
    avg[1, :, 0] = np.conj(pl.HAND_AVG[::-1])
    avg[:, :, 1:] = 2.0 + 1.0j
    got = md.apply_inversion(avg, opts=opts_for(leap, precision, batched))
    assert wrapped_max(got[0], pl.hand_cal(pl.HAND_AVG)) < tol
    assert wrapped_max(got[1], pl.hand_cal(avg[1, :, 0])) < tol
    assert wrapped_max(got[0], pl.hand_cal(pl.HAND_AVG, rows=(1, 2, 4))) > 0.05


@pytest.mark.parametrize("precision,batched,tol", ROUTES, ids=ROUTE_IDS)
def test_planted_closed_form(leap, ctx, precision, batched, tol):
    n_ant = 24
    ms = pl.inversion_ms(n_ant, autos=False)
    rng = np.random.default_rng(21)
    phis = rng.uniform(-0.25, 0.25, (5, n_ant))
    avg = np.stack([pl.planted_avg(ms, phi, rng) for phi in phis])
    md = ctx.metadata(leap.view_from_arrays(ms), ref_ant=7)
    got = md.apply_inversion(avg, opts=opts_for(leap, precision, batched))
    for d in range(len(phis)):
        assert np.abs(got[d] - pl.closed_form_cal(phis[d], 7)).max() < tol


@pytest.mark.parametrize("precision,batched,tol", ROUTES, ids=ROUTE_IDS)
def test_planted_with_autocorrelations_flags_and_wraps(leap, ctx, precision, batched, tol):
    n_ant = 20
    rng = np.random.default_rng(22)
    base = pl.inversion_ms(n_ant, autos=True)
    cross = np.nonzero(base["antenna1"] != base["antenna2"])[0]
    ms = pl.inversion_ms(n_ant, autos=True, flagged=rng.choice(cross, 9, replace=False))
    phis = rng.uniform(-np.pi, np.pi, (4, n_ant))
    avg = np.stack([pl.planted_avg(ms, phi, rng) for phi in phis])
    for ref in (None, 2):
        md = ctx.metadata(leap.view_from_arrays(ms), ref_ant=ref)
        m = md.copy()
        got = md.apply_inversion(avg, opts=opts_for(leap, precision, batched))
        for d in range(len(phis)):
            want = pl.planted_cal(ms, phis[d], md.reference_antenna, m["A"], m["I"])
            assert wrapped_max(got[d], want) < tol


def test_short_baselines_hand_table(leap, ctx):
    """u of rows i and i + 1 (the reference's (1, rows) matrix read as uv(i, 0), uv(i, 1)):
    baseline lengths 50, 501.6 and 1300 m; v = 5 km is never read."""
    ms = pl.inversion_ms(3, autos=False)
    ms["num_rows"] = 6
    ms["antenna1"] = np.tile(ms["antenna1"], 2)
    ms["antenna2"] = np.tile(ms["antenna2"], 2)
    ms["time"] = np.repeat([0.0, 1.0], 3)
    ms["vis"] = np.zeros((6, 1, 4), dtype=np.complex128)
    ms["uvw"] = np.column_stack([[30.0, 40.0, 500.0, 1200.0, 7.0, 9.0], np.full(6, 5000.0), np.full(6, 5000.0)])
    view = leap.view_from_arrays(ms)
    for thr, I in ((0.0, [0, 1, 2]), (50.01, [1, 2]), (600.0, [2])):
        m = ctx.metadata(view, ref_ant=2, threshold=thr).copy()
        assert m["I"].tolist() == I, thr

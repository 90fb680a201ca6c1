"""Parity of the HIP calibrator (through the C ABI) with the CPU oracle.

Bars (written in each test): A, I, A1, I1 and flags bit-exact; pseudo-inverses
within 1e-11; avgData within 1e-12 of its largest magnitude; calibrations
within 1e-9 rad after 2 pi wrapping.
"""
import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu

PHASE_TOL = 1e-9  # rad, wrapped


def wrapped_err(a, b):
    return float(np.abs(npr.wrap(np.asarray(a) - np.asarray(b))).max())


def directions_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


def make_ms(leap, dtype=np.complex128, **kw):
    s = leap.SynthMS(**kw)
    s.fill_host(dtype)
    return s


@pytest.fixture(scope="module")
def c0(leap):
    """The 128-antenna MWA-like set: 1.5 km disk, 8256 baselines with autos, 14 x 2 s,
    48 channels from 169.6 MHz at 40 kHz, 4 pols, complex128, 1% flagged baselines."""
    return make_ms(leap, n_antennas=128, n_timesteps=14, n_channels=48, flag_fraction=0.01, seed=20201)


def test_dd_kernel_matches_oracle(ctx, oracle):
    centre = (0.57595865315812877, 0.10471975511965978)
    dirs = directions_near(centre, 64, 1, 0.5)
    dd = ctx.dd(centre, dirs)
    for d in range(len(dirs)):
        assert np.abs(dd[d] - oracle.dd(centre, dirs[d])).max() < 2e-16
    gold = np.array([[0.50913780874486769, -0.089966081772685239, 0.85597009050371897],
                     [-0.2520402307174327, 0.93533988977932658, 0.24822371499818516],
                     [-0.82295468514759529, -0.34211897743046571, 0.45354182990718139]])
    g = ctx.dd(centre, [(-0.4606549305661674, -0.29719233792392513)])[0]
    assert np.abs(g - gold).max() < 4 * np.spacing(1.0)


@pytest.mark.parametrize("kw,ref", [
    (dict(n_antennas=32, n_timesteps=2, n_channels=4, flag_fraction=0.05, seed=1), None),
    (dict(n_antennas=32, n_timesteps=2, n_channels=4, flag_fraction=0.05, seed=1), 0),
    (dict(n_antennas=40, autocorrelations=0, n_timesteps=2, n_channels=4, flag_fraction=0.1, seed=2), 5),
])
def test_metadata_matches_oracle(leap, ctx, oracle, kw, ref):
    s = make_ms(leap, **kw)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view, ref_ant=ref)
    got = md.copy()
    want = oracle.metadata(ms, ref_ant=ref)
    assert md.reference_antenna == want.reference_antenna
    for name in ("A", "I", "A1", "I1"):
        assert np.array_equal(got[name], getattr(want, name)), name
    assert np.abs(got["Ad"] - want.Ad).max() < 1e-11
    assert np.abs(got["Ad1"] - want.Ad1).max() < 1e-11


def test_metadata_short_baseline_threshold(leap, ctx, oracle):
    s = make_ms(leap, n_antennas=20, n_timesteps=3, n_channels=2, seed=5)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    thr = 300.0
    filtered = oracle.filtered_baselines(ms, thr)
    assert filtered.sum() > ms["flag"].sum()
    # the default reference antenna (19) indexes a filtered baseline: both sides refuse
    assert filtered[19]
    with pytest.raises(leap.LeapError, match="refAnt 19 is flagged"):
        ctx.metadata(view, threshold=thr)
    ref = int(np.nonzero(~filtered[:20])[0][0])
    md = ctx.metadata(view, ref_ant=ref, threshold=thr)
    want = oracle.metadata(ms, ref_ant=ref, threshold=thr)
    got = md.copy()
    assert np.array_equal(got["I"], want.I) and np.array_equal(got["I1"], want.I1)
    assert np.array_equal(got["A"], want.A) and np.array_equal(got["A1"], want.A1)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("pols", [4, 2, 1])
def test_rotate_matches_oracle(leap, ctx, oracle, dtype, pols):
    s = make_ms(leap, dtype, n_antennas=24, n_timesteps=5, n_channels=37, n_pols=pols, seed=100 + pols)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = directions_near(ms["phase_centre"], 3, pols)
    got = md.rotate(view, dirs)
    for d in range(len(dirs)):
        want = omd.rotate(dirs[d])
        assert np.abs(got[d] - want).max() <= 1e-12 * np.abs(want).max()


def test_rotate_nan_rule_and_row_ranges(leap, ctx, oracle):
    s = make_ms(leap, n_antennas=16, n_timesteps=6, n_channels=300, nan_fraction=0.01, seed=8)
    ms = s.host_dict()
    assert np.isnan(ms["vis"]).any()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    d = directions_near(ms["phase_centre"], 1, 3)
    nbl = ms["num_baselines"]
    for row_start, n_rows in ((0, None), (2 * nbl, 3 * nbl)):
        got = md.rotate(view, d, row_start, n_rows)[0]
        want = omd.rotate(d[0], row_start, n_rows)
        assert np.isfinite(got).all()
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_rotate_is_deterministic(leap, ctx):
    s = make_ms(leap, n_antennas=20, n_timesteps=4, n_channels=64, seed=9)
    view = leap.view_from_arrays(s.host_dict())
    md = ctx.metadata(view)
    d = directions_near(s.host_dict()["phase_centre"], 2, 4)
    a = md.rotate(view, d)
    b = md.rotate(view, d)
    assert np.array_equal(a, b)


def test_inversion_matches_oracle(leap, ctx, oracle):
    s = make_ms(leap, n_antennas=48, n_timesteps=2, n_channels=8, flag_fraction=0.05, seed=10)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    for ref in (None, 0, 17):
        md = ctx.metadata(view, ref_ant=ref)
        omd = oracle.metadata(ms, ref_ant=ref)
        rng = np.random.default_rng(ref or 1)
        avg = rng.standard_normal((3, ms["num_baselines"], 4)) + 1j * rng.standard_normal((3, ms["num_baselines"], 4))
        got = md.apply_inversion(avg)
        for d in range(3):
            assert wrapped_err(got[d], omd.apply_inversion(avg[d])) < 1e-10


@pytest.mark.parametrize("slice_", [(0, 1, 1), (0, 14, 14), (0, -1, -1), (0, 1, -1), (0, 4, -1)],
                         ids=["first", "all14", "all", "each", "every4"])
def test_calibrate_c0_matches_oracle(leap, ctx, oracle, c0, slice_):
    """End-to-end on the 128-antenna set with the PhaseRotateTests slices and four directions."""
    ms = c0.host_dict()
    view = leap.view_from_arrays(ms)
    dirs = directions_near(ms["phase_centre"], 2, 7)
    dirs = np.vstack([dirs, dirs])  # repeated directions, as PhaseRotateTests.cc:97-103
    got, ge = ctx.calibrate(view, dirs, slice_, ref_ant=0)
    want, we = oracle.calibrate(ms, dirs, slice_, ref_ant=0)
    assert got.shape == want.shape
    assert np.array_equal(ge, we)
    assert wrapped_err(got, want) < PHASE_TOL
    assert np.array_equal(got[:, 0], got[:, 2])


def test_calibrate_default_reference_and_c64(leap, ctx, oracle):
    s = make_ms(leap, np.complex64, n_antennas=40, n_timesteps=3, n_channels=20, flag_fraction=0.02, seed=12)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    dirs = directions_near(ms["phase_centre"], 3, 8)
    got, _ = ctx.calibrate(view, dirs)
    want, _ = oracle.calibrate(ms, dirs)
    assert wrapped_err(got, want) < PHASE_TOL


def test_calibrate_device_resident_cube(leap, ctx, oracle):
    """Visibilities generated on the device; parity against the oracle on the same bytes."""
    s = leap.SynthMS(n_antennas=64, n_timesteps=6, n_channels=96, flag_fraction=0.02, seed=14)
    nbytes = s.num_rows * s.num_channels * s.num_pols * 16
    vis_p = ctx.malloc(nbytes)
    uvw_p = ctx.malloc(s.num_rows * 3 * 8)
    try:
        ctx.synth_fill_device(s, leap.LEAP_VIS_C128, vis_p, uvw_p)
        host = np.zeros((s.num_rows, s.num_channels, s.num_pols), dtype=np.complex128)
        ctx.d2h(host, vis_p)
        ref_host = s.fill_host()
        assert np.abs(host - ref_host).max() < 1e-9 * np.abs(ref_host).max()
        ms = s.host_dict(host)
        view = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=leap.LEAP_VIS_C128, uvw_device=uvw_p)
        dirs = directions_near(ms["phase_centre"], 2, 9)
        got, _ = ctx.calibrate(view, dirs, (0, 3, -1))
        want, _ = oracle.calibrate(ms, dirs, (0, 3, -1))
        assert wrapped_err(got, want) < PHASE_TOL
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)


def test_errors_are_reported_not_fatal(leap, ctx):
    s = make_ms(leap, n_antennas=8, n_timesteps=4, n_channels=2, seed=15)
    view = leap.view_from_arrays(s.host_dict())
    with pytest.raises(leap.LeapError, match="expected a non zero integer"):
        ctx.calibrate(view, [(0.5, -0.4)], (0, 0, -1))
    with pytest.raises(leap.LeapError, match="range increment out of bounds"):
        ctx.calibrate(view, [(0.5, -0.4)], (0, 6, 4))
    with pytest.raises(leap.LeapError, match="refAnt 99 is out of bounds"):
        ctx.calibrate(view, [(0.5, -0.4)], ref_ant=99)
    with pytest.raises(leap.LeapError, match="unknown strategy"):
        ctx.calibrate(view, [(0.5, -0.4)], opts=leap.options(strategy=7))
    # the context still works afterwards
    cal, _ = ctx.calibrate(view, [(0.5, -0.47)])
    assert np.isfinite(cal).all()


def test_degeneracy_diagnostics_match_reference_rule(leap, ctx, oracle):
    """"Ad is degenerate" / "Ad1 is degenerate" (MetaData.cc:129-136): the product's
    combinatorial test against the oracle's (Ad A).isApprox(I, 0.001) on the dense
    matrices, and the messages through the log callback."""
    import planted as pl

    base = pl.inversion_ms(6, autos=True)
    a1, a2 = base["antenna1"], base["antenna2"]
    touches3 = np.nonzero(((a1 == 3) | (a2 == 3)) & (a1 != a2))[0]
    b03 = int(np.nonzero((a1 == 0) & (a2 == 3))[0][0])
    cases = {
        "clean": (pl.inversion_ms(6, autos=True), (False, False)),
        "antenna 3 isolated": (pl.inversion_ms(6, autos=True, flagged=touches3), (True, True)),
        "no (0, 3) baseline": (pl.inversion_ms(6, autos=True, flagged=(b03,)), (False, True)),
    }
    messages = []
    ctx.set_log(lambda level, msg: messages.append((level, msg)))
    try:
        for name, (ms, want) in cases.items():
            messages.clear()
            md = ctx.metadata(leap.view_from_arrays(ms), ref_ant=0)
            assert oracle.metadata(ms, ref_ant=0).degenerate() == want, name
            assert md.diagnostics() == want, name
            expect = [(1, m) for m, on in zip(("Ad is degenerate", "Ad1 is degenerate"), want) if on]
            assert messages == expect, name
    finally:
        ctx.set_log(None)


@pytest.mark.parametrize("case", ["connected", "isolated", "components"])
def test_device_pseudo_inverse_matches_host_gram_routine(leap, ctx, case):
    """The metadata's Ad and Ad1, computed from G+ = pinv(A^T A) on the device,
    against the host routine leap_phase_matrix_pinv (same Gram construction on
    the CPU) and LAPACK, including components without the constraint antenna
    and an unused antenna."""
    from test_capi import _host_phase_matrix, _incidence

    rng = np.random.default_rng(7)
    if case == "connected":
        a1, a2, fg = _incidence(30, True, rng, 0.1)
    elif case == "isolated":
        a1, a2, fg = _incidence(24, True, rng, 0.05, isolate=(5, 17))
    else:  # {0..3} with the constraint, {4..7} without it, antenna 8 unused
        a1 = np.array([0, 0, 1, 2, 4, 4, 5, 6, 0], dtype=np.int32)
        a2 = np.array([1, 2, 3, 3, 5, 6, 7, 7, 8], dtype=np.int32)
        fg = np.array([0, 0, 0, 0, 0, 0, 0, 0, 1], dtype=np.uint8)
    nbl = len(a1)
    n_ant = int(max(a1.max(), a2.max())) + 1
    ms = {"num_stations": n_ant, "total_antennas": n_ant, "num_baselines": nbl, "num_rows": nbl,
          "num_channels": 1, "num_pols": 1, "freq_start_hz": 1e8, "freq_inc_hz": 1e6, "phase_centre": (0.1, -0.5),
          "antenna1": a1, "antenna2": a2, "time": np.zeros(nbl), "flag": fg.astype(np.uint8),
          "uvw": np.zeros((nbl, 3)), "vis": np.zeros((nbl, 1, 1), dtype=np.complex128)}
    ref = 0
    md = ctx.metadata(leap.view_from_arrays(ms), ref_ant=ref)
    got = md.copy()
    A, _, Ad_host = _host_phase_matrix(leap, a1, a2, fg.astype(np.uint8), None)
    A1, _, Ad1_host = _host_phase_matrix(leap, a1, a2, fg.astype(np.uint8), ref)
    assert np.array_equal(got["A"], A) and np.array_equal(got["A1"], A1)
    err = max(np.abs(got["Ad"] - Ad_host).max(), np.abs(got["Ad1"] - Ad1_host).max())
    print(f"device vs host Gram pseudo-inverse ({case}): {err:.2e}")
    assert err < 1e-13
    assert np.abs(got["Ad"] - np.linalg.pinv(A)).max() < 1e-11
    assert np.abs(got["Ad1"] - np.linalg.pinv(A1)).max() < 1e-11

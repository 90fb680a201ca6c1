"""Pins of the CPU oracle (oracle/leap_oracle.cpp) — no GPU needed.

Each test checks the oracle against something other than its own formula:
the reference's known answers (tests/golden/reference_goldens.json), LAPACK
(numpy.linalg.pinv / svd), closed forms, or an independent vectorised numpy
restatement (tests/np_restatement.py).
"""
import json
import os

import numpy as np
import pytest

import np_restatement as npr
from oracle_bindings import OracleError

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


# ---------------------------------------------------------------------------
# DD matrix and wavelength goldens
# ---------------------------------------------------------------------------
def test_dd_matrix_golden(oracle):
    g = GOLDEN["dd_matrix"]
    dd = oracle.dd(g["phase_centre"], g["direction"])
    assert ulps(dd, g["dd"]).max() <= 4  # EXPECT_DOUBLE_EQ
    assert g["direction"][0] - g["phase_centre"][0] == g["dlm_ra"]
    assert g["direction"][1] - g["phase_centre"][1] == g["dlm_dec"]


def test_dd_matrix_is_rotation_and_identity_at_centre(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        centre = rng.uniform([-np.pi, -1.4], [np.pi, 1.4])
        d = centre + rng.uniform(-0.3, 0.3, 2)
        dd = oracle.dd(centre, d)
        assert np.abs(dd @ dd.T - np.eye(3)).max() < 1e-15
        assert abs(np.linalg.det(dd) - 1) < 1e-15
        assert np.abs(dd - npr.dd_matrix(centre, d)).max() < 1e-15
        assert np.abs(oracle.dd(centre, centre) - np.eye(3)).max() < 1e-15


def test_channel_wavelength_golden(oracle):
    g = GOLDEN["channel_wavelength"]
    assert oracle.wavelength(g["freq_start_hz"], g["freq_inc_hz"], g["channel"]) == g["wavelength_m"]
    assert oracle.wavelength(1e8, 1e6, 10) == 299792458.0 / (1e8 + 10 * 1e6)


# ---------------------------------------------------------------------------
# Pseudo-inverse
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", GOLDEN["pseudo_inverse"]["cases"], ids=lambda c: c["name"])
def test_pinv_known_answers(oracle, case):
    A = np.array(case["A"], dtype=np.float64)
    Ad = oracle.pinv(A)
    assert np.abs(Ad - np.array(case["Ad"])).max() < GOLDEN["pseudo_inverse"]["tolerance"]
    assert np.abs(A @ Ad @ A - A).max() < 1e-13
    assert np.abs(Ad - np.linalg.pinv(A)).max() < 1e-13


@pytest.mark.parametrize("shape", [(5, 3), (3, 5), (12, 12), (40, 7), (7, 40), (1, 4), (4, 1)])
def test_pinv_random_matches_lapack_and_penrose(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    A = rng.standard_normal(shape)
    Ad = oracle.pinv(A)
    assert np.abs(Ad - np.linalg.pinv(A)).max() < 1e-12
    assert np.abs(A @ Ad @ A - A).max() < 1e-12
    assert np.abs(Ad @ A @ Ad - Ad).max() < 1e-12
    assert np.abs((A @ Ad).T - A @ Ad).max() < 1e-12
    assert np.abs((Ad @ A).T - Ad @ A).max() < 1e-12
    s = oracle.singular_values(A)
    assert np.abs(s - np.linalg.svd(A, compute_uv=False)).max() < 1e-12


def test_pinv_rank_deficient_matches_lapack(oracle):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((30, 4)) @ rng.standard_normal((4, 9))  # rank 4
    assert np.abs(oracle.pinv(A) - np.linalg.pinv(A)).max() < 1e-11
    Z = np.zeros((6, 3))
    assert np.abs(oracle.pinv(Z)).max() == 0.0


def _random_incidence(rng, n_ant, autos=True, flag_frac=0.05, isolate=()):
    a1, a2 = [], []
    for i in range(n_ant):
        for j in range(i if autos else i + 1, n_ant):
            a1.append(i)
            a2.append(j)
    a1 = np.array(a1, dtype=np.int32)
    a2 = np.array(a2, dtype=np.int32)
    fg = rng.random(len(a1)) < flag_frac
    fg[:n_ant] = False
    for ant in isolate:
        fg |= (a1 == ant) | (a2 == ant)
    return a1, a2, fg.astype(np.uint8)


@pytest.mark.parametrize("ref_ant", [None, 0, 3])
def test_phase_matrix_and_pinv_vs_brute_force_and_lapack(oracle, ref_ant):
    rng = np.random.default_rng(11 if ref_ant is None else ref_ant)
    a1, a2, fg = _random_incidence(rng, 24, isolate=(7,))
    A, I = oracle.phase_matrix(a1, a2, fg, ref_ant)
    A_np, I_np = npr.phase_matrix(a1, a2, fg, ref_ant)
    assert np.array_equal(A, A_np) and np.array_equal(I, I_np)
    Ad = oracle.pinv(A)
    assert np.abs(Ad - np.linalg.pinv(A)).max() < 1e-12
    assert np.abs(A @ Ad @ A - A).max() < 1e-12
    assert np.abs(Ad[7]).max() < 1e-13  # an antenna without baselines has a zero pinv row


def test_pinv_disconnected_graph(oracle):
    # Two components (0..3 and 4..7); the constraint sits on antenna 0.
    a1 = np.array([0, 0, 1, 2, 4, 4, 5, 6], dtype=np.int32)
    a2 = np.array([1, 2, 3, 3, 5, 6, 7, 7], dtype=np.int32)
    fg = np.zeros(8, dtype=np.uint8)
    A, _ = oracle.phase_matrix(a1, a2, fg, None)
    assert np.abs(oracle.pinv(A) - np.linalg.pinv(A)).max() < 1e-12


def test_phase_matrix_reference_rules(oracle):
    # autocorrelations skipped, flags honoured, constraint row last.
    a1 = np.array([0, 0, 0, 1, 1, 2], dtype=np.int32)
    a2 = np.array([0, 1, 2, 1, 2, 2], dtype=np.int32)
    fg = np.array([0, 0, 1, 0, 0, 0], dtype=np.uint8)
    A, I = oracle.phase_matrix(a1, a2, fg, None)
    assert I.tolist() == [1, 4]
    assert A.tolist() == [[1, -1, 0], [0, 1, -1], [1, 0, 0]]
    A1, I1 = oracle.phase_matrix(a1, a2, fg, 1)
    assert I1.tolist() == [1, 4]
    assert A1.tolist() == [[1, -1, 0], [0, 1, -1], [0, 1, 0]]
    # the flag is looked up at index refAnt as if it were a baseline (PhaseMatrixFunction.cc:52)
    with pytest.raises(OracleError, match="refAnt 2 is flagged"):
        oracle.phase_matrix(a1, a2, fg, 2)


def test_phase_matrix_errors(oracle):
    # PhaseRotateTests.cc:248-274: empty input with refAnt 0 throws.
    with pytest.raises(OracleError, match="out of bounds"):
        oracle.phase_matrix(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint8), 0)
    a1 = np.array([0, 0, 1], dtype=np.int32)
    a2 = np.array([1, 2, 2], dtype=np.int32)
    with pytest.raises(OracleError, match="refAnt 3 is out of bounds"):
        oracle.phase_matrix(a1, a2, np.zeros(3, np.uint8), 3)
    with pytest.raises(OracleError, match="refAnt 1 is flagged"):
        oracle.phase_matrix(a1, a2, np.array([0, 1, 0], np.uint8), 1)


# ---------------------------------------------------------------------------
# Slices (common/Slice.cc, Range.h) — the PhaseRotateTests slices
# ---------------------------------------------------------------------------
def test_slices(oracle):
    T = 14
    assert oracle.slice_solutions(0, 1, 1, T) == 1  # Slice(0, 1)
    assert oracle.slice_solutions(0, 14, 14, T) == 1  # Slice(0, 14)
    assert oracle.slice_solutions(0, -1, -1, T) == 1  # Slice(0, -1)
    assert oracle.slice_solutions(0, 1, -1, T) == 14  # Slice(1)
    assert oracle.slice_solutions(0, 4, -1, T) == 3
    assert oracle.slice_solutions(0, 0, -1, T) == -1  # interval 0
    assert oracle.slice_solutions(5, 1, 5, T) == -1  # start >= end
    assert oracle.slice_solutions(0, 6, 4, T) == -1  # interval out of bounds
    assert oracle.slice_solutions(-1, 1, 4, T) == -1


# ---------------------------------------------------------------------------
# RotateVisibilities
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def small_ms(leap):
    s = leap.SynthMS(n_antennas=12, n_timesteps=4, n_channels=16, flag_fraction=0.1, seed=7)
    s.fill_host()
    return s


def test_rotate_at_phase_centre_is_plain_sum(oracle, small_ms):
    ms = small_ms.host_dict()
    md = oracle.metadata(ms)
    avg = md.rotate(ms["phase_centre"])
    nbl = ms["num_baselines"]
    expect = ms["vis"].reshape(-1, nbl, ms["num_channels"], ms["num_pols"]).sum(axis=(0, 2))
    assert np.abs(avg - expect).max() <= 1e-9 * np.abs(expect).max()


def test_rotate_single_source_is_phase_flat(oracle, leap):
    s = leap.SynthMS(n_antennas=10, n_timesteps=3, n_channels=12, n_sources=1, noise_sigma=0.0, seed=9)
    s.fill_host()
    ms = s.host_dict()
    src = s.sources[0]
    md = oracle.metadata(ms)
    avg = md.rotate(src[:2])
    T, C = 3, 12
    bl_phase = s.antenna_phase[s.antenna1] - s.antenna_phase[s.antenna2]
    w = np.array([1.0, 0.1 * np.exp(0.3j), 0.1 * np.exp(-0.7j), 0.9])
    expect = T * C * src[2] * np.exp(1j * bl_phase)[:, None] * w[None, :]
    assert np.abs(avg - expect).max() < 1e-9 * np.abs(expect).max()


def test_rotate_nan_skips_whole_channel(oracle, leap):
    s = leap.SynthMS(n_antennas=8, n_timesteps=3, n_channels=10, nan_fraction=0.02, seed=13)
    s.fill_host()
    ms = s.host_dict()
    assert np.isnan(ms["vis"]).any()
    d = (ms["phase_centre"][0] + 0.01, ms["phase_centre"][1] - 0.015)
    avg = oracle.metadata(ms).rotate(d)
    assert np.isfinite(avg).all()
    expect = npr.rotate_visibilities(ms, d)
    assert np.abs(avg - expect).max() < 1e-9 * np.abs(expect).max()


def test_rotate_infinities_follow_the_complex_product(oracle, leap):
    """An infinite pol is skipped exactly when its rotated product has a NaN part
    (CpuLeapCalibrator.cc:196-206 with std::complex multiplication): (inf + 0.25i) times
    exp(i 0) = 1 + 0i has the part inf * 0 = NaN, so at the phase centre it is skipped like a NaN;
    off the centre both parts are infinite and the baseline's sum turns infinite; inf + inf i
    always has one NaN part (inf - inf) and is skipped."""
    s = leap.SynthMS(n_antennas=8, n_timesteps=3, n_channels=10, seed=14)
    s.fill_host()
    nbl = s.host_dict()["num_baselines"]
    rows = {"inf": (1 * nbl + 2, 4, 0, complex(np.inf, 0.25)), "infinf": (2 * nbl + 5, 6, 1, complex(np.inf, np.inf))}

    def cube(kind, nan=False):
        v = s.vis.copy()
        r, c, p, x = rows[kind]
        v[r, c, p] = complex(np.nan, 0.0) if nan else x
        return s.host_dict(v)

    centre = s.host_dict()["phase_centre"]
    off = (centre[0] + 0.01, centre[1] - 0.015)
    ms_inf, ms_nan = cube("inf"), cube("inf", nan=True)
    # at the phase centre: skipped, the same sums as a NaN in its place
    a = oracle.metadata(ms_inf).rotate(centre)
    b = oracle.metadata(ms_nan).rotate(centre)
    assert np.array_equal(a, b) and np.isfinite(a).all()
    # off the centre: summed, baseline 2's pol 0 infinite in both parts, the rest as with a NaN
    a = oracle.metadata(ms_inf).rotate(off)
    b = oracle.metadata(ms_nan).rotate(off)
    assert np.isinf(a[2, 0].real) and np.isinf(a[2, 0].imag)
    mask = np.ones(a.shape, dtype=bool)
    mask[2, :] = False  # the other pols of that (row, channel) are summed too
    assert np.array_equal(a[mask], b[mask])
    # inf + inf i: skipped off the centre as well
    a = oracle.metadata(cube("infinf")).rotate(off)
    b = oracle.metadata(cube("infinf", nan=True)).rotate(off)
    assert np.array_equal(a, b) and np.isfinite(a).all()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_rotate_matches_numpy_restatement(oracle, small_ms, dtype):
    ms = small_ms.host_dict(small_ms.vis.astype(dtype))
    d = (ms["phase_centre"][0] - 0.02, ms["phase_centre"][1] + 0.01)
    md = oracle.metadata(ms)
    for row_start, n_rows in ((0, None), (ms["num_baselines"], 2 * ms["num_baselines"])):
        avg = md.rotate(d, row_start, n_rows)
        expect = npr.rotate_visibilities(ms, d, row_start, n_rows)
        assert np.abs(avg - expect).max() < 1e-10 * np.abs(expect).max()


# ---------------------------------------------------------------------------
# Inversion and the whole Calibrate
# ---------------------------------------------------------------------------
def test_cal1_closed_form_and_inversion(oracle, small_ms):
    ms = small_ms.host_dict()
    md = oracle.metadata(ms, ref_ant=0)
    rng = np.random.default_rng(4)
    nbl = ms["num_baselines"]
    avg = np.exp(1j * rng.uniform(-np.pi, np.pi, (nbl, ms["num_pols"]))) * rng.uniform(0.5, 2, (nbl, 1))
    phi1 = np.append(np.angle(avg[md.I1, 0]), 0.0)
    cal1 = npr.star_cal1(phi1, ms["antenna1"], ms["antenna2"], md.I1, 0, md.N)
    assert np.abs(md.Ad1 @ phi1 - cal1).max() < 1e-12
    cal = md.apply_inversion(avg)
    expect = npr.apply_inversion(avg, md.A, md.I, md.A1, md.I1)
    assert np.abs(npr.wrap(cal - expect)).max() < 1e-10


def test_metadata_matches_brute_force(oracle, small_ms):
    ms = small_ms.host_dict()
    md = oracle.metadata(ms)
    assert md.reference_antenna == ms["total_antennas"] - 1  # MetaData.cc:51
    A, I = npr.phase_matrix(ms["antenna1"], ms["antenna2"], ms["flag"], None)
    A1, I1 = npr.phase_matrix(ms["antenna1"], ms["antenna2"], ms["flag"], md.reference_antenna)
    assert np.array_equal(md.A, A) and np.array_equal(md.I, I)
    assert np.array_equal(md.A1, A1) and np.array_equal(md.I1, I1)
    assert np.abs(md.Ad - np.linalg.pinv(A)).max() < 1e-12
    assert np.abs(md.Ad1 - np.linalg.pinv(A1)).max() < 1e-12


def test_short_baselines(oracle, small_ms):
    ms = small_ms.host_dict()
    f0 = oracle.filtered_baselines(ms, 0.0)
    assert np.array_equal(f0, ms["flag"].astype(bool))
    thr = 400.0
    f = oracle.filtered_baselines(ms, thr)
    u = ms["uvw"][:, 0]
    nbl = ms["num_baselines"]
    short = np.hypot(u[:nbl], u[1:nbl + 1]) < thr  # the reference's (u_i, u_i+1) reading
    assert np.array_equal(f, ms["flag"].astype(bool) | short)


@pytest.mark.parametrize("slice_", [(0, 1, 1), (0, -1, -1), (0, 1, -1), (0, 2, -1)])
def test_calibrate_composes_rotate_and_inversion(oracle, small_ms, slice_):
    ms = small_ms.host_dict()
    c0 = ms["phase_centre"]
    dirs = np.array([[c0[0] + 0.01, c0[1] - 0.01], [c0[0] - 0.015, c0[1] + 0.005]])
    cal, epochs = oracle.calibrate(ms, dirs, slice_, ref_ant=0)
    T = ms["num_rows"] // ms["num_baselines"]
    interval = T if slice_[1] == -1 else slice_[1]
    n_sol = oracle.slice_solutions(*slice_, T)
    assert cal.shape == (n_sol, 2, 12)
    md = oracle.metadata(ms, ref_ant=0)
    nbl = ms["num_baselines"]
    for s in range(n_sol):
        assert epochs[s, 0] == ms["time"][s * interval * nbl]
        assert epochs[s, 1] == ms["time"][((s + 1) * interval - 1) * nbl]
        for d in range(2):
            avg = npr.rotate_visibilities(ms, dirs[d], s * interval * nbl, interval * nbl)
            expect = npr.apply_inversion(avg, md.A, md.I, md.A1, md.I1)
            assert np.abs(npr.wrap(cal[s, d] - expect)).max() < 1e-9


# ---------------------------------------------------------------------------
# Residual loss and gradient (the checker of the product's adjoint)
# ---------------------------------------------------------------------------
def test_oracle_gradient_matches_central_differences(oracle, leap):
    s = leap.SynthMS(n_antennas=8, n_timesteps=3, n_channels=6, seed=5)
    s.fill_host()
    ms = s.host_dict()
    c0 = ms["phase_centre"]
    dirs = np.array([[c0[0] + 0.01, c0[1] - 0.01], [c0[0] - 0.012, c0[1] + 0.004]])
    rng = np.random.default_rng(0)
    off = rng.normal(0, 1e-3, (2, 2))
    dp = rng.normal(0, 0.01, (8, 3))
    M = s.enu_to_uvw
    _, g_off, g_ant = oracle.loss_grad(ms, dirs, M, off, dp)
    h = 1e-6

    def loss(o, a):
        return oracle.loss_grad(ms, dirs, M, o, a, want_grad=False)[0]

    for i in range(2):
        for j in range(2):
            e = np.zeros_like(off)
            e[i, j] = h
            fd = (loss(off + e, dp) - loss(off - e, dp)) / (2 * h)
            assert abs(fd - g_off[i, j]) <= 1e-6 * abs(g_off).max()
    for a in (0, 3, 7):
        for k in range(3):
            e = np.zeros_like(dp)
            e[a, k] = h
            fd = (loss(off, dp + e) - loss(off, dp - e)) / (2 * h)
            assert abs(fd - g_ant[a, k]) <= 1e-5 * abs(g_ant).max()

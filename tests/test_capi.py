"""The C ABI library without a GPU: it loads, exports every declared entry
point, and its host-side model (phase matrices, pseudo-inverse, synthetic
sets, argument errors) matches the oracle and LAPACK.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import np_restatement as npr

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "leap_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(leap_\w+)\s*\(", text)) - {"leap_solution_cb"})


def test_library_exports_every_declared_symbol(leap):
    lib = leap.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(leap.EXPORTS) <= set(names)
    assert lib.leap_hip_version().startswith(b"leap-hip")


def test_context_creation_reports_status(leap):
    h = ctypes.c_void_p()
    rc = leap.lib().leap_hip_create(ctypes.byref(h), 0)
    if rc == 0:
        leap.lib().leap_hip_destroy(h)
    else:
        assert rc < 0
        assert leap.lib().leap_hip_last_error(None)
    rc = leap.lib().leap_hip_create(ctypes.byref(h), 1 << 20)
    assert rc < 0


def _host_phase_matrix(leap, a1, a2, fg, ref):
    lib = leap.lib()
    n = len(a1)
    ncols = int(max(a1.max(initial=0), a2.max(initial=0))) + 1
    A = np.zeros((n + 1) * ncols)
    I = np.zeros(max(n, 1), dtype=np.int32)
    K = ctypes.c_int32()
    N = ctypes.c_int32()
    i32 = ctypes.POINTER(ctypes.c_int32)
    rc = lib.leap_phase_matrix(a1.ctypes.data_as(i32), a2.ctypes.data_as(i32),
                               fg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n,
                               -1 if ref is None else ref, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               I.ctypes.data_as(i32), ctypes.byref(K), ctypes.byref(N))
    if rc != 0:
        raise leap.LeapError(rc, lib.leap_hip_last_error(None).decode())
    A = A.reshape(ncols, n + 1).T[: K.value, : N.value]
    Ad = np.zeros((K.value, N.value))  # N x K column-major
    rc = lib.leap_phase_matrix_pinv(a1.ctypes.data_as(i32), a2.ctypes.data_as(i32),
                                    fg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n,
                                    -1 if ref is None else ref, Ad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return A, I[: K.value - 1], Ad.T


def _incidence(n_ant, autos, rng, flag_frac, isolate=()):
    a1, a2 = np.triu_indices(n_ant, 0 if autos else 1)
    a1 = a1.astype(np.int32)
    a2 = a2.astype(np.int32)
    fg = rng.random(len(a1)) < flag_frac
    fg[:n_ant] = False
    for ant in isolate:
        fg |= (a1 == ant) | (a2 == ant)
    return a1, a2, fg.astype(np.uint8)


@pytest.mark.parametrize("n_ant,autos,ref,isolate", [
    (16, True, None, ()), (16, True, 15, ()), (40, False, 0, (5,)), (64, True, None, (3, 9)), (33, True, 20, ()),
])
def test_host_phase_matrix_and_pinv(leap, oracle, n_ant, autos, ref, isolate):
    rng = np.random.default_rng(n_ant)
    a1, a2, fg = _incidence(n_ant, autos, rng, 0.1, isolate)
    A, I, Ad = _host_phase_matrix(leap, a1, a2, fg, ref)
    A_o, I_o = oracle.phase_matrix(a1, a2, fg, ref)
    assert np.array_equal(A, A_o) and np.array_equal(I, I_o)  # bit-exact
    assert np.abs(Ad - oracle.pinv(A_o)).max() < 1e-11
    assert np.abs(Ad - np.linalg.pinv(A)).max() < 1e-11
    assert np.abs(A @ Ad @ A - A).max() < 1e-11


def test_host_pinv_disconnected_components(leap):
    # components {0,1,2,3} (with the constraint on 0), {4..7}, and antenna 8 unused.
    a1 = np.array([0, 0, 1, 2, 4, 4, 5, 6, 0], dtype=np.int32)
    a2 = np.array([1, 2, 3, 3, 5, 6, 7, 7, 8], dtype=np.int32)
    fg = np.array([0, 0, 0, 0, 0, 0, 0, 0, 1], dtype=np.uint8)
    A, _, Ad = _host_phase_matrix(leap, a1, a2, fg, None)
    assert np.abs(Ad - np.linalg.pinv(A)).max() < 1e-13
    assert not Ad[8].any()


def test_host_phase_matrix_errors(leap):
    a1 = np.array([0, 0, 1], dtype=np.int32)
    a2 = np.array([1, 2, 2], dtype=np.int32)
    with pytest.raises(leap.LeapError, match="refAnt 3 is out of bounds"):
        _host_phase_matrix(leap, a1, a2, np.zeros(3, np.uint8), 3)
    with pytest.raises(leap.LeapError, match="refAnt 1 is flagged"):
        _host_phase_matrix(leap, a1, a2, np.array([0, 1, 0], np.uint8), 1)
    e = np.zeros(0, np.int32)
    with pytest.raises(leap.LeapError, match="out of bounds"):
        _host_phase_matrix(leap, e, e, np.zeros(0, np.uint8), 0)


def test_synth_is_deterministic_and_well_formed(leap):
    a = leap.SynthMS(n_antennas=9, n_timesteps=3, n_channels=5, seed=42, nan_fraction=0.01)
    b = leap.SynthMS(n_antennas=9, n_timesteps=3, n_channels=5, seed=42, nan_fraction=0.01)
    va, vb = a.fill_host(), b.fill_host()
    assert np.array_equal(np.isnan(va), np.isnan(vb))
    assert np.array_equal(np.nan_to_num(va), np.nan_to_num(vb))
    assert a.num_baselines == 45 and a.num_rows == 135
    assert np.array_equal(a.antenna1, np.triu_indices(9)[0]) and np.array_equal(a.antenna2, np.triu_indices(9)[1])
    # autocorrelation rows have zero uvw; uvw is antisymmetric in the antenna order
    autos = a.antenna1 == a.antenna2
    assert np.abs(a.uvw[:45][autos]).max() == 0.0
    part = a.fill_host(np.complex64, row_start=45, n_rows=45)
    assert np.allclose(part, np.nan_to_num(va[45:90]).astype(np.complex64) * ~np.isnan(va[45:90]), equal_nan=True) \
        or np.array_equal(np.isnan(part), np.isnan(va[45:90]))
    noauto = leap.SynthMS(n_antennas=9, autocorrelations=0, n_timesteps=2, n_channels=3)
    assert noauto.num_baselines == 36


def test_synth_one_source_phase_matches_geometry(leap, oracle):
    """A noiseless single source seen through the reference's own rotation is phase-flat."""
    s = leap.SynthMS(n_antennas=6, n_timesteps=2, n_channels=4, n_sources=1, noise_sigma=0.0, seed=3)
    s.fill_host()
    ms = s.host_dict()
    avg = npr.rotate_visibilities(ms, s.sources[0, :2])
    bl_phase = s.antenna_phase[s.antenna1] - s.antenna_phase[s.antenna2]
    assert np.abs(npr.wrap(np.angle(avg[:, 0]) - bl_phase)).max() < 1e-9


def test_synth_enu_to_uvw_is_the_uvw_geometry(leap):
    """uvw(t, b) = M_t (enu_a2 - enu_a1): recover the baseline vectors at t = 0, then predict every t."""
    s = leap.SynthMS(n_antennas=10, n_timesteps=5, n_channels=2, seed=8)
    nbl = s.num_baselines
    uvw = s.uvw.reshape(5, nbl, 3)
    M = s.enu_to_uvw
    enu_d = np.linalg.solve(M[0], uvw[0].T).T  # [nbl, 3]
    for t in range(5):
        assert np.abs(uvw[t] - enu_d @ M[t].T).max() < 1e-9 * np.abs(uvw).max()
    assert np.abs(enu_d[:, 2]).max() < 1e-9  # a flat array


def test_view_from_arrays_checks_dtype_and_sizes(leap):
    """Arrays the library would misread or read past are refused before any call."""
    s = leap.SynthMS(n_antennas=6, n_timesteps=2, n_channels=3, seed=1)
    s.fill_host()
    ms = s.host_dict()
    leap.view_from_arrays(ms)
    for bad in (ms["vis"].real.copy(), ms["vis"].astype(np.clongdouble), ms["vis"][:-1]):
        with pytest.raises(ValueError, match="visibilities"):
            leap.view_from_arrays(ms, vis=bad)
    for name in ("antenna1", "flag", "time", "uvw"):
        short = dict(ms)
        short[name] = ms[name][:-1]
        with pytest.raises(ValueError, match=name):
            leap.view_from_arrays(short)

"""ctypes bindings for the CPU oracle (oracle/build/liboracle.so).

Test infrastructure: the oracle is the parity checker, never the product.
Matrices cross the boundary column-major (Eigen's layout); avgData is
[baseline][pol] complex128; the cube is MS-native [row][channel][pol].
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "build", "liboracle.so")

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_u8p = ctypes.POINTER(ctypes.c_uint8)


class MsView(ctypes.Structure):
    """Layout shared by oracle_ms_view (oracle/leap_oracle.h) and leap_ms_view (include/leap_hip.h)."""

    _fields_ = [
        ("num_stations", ctypes.c_uint32),
        ("total_antennas", ctypes.c_uint32),
        ("num_baselines", ctypes.c_uint32),
        ("num_rows", ctypes.c_uint32),
        ("num_channels", ctypes.c_uint32),
        ("num_pols", ctypes.c_uint32),
        ("freq_start_hz", ctypes.c_double),
        ("freq_inc_hz", ctypes.c_double),
        ("phase_centre_ra_rad", ctypes.c_double),
        ("phase_centre_dec_rad", ctypes.c_double),
        ("antenna1", _c_i32p),
        ("antenna2", _c_i32p),
        ("time", _c_dp),
        ("flag", _c_u8p),
        ("uvw", _c_dp),
        ("vis", ctypes.c_void_p),
        ("vis_dtype", ctypes.c_int32),
        ("vis_on_device", ctypes.c_int32),
    ]


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def make_view(ms):
    """Build an MsView over a host MS dict (see tests/synth_ms.py); keeps arrays alive."""
    v = MsView()
    v.num_stations = ms["num_stations"]
    v.total_antennas = ms["total_antennas"]
    v.num_baselines = ms["num_baselines"]
    v.num_rows = ms["num_rows"]
    v.num_channels = ms["num_channels"]
    v.num_pols = ms["num_pols"]
    v.freq_start_hz = ms["freq_start_hz"]
    v.freq_inc_hz = ms["freq_inc_hz"]
    v.phase_centre_ra_rad = ms["phase_centre"][0]
    v.phase_centre_dec_rad = ms["phase_centre"][1]
    v.antenna1 = _ptr(ms["antenna1"], ctypes.c_int32)
    v.antenna2 = _ptr(ms["antenna2"], ctypes.c_int32)
    v.time = _ptr(ms["time"], ctypes.c_double)
    v.flag = _ptr(ms["flag"], ctypes.c_uint8)
    v.uvw = _ptr(ms["uvw"], ctypes.c_double)
    vis = ms["vis"]
    v.vis = vis.ctypes.data
    v.vis_dtype = 1 if vis.dtype == np.complex64 else 0
    v.vis_on_device = 0
    v._keep = ms  # noqa: SLF001 — keep the arrays referenced
    return v


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


class OracleError(RuntimeError):
    pass


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            build_oracle()
        lib = ctypes.CDLL(path)
        self.lib = lib
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_channel_wavelength.restype = ctypes.c_double
        lib.oracle_channel_wavelength.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int32]
        lib.oracle_dd_matrix.argtypes = [ctypes.c_double] * 4 + [_c_dp]
        lib.oracle_dd_matrix.restype = None
        lib.oracle_metadata_create.argtypes = [ctypes.POINTER(MsView), ctypes.c_int32, ctypes.c_double,
                                               ctypes.POINTER(ctypes.c_void_p)]
        lib.oracle_metadata_destroy.argtypes = [ctypes.c_void_p]
        lib.oracle_metadata_destroy.restype = None
        lib.oracle_metadata_dims.argtypes = [ctypes.c_void_p, _c_i32p]
        lib.oracle_metadata_dims.restype = None
        lib.oracle_metadata_copy.argtypes = [ctypes.c_void_p, _c_dp, _c_i32p, _c_dp, _c_dp, _c_i32p, _c_dp]
        lib.oracle_metadata_copy.restype = None
        lib.oracle_metadata_degenerate.argtypes = [ctypes.c_void_p, _c_i32p, _c_i32p]
        lib.oracle_metadata_degenerate.restype = None
        lib.oracle_rotate_visibilities.argtypes = [ctypes.c_void_p, ctypes.POINTER(MsView), ctypes.c_uint32,
                                                   ctypes.c_uint32, _c_dp, _c_dp]
        lib.oracle_apply_inversion.argtypes = [ctypes.c_void_p, _c_dp, _c_dp]
        lib.oracle_calibrate.argtypes = [ctypes.POINTER(MsView), _c_dp, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                         _c_dp, _c_dp, _c_i32p, _c_i32p]
        lib.oracle_slice_solutions.argtypes = [ctypes.c_int32] * 4
        lib.oracle_filtered_baselines.argtypes = [ctypes.POINTER(MsView), ctypes.c_double, _c_u8p]
        lib.oracle_residual_loss_grad.argtypes = [ctypes.POINTER(MsView), _c_dp, _c_dp, ctypes.c_int32, _c_dp, _c_dp,
                                                  ctypes.c_int32, _c_dp, _c_dp, _c_dp]

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self.lib.oracle_last_error().decode())

    # -- small functions -------------------------------------------------
    def wavelength(self, f0, df, c):
        return self.lib.oracle_channel_wavelength(f0, df, c)

    def dd(self, centre, direction):
        out = np.zeros(9)
        self.lib.oracle_dd_matrix(centre[0], centre[1], direction[0], direction[1], _ptr(out, ctypes.c_double))
        return out.reshape(3, 3)

    def phase_matrix(self, a1, a2, fg, ref_ant=None):
        a1 = np.ascontiguousarray(a1, dtype=np.int32)
        a2 = np.ascontiguousarray(a2, dtype=np.int32)
        fg = np.ascontiguousarray(fg, dtype=np.uint8)
        n = len(a1)
        ncols = int(max(a1.max(initial=0), a2.max(initial=0))) + 1
        A = np.zeros(((n + 1) * ncols,), dtype=np.float64)
        I = np.zeros(max(n, 1), dtype=np.int32)
        K = ctypes.c_int32()
        N = ctypes.c_int32()
        rc = self.lib.oracle_phase_matrix(_ptr(a1, ctypes.c_int32), _ptr(a2, ctypes.c_int32),
                                          _ptr(fg, ctypes.c_uint8), n, -1 if ref_ant is None else ref_ant,
                                          _ptr(A, ctypes.c_double), _ptr(I, ctypes.c_int32),
                                          ctypes.byref(K), ctypes.byref(N))
        self._check(rc)
        A = A.reshape(ncols, n + 1).T[: K.value, : N.value]
        return np.ascontiguousarray(A), I[: K.value - 1].copy()

    def pinv(self, A):
        A = np.asfortranarray(A, dtype=np.float64)
        m, n = A.shape
        out = np.zeros((n, m), dtype=np.float64, order="F")
        self._check(self.lib.oracle_pinv(A.ctypes.data_as(_c_dp), m, n, out.ctypes.data_as(_c_dp)))
        return np.ascontiguousarray(out)

    def singular_values(self, A):
        A = np.asfortranarray(A, dtype=np.float64)
        m, n = A.shape
        out = np.zeros(min(m, n))
        self._check(self.lib.oracle_singular_values(A.ctypes.data_as(_c_dp), m, n, out.ctypes.data_as(_c_dp)))
        return out

    def slice_solutions(self, start, interval, end, timesteps):
        return self.lib.oracle_slice_solutions(start, interval, end, timesteps)

    def filtered_baselines(self, ms, threshold):
        v = make_view(ms)
        out = np.zeros(ms["num_baselines"], dtype=np.uint8)
        self._check(self.lib.oracle_filtered_baselines(ctypes.byref(v), threshold, _ptr(out, ctypes.c_uint8)))
        return out.astype(bool)

    def loss_grad(self, ms, directions, enu_to_uvw, offsets=None, antenna_delta=None, ref_ant=None, want_grad=True):
        v = make_view(ms)
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        D = len(dirs)
        n_ant = int(max(ms["antenna1"].max(), ms["antenna2"].max())) + 1
        off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.float64).reshape(D, 2)
        dp = None if antenna_delta is None else np.ascontiguousarray(antenna_delta, dtype=np.float64).reshape(n_ant, 3)
        M = np.ascontiguousarray(enu_to_uvw, dtype=np.float64)
        loss = ctypes.c_double()
        g_off = np.zeros((D, 2))
        g_ant = np.zeros((n_ant, 3))
        null = ctypes.POINTER(ctypes.c_double)()
        self._check(self.lib.oracle_residual_loss_grad(
            ctypes.byref(v), dirs.ctypes.data_as(_c_dp), null if off is None else off.ctypes.data_as(_c_dp), D,
            null if dp is None else dp.ctypes.data_as(_c_dp), M.ctypes.data_as(_c_dp), -1 if ref_ant is None else ref_ant,
            ctypes.byref(loss), g_off.ctypes.data_as(_c_dp) if want_grad else null,
            g_ant.ctypes.data_as(_c_dp) if want_grad else null))
        return loss.value, g_off, g_ant

    # -- metadata / calibration -------------------------------------------
    def metadata(self, ms, ref_ant=None, threshold=0.0):
        return OracleMetaData(self, ms, ref_ant, threshold)

    def calibrate(self, ms, directions, slice_=(0, -1, -1), threshold=0.0, ref_ant=None):
        v = make_view(ms)
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        nsol = ctypes.c_int32()
        nant = ctypes.c_int32()
        null = ctypes.POINTER(ctypes.c_double)()
        ra = -1 if ref_ant is None else ref_ant
        self._check(self.lib.oracle_calibrate(ctypes.byref(v), dirs.ctypes.data_as(_c_dp), len(dirs), *slice_,
                                              threshold, ra, null, null, ctypes.byref(nsol), ctypes.byref(nant)))
        cal = np.zeros((nsol.value, len(dirs), nant.value))
        epochs = np.zeros((nsol.value, 2))
        self._check(self.lib.oracle_calibrate(ctypes.byref(v), dirs.ctypes.data_as(_c_dp), len(dirs), *slice_,
                                              threshold, ra, cal.ctypes.data_as(_c_dp),
                                              epochs.ctypes.data_as(_c_dp), ctypes.byref(nsol),
                                              ctypes.byref(nant)))
        return cal, epochs


class OracleMetaData:
    def __init__(self, oracle, ms, ref_ant, threshold):
        self.o = oracle
        self.ms = ms
        self.view = make_view(ms)
        h = ctypes.c_void_p()
        ra = -1 if ref_ant is None else ref_ant
        oracle._check(oracle.lib.oracle_metadata_create(ctypes.byref(self.view), ra, threshold, ctypes.byref(h)))
        self.h = h
        dims = np.zeros(5, dtype=np.int32)
        oracle.lib.oracle_metadata_dims(h, _ptr(dims, ctypes.c_int32))
        self.K, self.N, self.K1, self.nbl, self.reference_antenna = (int(x) for x in dims)
        A = np.zeros((self.N, self.K))
        Ad = np.zeros((self.K, self.N))
        A1 = np.zeros((self.N, self.K1))
        Ad1 = np.zeros((self.K1, self.N))
        I = np.zeros(max(self.K - 1, 1), dtype=np.int32)
        I1 = np.zeros(max(self.K1 - 1, 1), dtype=np.int32)
        oracle.lib.oracle_metadata_copy(h, _ptr(A, ctypes.c_double), _ptr(I, ctypes.c_int32),
                                        _ptr(Ad, ctypes.c_double), _ptr(A1, ctypes.c_double),
                                        _ptr(I1, ctypes.c_int32), _ptr(Ad1, ctypes.c_double))
        # column-major buffers -> numpy (rows, cols)
        self.A = A.T.copy()
        self.Ad = Ad.T.copy()
        self.A1 = A1.T.copy()
        self.Ad1 = Ad1.T.copy()
        self.I = I[: self.K - 1].copy()
        self.I1 = I1[: self.K1 - 1].copy()

    def rotate(self, direction, row_start=0, n_rows=None):
        if n_rows is None:
            n_rows = self.ms["num_rows"] - row_start
        out = np.zeros((self.nbl, self.ms["num_pols"]), dtype=np.complex128)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        self.o._check(self.o.lib.oracle_rotate_visibilities(self.h, ctypes.byref(self.view), row_start, n_rows,
                                                            _ptr(d, ctypes.c_double),
                                                            out.ctypes.data_as(_c_dp)))
        return out

    def degenerate(self):
        """(Ad degenerate, Ad1 degenerate) by the reference's isApprox test on the dense matrices."""
        a, a1 = ctypes.c_int32(), ctypes.c_int32()
        self.o.lib.oracle_metadata_degenerate(self.h, ctypes.byref(a), ctypes.byref(a1))
        return bool(a.value), bool(a1.value)

    def apply_inversion(self, avg):
        avg = np.ascontiguousarray(avg, dtype=np.complex128)
        cal = np.zeros(self.N)
        self.o._check(self.o.lib.oracle_apply_inversion(self.h, avg.ctypes.data_as(_c_dp), _ptr(cal, ctypes.c_double)))
        return cal

    def __del__(self):
        try:
            self.o.lib.oracle_metadata_destroy(self.h)
        except Exception:  # noqa: BLE001 — interpreter shutdown
            pass

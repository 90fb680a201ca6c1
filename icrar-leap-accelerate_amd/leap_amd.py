"""Python binding of the gfx950 LEAP calibrator's C ABI (include/leap_hip.h).

A thin ctypes layer used by the test-suite, the smoke check and the
benchmark.  The shared library is built in-tree (build/libleap_hip.so); if
it is missing this module raises instead of falling back to anything else.
"""
import ctypes
import json
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "build", "libleap_hip.so")

LEAP_OK = 0
LEAP_VIS_C128 = 0
LEAP_VIS_C64 = 1
LEAP_PRECISION_FP64 = 0
LEAP_PRECISION_MIXED = 1
LEAP_CHANNEL_SUM_AUTO = 0
LEAP_CHANNEL_SUM_HORNER = 1
LEAP_CHANNEL_SUM_NUFFT = 2
LEAP_CHANNEL_SUM_NUFFT_FP32 = 3
LEAP_CHANNEL_SUM_NUFFT_W16 = 4
LEAP_STRATEGY_PER_DIRECTION = 0
LEAP_STRATEGY_BATCHED = 1
LEAP_ADJOINT_AUTO = 0
LEAP_ADJOINT_RECOMPUTE = 1
LEAP_ADJOINT_ROW_SUMS = 2

_STATUS = {
    -1: "invalid argument",
    -2: "range",
    -3: "hip",
    -4: "out of memory",
    -5: "not implemented",
    -6: "callback",
    -7: "internal",
}

_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class LeapError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status
        self.message = message


class MsView(ctypes.Structure):
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
        ("antenna1", _i32p),
        ("antenna2", _i32p),
        ("time", _dp),
        ("flag", _u8p),
        ("uvw", _dp),
        ("vis", ctypes.c_void_p),
        ("vis_dtype", ctypes.c_int32),
        ("vis_on_device", ctypes.c_int32),
        ("uvw_device", _dp),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("strategy", ctypes.c_int32),
        ("direction_block", ctypes.c_int32),
        ("all_pols", ctypes.c_int32),
        ("fp32_blocks", ctypes.c_int32),
        ("lane_directions", ctypes.c_int32),
        ("workgroups_per_cu", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("direct_cube", ctypes.c_int32),
        ("pack_timesteps", ctypes.c_int32),
        ("channel_sum", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
    ]


def options(precision=LEAP_PRECISION_FP64, strategy=LEAP_STRATEGY_PER_DIRECTION, direction_block=0, all_pols=0,
            fp32_blocks=0, lane_directions=0, workgroups_per_cu=0, graph=0, direct_cube=0, pack_timesteps=0,
            channel_sum=0):
    o = Options()
    o.precision = precision
    o.strategy = strategy
    o.direction_block = direction_block
    o.all_pols = all_pols
    o.fp32_blocks = fp32_blocks
    o.lane_directions = lane_directions
    o.workgroups_per_cu = workgroups_per_cu
    o.graph = graph
    o.direct_cube = direct_cube
    o.pack_timesteps = pack_timesteps
    o.channel_sum = channel_sum
    return o


class Solution(ctypes.Structure):
    _fields_ = [
        ("index", ctypes.c_uint32),
        ("start_epoch", ctypes.c_double),
        ("end_epoch", ctypes.c_double),
        ("n_directions", ctypes.c_uint32),
        ("directions", _dp),
        ("n_antennas", ctypes.c_uint32),
        ("calibrations", _dp),
    ]


SOLUTION_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(Solution))
LOG_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p)


class BatchSet(ctypes.Structure):
    _fields_ = [
        ("metadata", ctypes.c_void_p),
        ("ms", ctypes.POINTER(MsView)),
        ("directions", _dp),
        ("n_directions", ctypes.c_uint32),
        ("slice_start", ctypes.c_int32),
        ("slice_interval", ctypes.c_int32),
        ("slice_end", ctypes.c_int32),
    ]


BATCH_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Solution))


class SynthConfig(ctypes.Structure):
    _fields_ = [
        ("n_antennas", ctypes.c_uint32),
        ("layout", ctypes.c_int32),
        ("array_radius_m", ctypes.c_double),
        ("autocorrelations", ctypes.c_int32),
        ("n_timesteps", ctypes.c_uint32),
        ("integration_s", ctypes.c_double),
        ("n_channels", ctypes.c_uint32),
        ("freq_start_hz", ctypes.c_double),
        ("freq_inc_hz", ctypes.c_double),
        ("n_pols", ctypes.c_uint32),
        ("phase_centre_ra_rad", ctypes.c_double),
        ("phase_centre_dec_rad", ctypes.c_double),
        ("latitude_rad", ctypes.c_double),
        ("n_sources", ctypes.c_uint32),
        ("source_flux_min", ctypes.c_double),
        ("source_flux_max", ctypes.c_double),
        ("source_offset_max_rad", ctypes.c_double),
        ("noise_sigma", ctypes.c_double),
        ("flag_fraction", ctypes.c_double),
        ("nan_fraction", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


# Exported symbols of include/leap_hip.h (checked by the CPU test-suite).
EXPORTS = [
    "leap_hip_create", "leap_hip_destroy", "leap_hip_last_error", "leap_hip_version",
    "leap_hip_calibrate", "leap_hip_calibrate_md", "leap_hip_calibrate_md_noalloc",
    "leap_hip_metadata_create", "leap_hip_metadata_destroy", "leap_hip_metadata_dims", "leap_hip_metadata_copy",
    "leap_hip_metadata_diagnostics", "leap_hip_set_log_callback",
    "leap_hip_rotate_visibilities", "leap_hip_apply_inversion", "leap_hip_dd_matrix",
    "leap_synth_create", "leap_synth_destroy", "leap_synth_view", "leap_synth_sky", "leap_synth_fill_vis_host",
    "leap_hip_synth_fill_vis_device",
    "leap_hip_malloc", "leap_hip_free", "leap_hip_memcpy_d2h", "leap_hip_memcpy_h2d",
    "leap_hip_last_kernel_times", "leap_hip_last_direction_block", "leap_hip_last_route",
    "leap_hip_last_workgroups_per_cu",
    "leap_hip_last_graph", "leap_hip_last_pack_times", "leap_hip_last_staged_bytes", "leap_hip_last_loss_grad_times",
    "leap_hip_set_adjoint_mode",
    "leap_hip_tuning_save", "leap_hip_tuning_load", "leap_phase_matrix", "leap_phase_matrix_pinv",
    "leap_hip_residual_loss_grad", "leap_synth_enu_to_uvw", "leap_hip_calibrate_batch", "leap_multi_schedule",
    "leap_multi_item_cost", "leap_phasor64", "leap_nufft_plan", "leap_nufft_plan32", "leap_nufft_plan_w16",
]


def build():
    subprocess.run(["make", "-s", "-C", PKG_DIR, "-j8"], check=True)


_lib = None


def lib():
    """Load build/libleap_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C {PKG_DIR}")
    L = ctypes.CDLL(LIB_PATH)
    L.leap_hip_last_error.restype = ctypes.c_char_p
    L.leap_hip_last_error.argtypes = [ctypes.c_void_p]
    L.leap_hip_version.restype = ctypes.c_char_p
    L.leap_hip_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32]
    L.leap_hip_destroy.argtypes = [ctypes.c_void_p]
    L.leap_hip_destroy.restype = None
    L.leap_hip_calibrate.argtypes = [ctypes.c_void_p, ctypes.POINTER(MsView), _dp, ctypes.c_uint32,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Options), SOLUTION_CB,
                                     ctypes.c_void_p]
    L.leap_hip_calibrate_md.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MsView), _dp,
                                        ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(Options), SOLUTION_CB, ctypes.c_void_p]
    L.leap_hip_calibrate_md_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MsView), _dp,
                                                ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_void_p, _dp, _i32p]
    L.leap_hip_calibrate_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchSet), ctypes.c_uint32,
                                           ctypes.POINTER(Options), BATCH_CB, ctypes.c_void_p]
    _u32p = ctypes.POINTER(ctypes.c_uint32)
    L.leap_multi_schedule.argtypes = [_u32p, _u32p, _u32p, _u32p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                      _u32p, _u32p, _u32p, _dp]
    L.leap_multi_item_cost.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32]
    L.leap_phasor64.argtypes = [_dp, ctypes.c_uint32, _dp, _dp]
    _u32p = ctypes.POINTER(ctypes.c_uint32)
    L.leap_nufft_plan.argtypes = [ctypes.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp]
    L.leap_nufft_plan32.argtypes = [ctypes.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp]
    L.leap_nufft_plan_w16.argtypes = [ctypes.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp]
    L.leap_multi_item_cost.restype = ctypes.c_double
    L.leap_hip_metadata_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(MsView), ctypes.c_int32,
                                           ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    L.leap_hip_metadata_destroy.argtypes = [ctypes.c_void_p]
    L.leap_hip_metadata_destroy.restype = None
    L.leap_hip_metadata_dims.argtypes = [ctypes.c_void_p, _i32p]
    L.leap_hip_metadata_dims.restype = None
    L.leap_hip_metadata_copy.argtypes = [ctypes.c_void_p, _dp, _i32p, _dp, _dp, _i32p, _dp]
    L.leap_hip_metadata_diagnostics.argtypes = [ctypes.c_void_p, _i32p, _i32p]
    L.leap_hip_set_log_callback.argtypes = [ctypes.c_void_p, LOG_CB, ctypes.c_void_p]
    L.leap_hip_rotate_visibilities.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MsView),
                                               ctypes.c_uint32, ctypes.c_uint32, _dp, ctypes.c_uint32,
                                               ctypes.POINTER(Options), _dp]
    L.leap_hip_apply_inversion.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, ctypes.c_uint32,
                                           ctypes.POINTER(Options), _dp]
    L.leap_hip_dd_matrix.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_uint32, _dp]
    L.leap_synth_create.argtypes = [ctypes.POINTER(SynthConfig), ctypes.POINTER(ctypes.c_void_p)]
    L.leap_synth_destroy.argtypes = [ctypes.c_void_p]
    L.leap_synth_destroy.restype = None
    L.leap_synth_view.argtypes = [ctypes.c_void_p, ctypes.POINTER(MsView)]
    L.leap_synth_view.restype = None
    L.leap_synth_sky.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.leap_synth_sky.restype = None
    L.leap_synth_fill_vis_host.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                           ctypes.c_void_p]
    L.leap_hip_synth_fill_vis_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_void_p, _dp]
    L.leap_hip_malloc.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
    L.leap_hip_free.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.leap_hip_memcpy_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.leap_hip_memcpy_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.leap_hip_last_kernel_times.argtypes = [ctypes.c_void_p, _dp, _i32p, _dp]
    L.leap_hip_last_direction_block.argtypes = [ctypes.c_void_p]
    L.leap_hip_last_route.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.leap_hip_last_workgroups_per_cu.argtypes = [ctypes.c_void_p]
    L.leap_hip_last_loss_grad_times.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.leap_hip_set_adjoint_mode.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.leap_hip_last_graph.argtypes = [ctypes.c_void_p]
    L.leap_hip_last_pack_times.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int32)]
    L.leap_hip_last_staged_bytes.argtypes = [ctypes.c_void_p]
    L.leap_hip_tuning_save.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.c_uint64)]
    L.leap_hip_tuning_load.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L.leap_hip_last_staged_bytes.restype = ctypes.c_uint64
    L.leap_hip_residual_loss_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MsView), _dp, _dp,
                                              ctypes.c_uint32, _dp, _dp, _dp, _dp, _dp]
    L.leap_synth_enu_to_uvw.argtypes = [ctypes.c_void_p, _dp]
    L.leap_synth_enu_to_uvw.restype = None
    _lib = L
    return L


def _check(rc, ctx=None):
    if rc != LEAP_OK:
        msg = lib().leap_hip_last_error(ctx).decode()
        raise LeapError(rc, msg)


def _dptr(a):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
# Synthetic measurement sets
# ---------------------------------------------------------------------------
def synth_config(**kw):
    """Defaults follow the 128-antenna MWA-like set (1.5 km disk, 14 x 2 s, 48 channels)."""
    c = SynthConfig()
    d = dict(n_antennas=128, layout=0, array_radius_m=750.0, autocorrelations=1, n_timesteps=14, integration_s=2.0,
             n_channels=48, freq_start_hz=169.6e6, freq_inc_hz=40e3, n_pols=4, phase_centre_ra_rad=0.5,
             phase_centre_dec_rad=np.deg2rad(-27.0), latitude_rad=np.deg2rad(-26.7033), n_sources=5,
             source_flux_min=1.0, source_flux_max=10.0, source_offset_max_rad=0.05, noise_sigma=0.1,
             flag_fraction=0.01, nan_fraction=0.0, seed=20201)
    d.update(kw)
    for k, v in d.items():
        setattr(c, k, v)
    return c


class SynthMS:
    """A synthetic measurement set: host columns, plus host or device visibilities."""

    def __init__(self, **kw):
        self.cfg = synth_config(**kw)
        h = ctypes.c_void_p()
        rc = lib().leap_synth_create(ctypes.byref(self.cfg), ctypes.byref(h))
        _check(rc)
        self.h = h
        v = MsView()
        lib().leap_synth_view(h, ctypes.byref(v))
        self.view_template = v
        self.num_rows = v.num_rows
        self.num_baselines = v.num_baselines
        self.num_channels = v.num_channels
        self.num_pols = v.num_pols
        n = self.cfg.n_antennas
        nb = v.num_baselines
        self.antenna1 = np.ctypeslib.as_array(v.antenna1, shape=(nb,)).copy()
        self.antenna2 = np.ctypeslib.as_array(v.antenna2, shape=(nb,)).copy()
        self.flag = np.ctypeslib.as_array(v.flag, shape=(nb,)).copy()
        self.uvw = np.ctypeslib.as_array(v.uvw, shape=(v.num_rows, 3)).copy()
        self.time = np.ctypeslib.as_array(v.time, shape=(v.num_rows,)).copy()
        self.sources = np.zeros((self.cfg.n_sources, 3))
        self.antenna_phase = np.zeros(n)
        lib().leap_synth_sky(h, _dptr(self.sources) if self.cfg.n_sources else None, _dptr(self.antenna_phase))
        self.enu_to_uvw = np.zeros((self.cfg.n_timesteps, 3, 3))
        lib().leap_synth_enu_to_uvw(h, _dptr(self.enu_to_uvw))
        self.vis = None

    def fill_host(self, dtype=np.complex128, row_start=0, n_rows=None):
        if n_rows is None:
            n_rows = self.num_rows - row_start
        out = np.zeros((n_rows, self.num_channels, self.num_pols), dtype=dtype)
        vd = LEAP_VIS_C64 if dtype == np.complex64 else LEAP_VIS_C128
        _check(lib().leap_synth_fill_vis_host(self.h, row_start, n_rows, vd, out.ctypes.data))
        if row_start == 0 and n_rows == self.num_rows:
            self.vis = out
        return out

    def host_dict(self, vis=None):
        """Dictionary form used by the oracle bindings in tests/."""
        vis = self.vis if vis is None else vis
        v = self.view_template
        return {
            "num_stations": v.num_stations, "total_antennas": v.total_antennas, "num_baselines": v.num_baselines,
            "num_rows": v.num_rows, "num_channels": v.num_channels, "num_pols": v.num_pols,
            "freq_start_hz": v.freq_start_hz, "freq_inc_hz": v.freq_inc_hz,
            "phase_centre": (v.phase_centre_ra_rad, v.phase_centre_dec_rad),
            "antenna1": self.antenna1, "antenna2": self.antenna2, "time": self.time, "flag": self.flag,
            "uvw": self.uvw, "vis": vis,
        }

    def __del__(self):
        try:
            lib().leap_synth_destroy(self.h)
        except Exception:  # noqa: BLE001
            pass


def view_from_arrays(ms, vis=None, vis_device=None, vis_dtype=None, uvw_device=None):
    """Build an MsView from a host MS dict; vis may be host numpy or a device pointer."""
    v = MsView()
    v.num_stations = ms["num_stations"]
    v.total_antennas = ms["total_antennas"]
    v.num_baselines = ms["num_baselines"]
    v.num_rows = ms["num_rows"]
    v.num_channels = ms["num_channels"]
    v.num_pols = ms["num_pols"]
    v.freq_start_hz = ms["freq_start_hz"]
    v.freq_inc_hz = ms["freq_inc_hz"]
    v.phase_centre_ra_rad, v.phase_centre_dec_rad = ms["phase_centre"]
    keep = {}
    nbl, rows = int(ms["num_baselines"]), int(ms["num_rows"])
    need = {"antenna1": nbl, "antenna2": nbl, "flag": nbl, "time": rows, "uvw": 3 * rows}
    for name, dt, ct in (("antenna1", np.int32, ctypes.c_int32), ("antenna2", np.int32, ctypes.c_int32),
                         ("time", np.float64, ctypes.c_double), ("flag", np.uint8, ctypes.c_uint8),
                         ("uvw", np.float64, ctypes.c_double)):
        a = np.ascontiguousarray(ms[name], dtype=dt)
        if a.size < need[name]:
            raise ValueError(f"{name} holds {a.size} values, the view reads {need[name]}")
        keep[name] = a
        setattr(v, name, a.ctypes.data_as(ctypes.POINTER(ct)))
    if vis_device is not None:
        v.vis = vis_device
        v.vis_on_device = 1
        v.vis_dtype = vis_dtype
    else:
        vis = ms["vis"] if vis is None else vis
        vis = np.ascontiguousarray(vis)
        if vis.dtype not in (np.complex64, np.complex128):
            raise ValueError(f"visibilities must be complex64 or complex128, not {vis.dtype}")
        want = int(ms["num_rows"]) * int(ms["num_channels"]) * int(ms["num_pols"])
        if vis.size != want:
            raise ValueError(f"visibilities hold {vis.size} values, the view needs num_rows x num_channels x "
                             f"num_pols = {want}")
        keep["vis"] = vis
        v.vis = vis.ctypes.data
        v.vis_on_device = 0
        v.vis_dtype = LEAP_VIS_C64 if vis.dtype == np.complex64 else LEAP_VIS_C128
    if uvw_device is not None:
        v.uvw_device = ctypes.cast(uvw_device, _dp)
    v._keep = keep  # noqa: SLF001
    return v


def multi_schedule(shapes, tile, n_blocks):
    """Host schedule of the multi-set kernel for shapes [(nbl, T, C, D)]:
    (offsets [n_blocks + 1], item_set, item_local, makespan)."""
    sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.uint32).reshape(-1, 4))
    cols = [np.ascontiguousarray(sh[:, i]) for i in range(4)]
    db = tile // 10
    n_items = int(sum(int(n) * ((int(d) + db - 1) // db) for n, _, _, d in sh if d > 0)) if db else 0
    off = np.zeros(n_blocks + 1, dtype=np.uint32)
    iset = np.zeros(max(1, n_items), dtype=np.uint32)
    iloc = np.zeros(max(1, n_items), dtype=np.uint32)
    mk = ctypes.c_double()
    u = ctypes.POINTER(ctypes.c_uint32)
    _check(lib().leap_multi_schedule(*[c.ctypes.data_as(u) for c in cols], len(sh), tile, n_blocks,
                                     off.ctypes.data_as(u), iset.ctypes.data_as(u), iloc.ctypes.data_as(u),
                                     ctypes.byref(mk)))
    return off, iset[:n_items], iloc[:n_items], mk.value


def write_raw_ms(path, *, antenna1, antenna2, time, flag_c0p0, uvw, data, antenna_table_rows, ref_frequency_hz,
                 chan_width_hz, phase_dir_rad, data_column="DATA"):
    """Writes a raw measurement-set directory (the format RawMeasurementSet.h reads): header.json and
    little-endian column files.  data is [rows][channels][pols] complex64 or complex128; every other
    column has one entry per row (uvw three)."""
    data = np.asarray(data)
    if data.dtype not in (np.complex64, np.complex128) or data.ndim != 3:
        raise ValueError("data must be a [rows][channels][pols] complex64 or complex128 array")
    R, C, P = data.shape
    cols = {"ANTENNA1.i32": np.asarray(antenna1, dtype="<i4"), "ANTENNA2.i32": np.asarray(antenna2, dtype="<i4"),
            "TIME.f64": np.asarray(time, dtype="<f8"), "FLAG_C0P0.u8": np.asarray(flag_c0p0, dtype=np.uint8),
            "UVW.f64": np.asarray(uvw, dtype="<f8").reshape(-1, 3)}
    for name, col in cols.items():
        if len(col) != R:
            raise ValueError(f"{name} has {len(col)} rows, DATA has {R}")
    os.makedirs(path, exist_ok=True)
    header = {"format": "leap-raw-ms", "version": 1, "rows": int(R), "channels": int(C), "pols": int(P),
              "antenna_table_rows": int(antenna_table_rows), "ref_frequency_hz": float(ref_frequency_hz),
              "chan_width_hz": float(chan_width_hz),
              "phase_dir_rad": [float(phase_dir_rad[0]), float(phase_dir_rad[1])],
              "data_column": data_column, "data_type": "complex64" if data.dtype == np.complex64 else "complex128"}
    for name, col in cols.items():
        np.ascontiguousarray(col).tofile(os.path.join(path, name))
    np.ascontiguousarray(data).tofile(os.path.join(path, "DATA.c64" if data.dtype == np.complex64 else "DATA.c128"))
    with open(os.path.join(path, "header.json"), "w") as f:
        json.dump(header, f)


def write_raw_ms_from_dict(path, ms, antenna_table_rows=None):
    """write_raw_ms from the dictionary form of a measurement set (SynthMS.host_dict): the first-epoch
    flags are repeated over every epoch."""
    nbl = ms["num_baselines"]
    rows = ms["num_rows"]
    a1 = np.tile(np.asarray(ms["antenna1"]), rows // nbl)
    a2 = np.tile(np.asarray(ms["antenna2"]), rows // nbl)
    flag = np.tile(np.asarray(ms["flag"], dtype=np.uint8), rows // nbl)
    write_raw_ms(path, antenna1=a1, antenna2=a2, time=ms["time"], flag_c0p0=flag, uvw=ms["uvw"], data=ms["vis"],
                 antenna_table_rows=ms["total_antennas"] if antenna_table_rows is None else antenna_table_rows,
                 ref_frequency_hz=ms["freq_start_hz"], chan_width_hz=ms["freq_inc_hz"],
                 phase_dir_rad=ms["phase_centre"])


def phasor64(theta):
    """(cos, sin) of theta as the batched kernels compute them (host build of the same routine)."""
    th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    c = np.empty_like(th)
    s = np.empty_like(th)
    _check(lib().leap_phasor64(_dptr(th), len(th), _dptr(c), _dptr(s)))
    return c, s


def multi_item_cost(T, C, tile):
    return lib().leap_multi_item_cost(T, C, tile)


def nufft_plan(C, fp32=False, w16=False):
    """The NUFFT plan of C channels (fp32: the FP32 transform of LEAP_CHANNEL_SUM_NUFFT_FP32;
    w16: the width-16 kernel of LEAP_CHANNEL_SUM_NUFFT_W16; else the FP64 transform of
    LEAP_CHANNEL_SUM_NUFFT): dict of N, W, h, beta, coef (W/2 x W weight-polynomial
    coefficients in s) and scale (C precorrection factors)."""
    if fp32 and w16:
        raise ValueError("the width-16 plan has an FP64 transform")
    N, W, h = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    beta = ctypes.c_double()
    width = 16 if w16 else 12
    coef = np.zeros((width // 2) * width)
    scale = np.zeros(max(int(C), 1))
    fn = lib().leap_nufft_plan_w16 if w16 else lib().leap_nufft_plan32 if fp32 else lib().leap_nufft_plan
    _check(fn(C, ctypes.byref(N), ctypes.byref(W), ctypes.byref(h), ctypes.byref(beta), _dptr(coef), _dptr(scale)))
    w = W.value
    return {"N": N.value, "W": w, "h": h.value, "beta": beta.value,
            "coef": coef.reshape(width // 2, width)[: w // 2, :w].copy(), "scale": scale[: int(C)].copy()}


# ---------------------------------------------------------------------------
# Calibrator context
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().leap_hip_create(ctypes.byref(h), device))
        self.h = h

    def close(self):
        if self.h:
            lib().leap_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def _chk(self, rc):
        _check(rc, self.h)

    # device memory --------------------------------------------------------
    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self._chk(lib().leap_hip_malloc(self.h, nbytes, ctypes.byref(p)))
        return p.value

    def free(self, ptr):
        self._chk(lib().leap_hip_free(self.h, ptr))

    def d2h(self, dst_array, src_ptr, nbytes=None):
        nbytes = dst_array.nbytes if nbytes is None else nbytes
        self._chk(lib().leap_hip_memcpy_d2h(self.h, dst_array.ctypes.data, src_ptr, nbytes))

    def h2d(self, dst_ptr, src_array):
        self._chk(lib().leap_hip_memcpy_h2d(self.h, dst_ptr, src_array.ctypes.data, src_array.nbytes))

    def synth_fill_device(self, synth, vis_dtype, vis_ptr, uvw_ptr):
        self._chk(lib().leap_hip_synth_fill_vis_device(self.h, synth.h, vis_dtype, vis_ptr,
                                                       ctypes.cast(uvw_ptr, _dp) if uvw_ptr else None))

    def set_log(self, fn):
        """fn(level, message) receives the context's diagnostics (None drops them)."""
        self._log_cb = LOG_CB(lambda _u, level, msg: fn(level, msg.decode())) if fn else LOG_CB()
        self._chk(lib().leap_hip_set_log_callback(self.h, self._log_cb, None))

    # metadata / kernels ---------------------------------------------------
    def metadata(self, view, ref_ant=None, threshold=0.0, use_cache=False):
        return Metadata(self, view, ref_ant, threshold, use_cache)

    def dd(self, centre, directions):
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((len(dirs), 3, 3))
        self._chk(lib().leap_hip_dd_matrix(self.h, centre[0], centre[1], _dptr(dirs), len(dirs), _dptr(out)))
        return out

    def calibrate(self, view, directions, slice_=(0, -1, -1), threshold=0.0, ref_ant=None, use_cache=False,
                  opts=None, md=None):
        """Returns (cal [solutions][directions][N], epochs [solutions][2])."""
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        cals, epochs = [], []

        def cb(_user, sol_p):
            s = sol_p.contents
            n = s.n_directions * s.n_antennas
            cals.append(np.ctypeslib.as_array(s.calibrations, shape=(n,)).copy().reshape(s.n_directions,
                                                                                          s.n_antennas)
                        if n else np.zeros((s.n_directions, s.n_antennas)))
            epochs.append((s.start_epoch, s.end_epoch))
            return 0

        cbf = SOLUTION_CB(cb)
        o = opts if opts is not None else options()
        if md is None:
            rc = lib().leap_hip_calibrate(self.h, ctypes.byref(view), _dptr(dirs), len(dirs), *slice_, threshold,
                                          -1 if ref_ant is None else ref_ant, 1 if use_cache else 0,
                                          ctypes.byref(o), cbf, None)
        else:
            rc = lib().leap_hip_calibrate_md(self.h, md.h, ctypes.byref(view), _dptr(dirs), len(dirs), *slice_,
                                             ctypes.byref(o), cbf, None)
        self._chk(rc)
        return np.array(cals), np.array(epochs)

    def calibrate_noalloc(self, view, directions, md, cal_device, slice_=(0, -1, -1)):
        """leap_hip_calibrate_md_noalloc: cal into the device buffer cal_device
        ([solutions][directions][N] doubles, the caller's); returns epochs [solutions][2]."""
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        n_sol = ctypes.c_int32()
        self._chk(lib().leap_hip_calibrate_md_noalloc(self.h, md.h, ctypes.byref(view), _dptr(dirs), len(dirs),
                                                      *slice_, None, None, ctypes.byref(n_sol)))
        epochs = np.zeros((n_sol.value, 2))
        self._chk(lib().leap_hip_calibrate_md_noalloc(self.h, md.h, ctypes.byref(view), _dptr(dirs), len(dirs),
                                                      *slice_, cal_device, _dptr(epochs), ctypes.byref(n_sol)))
        return epochs

    def calibrate_batch(self, sets, opts=None):
        """sets: [(md, view, directions, slice_)]; returns per set (cal [solutions][D][N], epochs [solutions][2])."""
        keep = []
        arr = (BatchSet * max(1, len(sets)))()
        for i, (md, view, directions, slice_) in enumerate(sets):
            dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
            keep.append(dirs)
            arr[i].metadata = md.h
            arr[i].ms = ctypes.pointer(view)
            arr[i].directions = _dptr(dirs)
            arr[i].n_directions = len(dirs)
            arr[i].slice_start, arr[i].slice_interval, arr[i].slice_end = slice_
        cals = [[] for _ in sets]
        epochs = [[] for _ in sets]
        order = []

        def cb(_user, set_index, sol_p):
            s = sol_p.contents
            n = s.n_directions * s.n_antennas
            cals[set_index].append(np.ctypeslib.as_array(s.calibrations, shape=(n,)).copy().reshape(
                s.n_directions, s.n_antennas) if n else np.zeros((s.n_directions, s.n_antennas)))
            epochs[set_index].append((s.start_epoch, s.end_epoch))
            order.append((s.index, set_index))
            return 0

        cbf = BATCH_CB(cb)
        o = opts if opts is not None else options(strategy=LEAP_STRATEGY_BATCHED)
        self._chk(lib().leap_hip_calibrate_batch(self.h, arr, len(sets), ctypes.byref(o), cbf, None))
        self.last_batch_order = order
        return [(np.array(c), np.array(e)) for c, e in zip(cals, epochs)]

    def last_direction_block(self):
        return int(lib().leap_hip_last_direction_block(self.h))

    def last_route(self):
        """(channel_sum, direct_cube) of the last batched call: 1 Horner / 2 NUFFT, 1 = read the cube directly."""
        cs, dc = ctypes.c_int32(), ctypes.c_int32()
        self._chk(lib().leap_hip_last_route(self.h, ctypes.byref(cs), ctypes.byref(dc)))
        return cs.value, dc.value

    def last_workgroups_per_cu(self):
        return int(lib().leap_hip_last_workgroups_per_cu(self.h))

    def set_adjoint_mode(self, mode):
        """LEAP_ADJOINT_AUTO / _RECOMPUTE / _ROW_SUMS for later loss_grad calls."""
        self._chk(lib().leap_hip_set_adjoint_mode(self.h, int(mode)))

    def last_loss_grad_times(self):
        """(forward kernel ms, adjoint kernel ms) of the last loss_grad call."""
        f, a = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().leap_hip_last_loss_grad_times(self.h, ctypes.byref(f), ctypes.byref(a)))
        return f.value, a.value

    def last_graph(self):
        return int(lib().leap_hip_last_graph(self.h))

    def tuning_json(self):
        """The autotuners' choices held by this context, as a JSON string."""
        n = ctypes.c_uint64()
        self._chk(lib().leap_hip_tuning_save(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._chk(lib().leap_hip_tuning_save(self.h, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def load_tuning(self, text):
        self._chk(lib().leap_hip_tuning_load(self.h, text.encode()))

    def last_staged_bytes(self):
        return int(lib().leap_hip_last_staged_bytes(self.h))

    def last_pack_times(self):
        a = ctypes.c_double()
        n = ctypes.c_int32()
        self._chk(lib().leap_hip_last_pack_times(self.h, ctypes.byref(a), ctypes.byref(n)))
        return a.value, n.value

    def last_kernel_times(self):
        a = ctypes.c_double()
        n = ctypes.c_int32()
        t = ctypes.c_double()
        self._chk(lib().leap_hip_last_kernel_times(self.h, ctypes.byref(a), ctypes.byref(n), ctypes.byref(t)))
        return a.value, n.value, t.value


class Metadata:
    def __init__(self, ctx, view, ref_ant, threshold, use_cache):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._chk(lib().leap_hip_metadata_create(ctx.h, ctypes.byref(view), -1 if ref_ant is None else ref_ant,
                                                threshold, 1 if use_cache else 0, ctypes.byref(h)))
        self.h = h
        dims = np.zeros(5, dtype=np.int32)
        lib().leap_hip_metadata_dims(h, dims.ctypes.data_as(_i32p))
        self.K, self.N, self.K1, self.nbl, self.reference_antenna = (int(x) for x in dims)
        self.num_pols = view.num_pols

    def copy(self):
        A = np.zeros((self.N, self.K))
        Ad = np.zeros((self.K, self.N))
        A1 = np.zeros((self.N, self.K1))
        Ad1 = np.zeros((self.K1, self.N))
        I = np.zeros(max(self.K - 1, 1), dtype=np.int32)
        I1 = np.zeros(max(self.K1 - 1, 1), dtype=np.int32)
        _check(lib().leap_hip_metadata_copy(self.h, _dptr(A), I.ctypes.data_as(_i32p), _dptr(Ad), _dptr(A1),
                                            I1.ctypes.data_as(_i32p), _dptr(Ad1)))
        return {"A": A.T.copy(), "I": I[: self.K - 1].copy(), "Ad": Ad.T.copy(), "A1": A1.T.copy(),
                "I1": I1[: self.K1 - 1].copy(), "Ad1": Ad1.T.copy()}

    def diagnostics(self):
        """(Ad degenerate, Ad1 degenerate): the reference's warnings of MetaData.cc:129-136."""
        a, a1 = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().leap_hip_metadata_diagnostics(self.h, ctypes.byref(a), ctypes.byref(a1)))
        return bool(a.value), bool(a1.value)

    def rotate(self, view, directions, row_start=0, n_rows=None, opts=None):
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        if n_rows is None:
            n_rows = view.num_rows - row_start
        out = np.zeros((len(dirs), self.nbl, self.num_pols), dtype=np.complex128)
        o = opts if opts is not None else options()
        self.ctx._chk(lib().leap_hip_rotate_visibilities(self.ctx.h, self.h, ctypes.byref(view), row_start, n_rows,
                                                         _dptr(dirs), len(dirs), ctypes.byref(o),
                                                         out.ctypes.data_as(_dp)))
        return out

    def apply_inversion(self, avg, opts=None):
        avg = np.ascontiguousarray(avg, dtype=np.complex128).reshape(-1, self.nbl, self.num_pols)
        cal = np.zeros((avg.shape[0], self.N))
        o = opts if opts is not None else options()
        self.ctx._chk(lib().leap_hip_apply_inversion(self.ctx.h, self.h, avg.ctypes.data_as(_dp), avg.shape[0],
                                                     ctypes.byref(o), _dptr(cal)))
        return cal

    def loss_grad(self, view, directions, enu_to_uvw, offsets=None, antenna_delta=None, want_grad=True):
        """Calibration-residual loss and its gradients (offsets [D, 2], antenna positions [N, 3])."""
        dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
        D = len(dirs)
        off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.float64).reshape(D, 2)
        dp = None if antenna_delta is None else np.ascontiguousarray(antenna_delta, dtype=np.float64).reshape(self.N, 3)
        M = np.ascontiguousarray(enu_to_uvw, dtype=np.float64)
        loss = ctypes.c_double()
        g_off = np.zeros((D, 2))
        g_ant = np.zeros((self.N, 3))
        self.ctx._chk(lib().leap_hip_residual_loss_grad(
            self.ctx.h, self.h, ctypes.byref(view), _dptr(dirs), None if off is None else _dptr(off), D,
            None if dp is None else _dptr(dp), _dptr(M), ctypes.byref(loss),
            _dptr(g_off) if want_grad else None, _dptr(g_ant) if want_grad else None))
        return loss.value, g_off, g_ant

    def __del__(self):
        try:
            lib().leap_hip_metadata_destroy(self.h)
        except Exception:  # noqa: BLE001
            pass

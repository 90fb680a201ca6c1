"""The batched strategy (one pass over the visibilities per block of
directions) against the oracle.

Bars: FP64 batched — avgData pol 0 within 1e-12 of its largest magnitude and
calibrations within 1e-9 rad wrapped; mixed precision — calibrations within
1e-5 rad wrapped (max and RMS are printed per case).
"""
import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_TOL = 1e-5


def wrapped(a, b):
    return np.abs(npr.wrap(np.asarray(a) - np.asarray(b)))


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


@pytest.fixture(scope="module")
def small(leap):
    s = leap.SynthMS(n_antennas=40, n_timesteps=5, n_channels=77, flag_fraction=0.02, nan_fraction=0.002, seed=31)
    s.fill_host()
    return s


@pytest.mark.parametrize("db", [0, 32, 64, 128, 256])
@pytest.mark.parametrize("n_dirs", [1, 37, 70])
def test_batched_fp64_rotate_matches_oracle(leap, ctx, oracle, small, db, n_dirs):
    ms = small.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], n_dirs, n_dirs)
    opts = leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db)
    got = md.rotate(view, dirs, opts=opts)
    for d in range(0, n_dirs, max(1, n_dirs // 5)):
        want = omd.rotate(dirs[d])
        assert np.abs(got[d, :, 0] - want[:, 0]).max() <= 1e-12 * np.abs(want[:, 0]).max()
        assert not got[d, :, 1:].any()  # only pol 0 is accumulated by default


@pytest.mark.parametrize("db,nd", [(64, 2), (128, 4), (256, 2), (256, 4)])
def test_batched_lane_directions(leap, ctx, oracle, small, db, nd):
    ms = small.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], 41, 11)
    got = md.rotate(view, dirs, opts=leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db,
                                                  lane_directions=nd))
    assert ctx.last_direction_block() == db * 10 + nd
    for d in (0, 1, 17, 40):
        want = omd.rotate(dirs[d])
        assert np.abs(got[d, :, 0] - want[:, 0]).max() <= 1e-12 * np.abs(want[:, 0]).max()


@pytest.mark.parametrize("tile", [(0, 0), (64, 2), (256, 1)])
@pytest.mark.parametrize("n_pols", [1, 2, 3, 4])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batched_device_cube_alignment(leap, ctx, oracle, dtype, n_pols, tile):
    """A device cube at a 16-byte-aligned address and the same cube 8 bytes past one give the
    same bits (the pol-0 tiles copy 16-byte-multiple elements of an aligned cube straight into
    LDS and stage the rest through registers), both within the FP64 bar of the oracle."""
    s = leap.SynthMS(n_antennas=12, n_timesteps=9, n_channels=83, n_pols=n_pols, nan_fraction=0.004, seed=41)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(dtype))
    omd = oracle.metadata(ms)
    md = ctx.metadata(leap.view_from_arrays(ms))
    dirs = dirs_near(ms["phase_centre"], 45, 7)
    db, nd = tile
    opts = leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db, lane_directions=nd)
    raw = np.ascontiguousarray(ms["vis"]).view(np.uint8)
    base = ctx.malloc(raw.nbytes + 64)
    try:
        got = []
        for off in (0, 8):
            ctx.h2d(base + off, raw)
            view = leap.view_from_arrays(ms, vis_device=base + off, vis_dtype=leap.LEAP_VIS_C64 if dtype == np.complex64
                                         else leap.LEAP_VIS_C128)
            got.append(md.rotate(view, dirs, opts=opts))
    finally:
        ctx.free(base)
    assert np.array_equal(got[0], got[1])
    for d in (0, 13, 44):
        want = omd.rotate(dirs[d])
        assert np.abs(got[0][d, :, 0] - want[:, 0]).max() <= 1e-12 * np.abs(want[:, 0]).max()


EDGE_SHAPES = [  # (antennas, timesteps, channels, pols)
    (4, 1, 1, 1),    # one row block shorter than the rows in flight, a single channel
    (5, 2, 5, 3),    # fewer channels than a chunk; 48-byte complex128 elements (rows not a power of two)
    (6, 3, 33, 4),   # one channel past a 32-channel chunk
    (3, 11, 64, 2),  # whole chunks, a partial last row block
    (2, 4, 200, 4),  # three baselines (two antennas and their autocorrelations)
]


@pytest.mark.parametrize("tile", [(64, 2), (32, 1), (256, 2)])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("shape", EDGE_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_batched_fp64_edge_shapes(leap, ctx, oracle, shape, dtype, tile):
    """The FP64 tiles straight from a device cube at the shapes the chunking and the cube copies
    handle specially, against the oracle."""
    n_ant, n_t, n_ch, n_pols = shape
    s = leap.SynthMS(n_antennas=n_ant, n_timesteps=n_t, n_channels=n_ch, n_pols=n_pols, nan_fraction=0.01, seed=43)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(dtype))
    omd = oracle.metadata(ms)
    md = ctx.metadata(leap.view_from_arrays(ms))
    dirs = dirs_near(ms["phase_centre"], 9, 5)
    raw = np.ascontiguousarray(ms["vis"]).view(np.uint8)
    p = ctx.malloc(max(raw.nbytes, 16))
    try:
        ctx.h2d(p, raw)
        view = leap.view_from_arrays(ms, vis_device=p,
                                     vis_dtype=leap.LEAP_VIS_C64 if dtype == np.complex64 else leap.LEAP_VIS_C128)
        got = md.rotate(view, dirs, opts=leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=tile[0],
                                                      lane_directions=tile[1]))
    finally:
        ctx.free(p)
    for d in range(len(dirs)):
        want = omd.rotate(dirs[d])
        scale = max(np.abs(want[:, 0]).max(), 1e-300)
        assert np.abs(got[d, :, 0] - want[:, 0]).max() <= 1e-12 * scale, d


@pytest.mark.parametrize("nd", [1, 2])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batched_all_pols_matches_oracle(leap, ctx, oracle, dtype, nd):
    s = leap.SynthMS(n_antennas=16, n_timesteps=7, n_channels=100, nan_fraction=0.003, seed=32)
    s.fill_host(dtype)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], 5, 2)
    got = md.rotate(view, dirs, opts=leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, all_pols=1,
                                                  direction_block=64, lane_directions=nd))
    for d in range(5):
        want = omd.rotate(dirs[d])
        assert np.abs(got[d] - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("slice_", [(0, -1, -1), (0, 2, -1)])
def test_batched_fp64_calibrate_matches_oracle(leap, ctx, oracle, small, slice_):
    ms = small.host_dict()
    view = leap.view_from_arrays(ms)
    dirs = dirs_near(ms["phase_centre"], 20, 5)
    got, ge = ctx.calibrate(view, dirs, slice_, ref_ant=0, opts=leap.options(strategy=leap.LEAP_STRATEGY_BATCHED))
    want, we = oracle.calibrate(ms, dirs, slice_, ref_ant=0)
    assert np.array_equal(ge, we)
    assert wrapped(got, want).max() < FP64_TOL


ROUTES = {"nufft": dict(channel_sum=2), "nufft-fp32": dict(channel_sum=3), "horner": dict(channel_sum=1),
          "fp32-blocks": dict(fp32_blocks=1)}


@pytest.mark.parametrize("route", list(ROUTES))
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_mixed_precision_calibrate_error_budget(leap, ctx, oracle, dtype, route):
    fp32_blocks = route == "fp32-blocks"
    s = leap.SynthMS(n_antennas=48, n_timesteps=8, n_channels=384, freq_start_hz=50e6, freq_inc_hz=781.25e3,
                     layout=1, array_radius_m=20000.0, flag_fraction=0.02, seed=33)
    s.fill_host(np.complex128)
    ms64 = s.host_dict()
    vis = s.vis.astype(np.complex64) if dtype == np.complex64 else s.vis
    ms = s.host_dict(vis)
    view = leap.view_from_arrays(ms)
    dirs = dirs_near(ms["phase_centre"], 24, 6)
    got, _ = ctx.calibrate(view, dirs, opts=leap.options(precision=leap.LEAP_PRECISION_MIXED,
                                                         strategy=leap.LEAP_STRATEGY_BATCHED, **ROUTES[route]))
    # against the FP64 oracle on the same stored visibilities
    want, _ = oracle.calibrate(ms, dirs)
    err = wrapped(got, want)
    per_dir_max = err.max(axis=(0, 2))
    per_dir_rms = np.sqrt((err ** 2).mean(axis=(0, 2)))
    print(f"mixed {np.dtype(dtype).name} {route}: wrapped error max {per_dir_max.max():.2e} rad, "
          f"rms {per_dir_rms.max():.2e} rad (worst direction)")
    # The FP32 blocked Horner and the FP32 NUFFT transform are kept for comparison (faster); they
    # do not meet the 1e-5 rad budget.
    assert per_dir_max.max() < {"fp32-blocks": 1e-2, "nufft-fp32": 1e-4}.get(route, MIXED_TOL)
    if dtype == np.complex64:
        # the cube itself was rounded to complex64: report the FP64-cube distance too
        ref64, _ = oracle.calibrate(ms64, dirs)
        print(f"  vs FP64 oracle on the complex128 cube: max {wrapped(got, ref64).max():.2e} rad")


def test_mixed_rotate_relative_error(leap, ctx, oracle, small):
    ms = small.host_dict(small.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], 8, 9)
    got = md.rotate(view, dirs, opts=leap.options(precision=leap.LEAP_PRECISION_MIXED,
                                                  strategy=leap.LEAP_STRATEGY_BATCHED))
    blk = md.rotate(view, dirs, opts=leap.options(precision=leap.LEAP_PRECISION_MIXED,
                                                  strategy=leap.LEAP_STRATEGY_BATCHED, fp32_blocks=1))
    for d in range(8):
        want = omd.rotate(dirs[d])[:, 0]
        rel = np.abs(got[d, :, 0] - want).max() / np.abs(want).max()
        assert rel < 1e-12, rel
        rel_blk = np.abs(blk[d, :, 0] - want).max() / np.abs(want).max()
        assert rel_blk < 2e-5, rel_blk


def test_mixed_requires_batched(leap, ctx, small):
    view = leap.view_from_arrays(small.host_dict())
    with pytest.raises(leap.LeapError, match="batched"):
        ctx.calibrate(view, [(0.5, -0.47)], opts=leap.options(precision=leap.LEAP_PRECISION_MIXED))
    with pytest.raises(leap.LeapError, match="direction block"):
        ctx.calibrate(view, [(0.5, -0.47)], opts=leap.options(strategy=leap.LEAP_STRATEGY_BATCHED,
                                                               direction_block=48))


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 2e-6)], ids=["fp64-gemm", "bf16x3-mfma"])
@pytest.mark.parametrize("n_dirs", [1, 16, 70, 300])
def test_batched_inversion_gemm(leap, ctx, oracle, precision, tol, n_dirs):
    """cal = Ad * dphi + cal1 as one GEMM (FP64) or split-3 x BF16 MFMA (mixed) against the oracle.
    Up to 256 directions and antennas the MFMA product runs as whole-output workgroups; 300
    directions take the one-wave-per-tile kernel."""
    s = leap.SynthMS(n_antennas=61, autocorrelations=0, n_timesteps=2, n_channels=4, flag_fraction=0.03, seed=40)
    s.fill_host()
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view, ref_ant=3)
    omd = oracle.metadata(ms, ref_ant=3)
    rng = np.random.default_rng(n_dirs)
    avg = rng.standard_normal((n_dirs, ms["num_baselines"], 4)) + 1j * rng.standard_normal((n_dirs, ms["num_baselines"], 4))
    got = md.apply_inversion(avg, opts=leap.options(precision=precision, strategy=leap.LEAP_STRATEGY_BATCHED))
    err = max(wrapped(got[d], omd.apply_inversion(avg[d])).max() for d in range(n_dirs))
    print(f"inversion precision={precision} D={n_dirs}: max wrapped error {err:.2e} rad")
    assert err < tol

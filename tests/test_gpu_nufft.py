"""NUFFT channel sums of the mixed-precision batched path (LEAP_CHANNEL_SUM_NUFFT)
against the oracle.

Both transform precisions are covered: LEAP_CHANNEL_SUM_NUFFT (FP64) and
LEAP_CHANNEL_SUM_NUFFT_FP32 (FP32 transform, grid and interpolation).

Bars:
- avgData: for every direction and baseline |got - oracle| <= f * A, where
  A = sum over the interval's rows and channels of |v| for that baseline (pol 0,
  non-finite (row, channel)s excluded), f = 5e-10 for the FP64 transform and
  2e-7 for the FP32 one.  The NUFFT's error for one row and direction is a
  fraction of sum_c |v_c| (1e-11 .. 1e-10 with the FP64 transform, 2e-8 .. 7e-8
  with the FP32 one, tests/test_nufft_plan.py), so the bound is independent of
  size and data.
- calibrations: within 1e-5 rad wrapped of the FP64 oracle (the mixed budget)
  for the FP64 transform; the FP32 transform is outside it (within 1e-4 rad).
- host-resident and device-resident cubes give identical results, whether the
  NUFFT kernel reads the complex64 cube itself or packed pol-0 rows; chunked
  packs agree with one chunk to summation order.
"""
import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu

NUFFT = 2
NUFFT_FP32 = 3
HORNER = 1
# (route, f of the avgData bound f * A)
ROUTES = [(NUFFT, 5e-10), (NUFFT_FP32, 2e-7)]
ROUTE_IDS = ["fp64-transform", "fp32-transform"]


def wrapped(a, b):
    return np.abs(npr.wrap(np.asarray(a) - np.asarray(b)))


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


def abs_mass(ms):
    """sum over rows and channels of |pol 0| per baseline, skipping non-finite (row, channel)s."""
    vis = ms["vis"]
    ok = np.isfinite(vis.real).all(axis=2) & np.isfinite(vis.imag).all(axis=2)
    a = np.where(ok, np.abs(np.nan_to_num(vis[:, :, 0].astype(np.complex128))), 0.0).sum(axis=1)
    nbl = ms["num_baselines"]
    return a.reshape(-1, nbl).sum(axis=0)


def mixed(**kw):
    return dict(precision=1, strategy=1, **kw)


@pytest.mark.parametrize("route,f", ROUTES, ids=ROUTE_IDS)
@pytest.mark.parametrize("tile", [641, 1282, 2564])
@pytest.mark.parametrize("n_ch,n_dirs", [(48, 37), (77, 300), (130, 70), (256, 16), (384, 64), (500, 20)])
def test_nufft_rotate_matches_oracle(leap, ctx, oracle, n_ch, n_dirs, tile, route, f):
    print(f"nufft rotate: {n_ch} channels, {n_dirs} directions, tile {tile}, route {route}", flush=True)
    s = leap.SynthMS(n_antennas=20, n_timesteps=5, n_channels=n_ch, freq_start_hz=50e6, freq_inc_hz=781.25e3,
                     layout=1, array_radius_m=20000.0, nan_fraction=0.003, seed=200 + n_ch)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], n_dirs, 5, spread=0.1)
    db, nd = tile // 10, tile % 10
    got = md.rotate(view, dirs, opts=leap.options(**mixed(direction_block=db, lane_directions=nd, channel_sum=route)))
    assert ctx.last_route()[0] == route
    bound = f * abs_mass(ms)
    worst = 0.0
    for d in range(n_dirs):
        want = omd.rotate(dirs[d])[:, 0]
        err = np.abs(got[d, :, 0] - want)
        worst = max(worst, (err / bound).max())
        assert np.all(err <= bound), d
        assert not np.any(got[d, :, 1:])  # pols 1.. are not accumulated
    print(f"  worst error / bound {worst:.3f}")


@pytest.mark.parametrize("route,f", ROUTES, ids=ROUTE_IDS)
def test_nufft_whole_rounds_and_chunked_tail(leap, ctx, oracle, route, f):
    # 64 antennas with autocorrelations = 2,080 baselines: more than one round of
    # resident workgroups on a 256-CU part (2,048 at two waves per SIMD), so the
    # launch runs whole rounds of baselines plus a few baselines split into row
    # chunks whose partial sums are added afterwards.
    print("nufft rotate: 64 antennas (2,080 baselines), 6 timesteps, 160 channels", flush=True)
    s = leap.SynthMS(n_antennas=64, n_timesteps=6, n_channels=160, freq_start_hz=50e6, freq_inc_hz=781.25e3,
                     layout=1, array_radius_m=20000.0, nan_fraction=0.002, seed=218)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    n_dirs = 9
    dirs = dirs_near(ms["phase_centre"], n_dirs, 6, spread=0.1)
    got = md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=route)))
    assert got.shape[1] == 2080
    bound = f * abs_mass(ms)
    for d in range(n_dirs):
        want = omd.rotate(dirs[d])[:, 0]
        err = np.abs(got[d, :, 0] - want)
        assert np.all(err <= bound), (d, int(err.argmax()), (err / bound).max())
        assert not np.any(got[d, :, 1:])


@pytest.mark.parametrize("route,f", ROUTES, ids=ROUTE_IDS)
@pytest.mark.parametrize("n_ch,n_dirs,n_t,n_pols", [(1, 1, 3, 4), (7, 5, 1, 2), (64, 65, 2, 1), (512, 3, 2, 4)])
def test_nufft_edge_shapes_match_oracle(leap, ctx, oracle, n_ch, n_dirs, n_t, n_pols, route, f):
    """Smallest and largest channel counts, one timestep, one direction and
    direction counts one past a block, on the device cube and on packed rows.
    The tile is fixed: the packed and the cube route are tuned separately, and
    with a single direction block the launch sums its last round in row chunks,
    so two different tuned tiles can differ in rounding."""
    print(f"nufft edge: {n_ch} ch, {n_dirs} dirs, {n_t} t, {n_pols} pols", flush=True)
    s = leap.SynthMS(n_antennas=10, n_timesteps=n_t, n_channels=n_ch, n_pols=n_pols, nan_fraction=0.01, seed=217)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    dirs = dirs_near(ms["phase_centre"], n_dirs, 11, spread=0.1)
    omd = oracle.metadata(ms)
    bound = f * abs_mass(ms)
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    opts = leap.options(**mixed(channel_sum=route, direction_block=64, lane_directions=1))
    packed = md.rotate(view, dirs, opts=opts)
    vis_p = ctx.malloc(ms["vis"].nbytes)
    try:
        ctx.h2d(vis_p, np.ascontiguousarray(ms["vis"]))
        dview = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=leap.LEAP_VIS_C64)
        cube = md.rotate(dview, dirs, opts=opts)
        assert ctx.last_route() == (route, 1)
    finally:
        ctx.free(vis_p)
    assert np.array_equal(packed, cube)
    for d in range(n_dirs):
        assert np.all(np.abs(cube[d, :, 0] - omd.rotate(dirs[d])[:, 0]) <= bound + 1e-300), d


@pytest.mark.parametrize("route", [NUFFT, NUFFT_FP32], ids=ROUTE_IDS)
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_nufft_host_and_device_cubes_agree(leap, ctx, dtype, route):
    s = leap.SynthMS(n_antennas=18, n_timesteps=6, n_channels=200, nan_fraction=0.004, seed=211)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(dtype))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    dirs = dirs_near(ms["phase_centre"], 90, 6)
    opts = leap.options(**mixed(channel_sum=route, direction_block=128, lane_directions=2))
    host = md.rotate(view, dirs, opts=opts)
    vis_p = ctx.malloc(ms["vis"].nbytes)
    try:
        ctx.h2d(vis_p, np.ascontiguousarray(ms["vis"]))
        dview = leap.view_from_arrays(ms, vis_device=vis_p,
                                      vis_dtype=leap.LEAP_VIS_C64 if dtype == np.complex64 else leap.LEAP_VIS_C128)
        dev = md.rotate(dview, dirs, opts=opts)
    finally:
        ctx.free(vis_p)
    assert np.array_equal(host, dev)


@pytest.mark.parametrize("route,f", ROUTES, ids=ROUTE_IDS)
@pytest.mark.parametrize("n_pols", [1, 2, 3, 4])
def test_nufft_cube_kernel_equals_packed(leap, ctx, oracle, n_pols, route, f):
    """A device-resident complex64 cube with 1, 2 or 4 pols is read by the NUFFT
    kernel itself (mask and precorrection in registers); packing first (asked
    for through pack_timesteps, and always for 3 pols) gives the same bits."""
    s = leap.SynthMS(n_antennas=16, n_timesteps=6, n_channels=150, n_pols=n_pols, nan_fraction=0.01, seed=215)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    dirs = dirs_near(ms["phase_centre"], 50, 10, spread=0.1)
    vis_p = ctx.malloc(ms["vis"].nbytes)
    try:
        ctx.h2d(vis_p, np.ascontiguousarray(ms["vis"]))
        dview = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=leap.LEAP_VIS_C64)
        md = ctx.metadata(dview)
        direct = md.rotate(dview, dirs, opts=leap.options(**mixed(channel_sum=route)))
        assert ctx.last_pack_times()[1] == (1 if n_pols == 3 else 0)
        packed = md.rotate(dview, dirs, opts=leap.options(**mixed(channel_sum=route, pack_timesteps=6)))
        assert ctx.last_pack_times()[1] == 1
        assert np.array_equal(direct, packed)
        if n_pols != 3:
            forced = md.rotate(dview, dirs, opts=leap.options(**mixed(channel_sum=route, direct_cube=1)))
            assert np.array_equal(forced, direct)
        else:
            with pytest.raises(leap.LeapError, match="reads the cube directly"):
                md.rotate(dview, dirs, opts=leap.options(**mixed(channel_sum=route, direct_cube=1)))
    finally:
        ctx.free(vis_p)
    omd = oracle.metadata(ms)
    bound = f * abs_mass(ms)
    for d in (0, 25, 49):
        assert np.all(np.abs(direct[d, :, 0] - omd.rotate(dirs[d])[:, 0]) <= bound)


@pytest.mark.parametrize("route", [NUFFT, NUFFT_FP32], ids=ROUTE_IDS)
@pytest.mark.parametrize("chunk", [3, 7])
def test_nufft_chunks_agree_with_one_chunk(leap, ctx, chunk, route):
    s = leap.SynthMS(n_antennas=16, n_timesteps=20, n_channels=96, nan_fraction=0.002, seed=212)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    dirs = dirs_near(ms["phase_centre"], 40, 7)
    one = md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=route)))
    got = md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=route, pack_timesteps=chunk)))
    assert ctx.last_kernel_times()[1] == -(-20 // chunk)
    assert np.abs(got - one).max() <= 1e-13 * np.abs(one).max()


@pytest.mark.parametrize("route", [NUFFT, NUFFT_FP32], ids=ROUTE_IDS)
def test_nufft_calibration_error_budget(leap, ctx, oracle, route):
    """The error-budget case of the mixed path (decorrelated long baselines,
    where the inversion amplifies phase errors) on the NUFFT channel sums."""
    s = leap.SynthMS(n_antennas=48, n_timesteps=8, n_channels=384, freq_start_hz=50e6, freq_inc_hz=781.25e3,
                     layout=1, array_radius_m=20000.0, flag_fraction=0.02, seed=33)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    dirs = dirs_near(ms["phase_centre"], 24, 6)
    got, _ = ctx.calibrate(view, dirs, opts=leap.options(**mixed(channel_sum=route)))
    want, _ = oracle.calibrate(ms, dirs)
    err = wrapped(got, want)
    print(f"nufft calibrate (route {route}): wrapped error max {err.max():.2e} rad, rms {np.sqrt((err ** 2).mean()):.2e} rad")
    assert err.max() < (1e-5 if route == NUFFT else 1e-4)


def test_nufft_is_the_automatic_mixed_choice(leap, ctx):
    s = leap.SynthMS(n_antennas=12, n_timesteps=4, n_channels=128, seed=213)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    dirs = dirs_near(ms["phase_centre"], 32, 8)
    auto = md.rotate(view, dirs, opts=leap.options(**mixed(direction_block=64, lane_directions=1)))
    nufft = md.rotate(view, dirs, opts=leap.options(**mixed(direction_block=64, lane_directions=1, channel_sum=NUFFT)))
    horner = md.rotate(view, dirs, opts=leap.options(**mixed(direction_block=64, lane_directions=1,
                                                             channel_sum=HORNER)))
    assert np.array_equal(auto, nufft)
    assert not np.array_equal(auto, horner)
    assert ctx.last_route() == (HORNER, 0)
    # few directions stay on Horner
    few = md.rotate(view, dirs[:4], opts=leap.options(**mixed(direction_block=64, lane_directions=1)))
    assert ctx.last_route()[0] == HORNER
    assert np.array_equal(few, horner[:4])
    # and so do few channels
    s2 = leap.SynthMS(n_antennas=12, n_timesteps=4, n_channels=80, seed=216)
    s2.fill_host(np.complex128)
    ms2 = s2.host_dict(s2.vis.astype(np.complex64))
    view2 = leap.view_from_arrays(ms2)
    ctx.metadata(view2).rotate(view2, dirs, opts=leap.options(**mixed()))
    assert ctx.last_route()[0] == HORNER


def test_nufft_option_errors(leap, ctx):
    s = leap.SynthMS(n_antennas=8, n_timesteps=2, n_channels=520, seed=214)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(np.complex64))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    dirs = dirs_near(ms["phase_centre"], 20, 9)
    for route in (NUFFT, NUFFT_FP32):
        with pytest.raises(leap.LeapError, match="mixed-precision"):
            md.rotate(view, dirs, opts=leap.options(strategy=1, channel_sum=route))
        for bad in (dict(all_pols=1), dict(fp32_blocks=1)):
            with pytest.raises(leap.LeapError, match="run on pol 0"):
                md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=route, **bad)))
        with pytest.raises(leap.LeapError, match="1..512 channels"):
            md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=route)))
    for bad in (-1, 5):
        with pytest.raises(leap.LeapError, match="channel_sum"):
            md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=bad)))
    # 520 channels: the automatic choice stays on Horner
    auto = md.rotate(view, dirs, opts=leap.options(**mixed()))
    horner = md.rotate(view, dirs, opts=leap.options(**mixed(channel_sum=HORNER)))
    assert np.array_equal(auto, horner)

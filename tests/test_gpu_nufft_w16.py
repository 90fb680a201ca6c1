"""The width-16 NUFFT channel sums (LEAP_CHANNEL_SUM_NUFFT_W16) against the oracle.

The width-16 plan keeps the kernel's truncation (exp(-beta) <= 1.3e-16) below FP64
rounding, so its channel sums are as close to the exact sums as the FP64 Horner
chains (tests/test_nufft_plan.py pins the plan against extended-precision sums).
It is the automatic choice of FP64 batched pol-0 calls from 16 directions and
96..512 channels, read from a complex64 or complex128 cube on the device or from
packed pol-0 rows.

Bars (those of the FP64 batched Horner path, tests/test_gpu_batched.py):
- avgData pol 0 within 1e-12 of the direction's largest magnitude of the oracle's
  on a 1.5 km array;
- within 1e-12 of the FP64 Horner route on the same call;
- calibrations within 1e-9 rad wrapped of the FP64 oracle;
- the cube sources and packed rows give identical bits (the precorrection is the
  same FP64 product in the kernel and in the pack kernel).
On a 40 km array the phase arguments reach 1e4 rad and more, whose own FP64
rounding (eps |phase| per term, in the oracle as in every route) puts the oracle
1e-12 to 3e-12 of the largest magnitude from both routes; there the width-16
route's distance from the oracle may not exceed the Horner route's by more than
a tenth (measured: 1.32e-12 against 1.32e-12, the routes 2e-14 apart at 0.02 rad
and 2e-13 at 0.1 rad).
"""
import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu

W16 = 4
HORNER = 1
FP64_TOL = 1e-9


def wrapped(a, b):
    return np.abs(npr.wrap(np.asarray(a) - np.asarray(b)))


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


def fp64(**kw):
    return dict(precision=0, strategy=1, **kw)


def synth(leap, n_ch, n_t=5, n_pols=4, n_ant=20, seed=300, nan_fraction=0.003, radius=1500.0):
    s = leap.SynthMS(n_antennas=n_ant, n_timesteps=n_t, n_channels=n_ch, n_pols=n_pols, freq_start_hz=50e6,
                     freq_inc_hz=781.25e3, layout=1, array_radius_m=radius, nan_fraction=nan_fraction,
                     seed=seed + n_ch)
    s.fill_host(np.complex128)
    return s


class DeviceCube:
    """The cube copied to the device (at a byte offset), freed on exit."""

    def __init__(self, leap, ctx, ms, offset=0):
        self.ctx = ctx
        raw = np.ascontiguousarray(ms["vis"]).view(np.uint8)
        self.p = ctx.malloc(raw.nbytes + 64)
        ctx.h2d(self.p + offset, raw)
        vd = leap.LEAP_VIS_C64 if ms["vis"].dtype == np.complex64 else leap.LEAP_VIS_C128
        self.view = leap.view_from_arrays(ms, vis_device=self.p + offset, vis_dtype=vd)

    def __enter__(self):
        return self.view

    def __exit__(self, *exc):
        self.ctx.free(self.p)


def check_against_oracle(got, omd, dirs, tol=1e-12):
    worst = 0.0
    for d in range(len(dirs)):
        want = omd.rotate(dirs[d])[:, 0]
        scale = np.abs(want).max()
        err = np.abs(got[d, :, 0] - want).max()
        worst = max(worst, err / scale)
        assert err <= tol * scale, (d, err / scale)
        assert not np.any(got[d, :, 1:])  # pols 1.. are not accumulated
    return worst


SOURCES = ["complex128-cube", "complex64-cube", "packed"]


@pytest.mark.parametrize("source", SOURCES)
@pytest.mark.parametrize("tile", [641, 1282, 2564])
@pytest.mark.parametrize("n_ch,n_dirs", [(48, 37), (130, 70), (384, 64), (500, 20)])
def test_w16_rotate_matches_oracle(leap, ctx, oracle, n_ch, n_dirs, tile, source):
    print(f"w16 rotate: {n_ch} channels, {n_dirs} directions, tile {tile}, {source}", flush=True)
    s = synth(leap, n_ch)
    ms = s.host_dict(s.vis.astype(np.complex64)) if source == "complex64-cube" else s.host_dict()
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], n_dirs, 5)
    db, nd = tile // 10, tile % 10
    opts = leap.options(**fp64(direction_block=db, lane_directions=nd, channel_sum=W16))
    if source == "packed":
        view = leap.view_from_arrays(ms)
        md = ctx.metadata(view)
        got = md.rotate(view, dirs, opts=opts)
        assert ctx.last_route() == (W16, 0)
    else:
        with DeviceCube(leap, ctx, ms) as view:
            md = ctx.metadata(view)
            got = md.rotate(view, dirs, opts=opts)
            assert ctx.last_route() == (W16, 1)
    worst = check_against_oracle(got, omd, dirs)
    print(f"  worst error / largest |avgData| {worst:.2e}")


@pytest.mark.parametrize("n_pols", [1, 2, 4])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_w16_cube_equals_packed_and_horner(leap, ctx, oracle, dtype, n_pols):
    """The kernel reading the cube gives the bits of the packed rows, and both
    stay within the FP64 bar of the Horner route on the same call."""
    s = synth(leap, 200, n_pols=n_pols, nan_fraction=0.01, radius=20000.0)
    ms = s.host_dict(s.vis.astype(dtype))
    dirs = dirs_near(ms["phase_centre"], 48, 7)
    opts = leap.options(**fp64(direction_block=64, lane_directions=1, channel_sum=W16))
    host_view = leap.view_from_arrays(ms)
    md = ctx.metadata(host_view)
    packed = md.rotate(host_view, dirs, opts=opts)
    assert ctx.last_route() == (W16, 0)
    with DeviceCube(leap, ctx, ms) as view:
        cube = md.rotate(view, dirs, opts=opts)
        assert ctx.last_route() == (W16, 1)
        horner = md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=HORNER)))
    assert np.array_equal(packed, cube)
    for d in range(len(dirs)):
        scale = np.abs(horner[d, :, 0]).max()
        assert np.abs(cube[d, :, 0] - horner[d, :, 0]).max() <= 1e-12 * scale, d
    check_against_oracle(cube[::6], oracle.metadata(ms), dirs[::6], tol=5e-12)


@pytest.mark.parametrize("spread", [0.02, 0.1])
@pytest.mark.parametrize("n_ch", [130, 384])
def test_w16_on_a_40km_array_as_close_as_horner(leap, ctx, oracle, n_ch, spread):
    """The 40 km array, directions up to 0.02 or 0.1 rad off the phase centre:
    within 1e-12 of the Horner route, and no further from the oracle than it."""
    s = synth(leap, n_ch, radius=20000.0)
    ms = s.host_dict()
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], 64, 18, spread=spread)
    with DeviceCube(leap, ctx, ms) as view:
        md = ctx.metadata(view)
        got = md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=W16)))
        assert ctx.last_route() == (W16, 1)
        horner = md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=HORNER)))
    for d in range(len(dirs)):
        scale = np.abs(horner[d, :, 0]).max()
        assert np.abs(got[d, :, 0] - horner[d, :, 0]).max() <= 1e-12 * scale, d
    worst = check_against_oracle(got[::4], omd, dirs[::4], tol=5e-12)
    worst_h = check_against_oracle(horner[::4], omd, dirs[::4], tol=5e-12)
    print(f"w16 {n_ch} ch at {spread} rad: worst error / largest |avgData| {worst:.2e} (Horner {worst_h:.2e})")
    assert worst <= 1.1 * worst_h + 1e-14


def test_w16_misaligned_cube_takes_the_packed_rows(leap, ctx):
    """A complex128 cube 8 bytes off a 16-byte boundary is packed first (the
    kernel loads 16 bytes at a time), with the same bits."""
    s = synth(leap, 130)
    ms = s.host_dict()
    dirs = dirs_near(ms["phase_centre"], 40, 8)
    opts = leap.options(**fp64(direction_block=64, lane_directions=1, channel_sum=W16))
    with DeviceCube(leap, ctx, ms, offset=0) as view:
        md = ctx.metadata(view)
        aligned = md.rotate(view, dirs, opts=opts)
        assert ctx.last_route() == (W16, 1)
    with DeviceCube(leap, ctx, ms, offset=8) as view:
        off = md.rotate(view, dirs, opts=opts)
        assert ctx.last_route() == (W16, 0)
        with pytest.raises(leap.LeapError, match="aligned"):
            md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=W16, direct_cube=1)))
    assert np.array_equal(aligned, off)


@pytest.mark.parametrize("n_ch,n_dirs,n_t,n_pols", [(1, 1, 3, 4), (7, 5, 1, 2), (64, 65, 2, 1), (512, 3, 2, 4),
                                                   (300, 20, 2, 3)])
def test_w16_edge_shapes_match_oracle(leap, ctx, oracle, n_ch, n_dirs, n_t, n_pols):
    """Smallest and largest channel counts, one timestep, one direction, a direction
    count one past a block, and 3 pols (which the kernel does not read itself: the
    device cube is packed first), on the device cube and on packed rows."""
    print(f"w16 edge: {n_ch} ch, {n_dirs} dirs, {n_t} t, {n_pols} pols", flush=True)
    s = leap.SynthMS(n_antennas=10, n_timesteps=n_t, n_channels=n_ch, n_pols=n_pols, nan_fraction=0.01, seed=317)
    s.fill_host(np.complex128)
    ms = s.host_dict()
    dirs = dirs_near(ms["phase_centre"], n_dirs, 11)
    opts = leap.options(**fp64(channel_sum=W16, direction_block=64, lane_directions=1))
    host_view = leap.view_from_arrays(ms)
    md = ctx.metadata(host_view)
    packed = md.rotate(host_view, dirs, opts=opts)
    with DeviceCube(leap, ctx, ms) as view:
        cube = md.rotate(view, dirs, opts=opts)
    assert np.array_equal(packed, cube)
    check_against_oracle(cube, oracle.metadata(ms), dirs)


@pytest.mark.parametrize("n_ch,nd", [(160, 1), (384, 1), (384, 2), (384, 4)])
def test_w16_skips_non_finite_pols_like_horner(leap, ctx, n_ch, nd):
    """NaN or infinity in any pol of a (row, channel), pol 0 finite, drops that
    channel from the sum on the complex128 cube, as the Horner route does: with
    the upper words of pols 1.. loaded (512-point grid; 4 directions per lane) and
    with every pol loaded whole (1024-point grid, 1 or 2 directions per lane)."""
    s = synth(leap, n_ch, n_t=4, nan_fraction=0.0)
    ms = s.host_dict()
    nbl = ms["num_baselines"]
    vis = np.array(ms["vis"]).reshape(ms["num_rows"], ms["num_channels"], ms["num_pols"])
    planted = [(1, 3, 5, 1, complex(np.nan, 0.0)), (2, 4, 70, 3, complex(0.5, -np.inf)),
               (0, 6, 159, 2, complex(np.inf, 1.0)), (3, 7, 64, 3, complex(-0.0, np.nan))]
    for t, bl, ch, pol, val in planted:
        vis[t * nbl + bl, ch, pol] = val
    ms = s.host_dict(vis.reshape(ms["vis"].shape))
    dirs = dirs_near(ms["phase_centre"], 64 * nd, 12)
    tile = dict(direction_block=64 * nd, lane_directions=nd)
    with DeviceCube(leap, ctx, ms) as view:
        md = ctx.metadata(view)
        w16 = md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=W16, **tile)))
        assert ctx.last_route() == (W16, 1)
        horner = md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=HORNER)))
    assert np.all(np.isfinite(w16))
    for d in range(len(dirs)):
        scale = np.abs(horner[d, :, 0]).max()
        assert np.abs(w16[d, :, 0] - horner[d, :, 0]).max() <= 1e-12 * scale, d
    # a planted baseline differs from the cube without the plant by that channel's term
    clean = s.host_dict()
    with DeviceCube(leap, ctx, clean) as view:
        ref = md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=W16, **tile)))
    for t, bl, ch, pol, _ in planted:
        assert np.abs(w16[:, bl, 0] - ref[:, bl, 0]).max() > 1e-6 * np.abs(ref[:, bl, 0]).max()


def test_w16_is_the_automatic_fp64_choice(leap, ctx):
    s = synth(leap, 160, n_t=3)
    ms = s.host_dict()
    with DeviceCube(leap, ctx, ms) as view:
        md = ctx.metadata(view)
        for n_dirs, want in [(15, HORNER), (16, W16), (200, W16)]:
            md.rotate(view, dirs_near(ms["phase_centre"], n_dirs, 13), opts=leap.options(**fp64()))
            assert ctx.last_route()[0] == want, n_dirs
    few = synth(leap, 90, n_t=2)
    ms = few.host_dict()
    with DeviceCube(leap, ctx, ms) as view:
        md = ctx.metadata(view)
        md.rotate(view, dirs_near(ms["phase_centre"], 64, 14), opts=leap.options(**fp64()))
        assert ctx.last_route()[0] == HORNER  # fewer than 96 channels


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_w16_calibrate_matches_oracle(leap, ctx, oracle, dtype):
    s = leap.SynthMS(n_antennas=48, n_timesteps=8, n_channels=384, freq_start_hz=50e6, freq_inc_hz=781.25e3,
                     layout=1, array_radius_m=20000.0, flag_fraction=0.02, seed=333)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(dtype))
    dirs = dirs_near(ms["phase_centre"], 40, 16, spread=0.02)
    want, we = oracle.calibrate(ms, dirs)
    with DeviceCube(leap, ctx, ms) as view:
        got, ge = ctx.calibrate(view, dirs, opts=leap.options(**fp64()))
        assert ctx.last_route() == (W16, 1)
        horner, _ = ctx.calibrate(view, dirs, opts=leap.options(**fp64(channel_sum=HORNER)))
    assert np.array_equal(ge, we)
    err = wrapped(got, want)
    print(f"w16 {np.dtype(dtype).name}: calibration error vs oracle max {err.max():.2e} rad, "
          f"vs Horner route max {wrapped(got, horner).max():.2e} rad")
    assert err.max() < FP64_TOL
    assert wrapped(horner, want).max() < FP64_TOL


def test_w16_option_errors(leap, ctx):
    s = synth(leap, 100, n_t=2)
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    dirs = dirs_near(ms["phase_centre"], 40, 17)
    with pytest.raises(leap.LeapError, match="width 16"):
        md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=2)))
    with pytest.raises(leap.LeapError, match="pol 0"):
        md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=W16, all_pols=1)))
    with pytest.raises(leap.LeapError, match="channel_sum"):
        md.rotate(view, dirs, opts=leap.options(**fp64(channel_sum=5)))
    # the multi-set call takes it too (tests/test_gpu_multi.py)
    ctx.calibrate_batch([(md, view, dirs, (0, -1, -1))], leap.options(**fp64(channel_sum=W16)))
    assert ctx.last_route()[0] == W16

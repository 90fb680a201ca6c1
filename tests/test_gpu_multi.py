"""Several measurement sets of uneven shape in one persistent launch.

Small sets run against the oracle (FP64 <= 1e-9 rad wrapped, mixed <= 1e-5
rad), on the FP64 Horner tile and, in mixed precision, on the NUFFT channel
sums (sets of 96..512 channels with 16 or more directions and complex64 cubes,
one launch per group of sets that share a grid and kernel width).  Every
(baseline, direction block) of the NUFFT launch runs the single-set kernel's
code in its order, so it equals the single-set NUFFT bit for bit wherever the
single-set launch does not split a baseline's rows into chunks (more than one
direction block).  The three sets of the uneven benchmark batch run at full
size against the single-set batched kernel that reads the cube directly with
the same tile, which must agree bit for bit on the Horner route (the same FP64
Horner tile code runs per work item), within rounding on the NUFFT route, and
against the oracle for a few directions of the two smaller sets.
"""
import ctypes

import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu


def wrapped(a, b):
    return np.abs(npr.wrap(np.asarray(a) - np.asarray(b)))


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


SMALL = [  # synth kwargs, directions, slice, reference antenna, visibility dtype
    (dict(n_antennas=12, autocorrelations=0, n_timesteps=6, n_channels=40, seed=51, nan_fraction=0.003), 5,
     (0, 2, -1), 0, np.complex128),
    (dict(n_antennas=20, n_timesteps=4, n_channels=9, seed=52, flag_fraction=0.05), 70, (0, -1, -1), None,
     np.complex128),
    (dict(n_antennas=9, n_timesteps=5, n_channels=300, seed=53), 1, (0, 1, -1), 3, np.complex64),
    (dict(n_antennas=7, n_timesteps=2, n_channels=3, seed=54), 0, (0, -1, -1), None, np.complex128),
]


@pytest.fixture(scope="module")
def small_sets(leap):
    out = []
    for kw, nd, slice_, ref, dt in SMALL:
        s = leap.SynthMS(**kw)
        s.fill_host(np.complex128)
        ms = s.host_dict(s.vis.astype(dt) if dt == np.complex64 else s.vis)
        out.append((s, ms, dirs_near(ms["phase_centre"], nd, nd + 1), slice_, ref))
    return out


def _run_small(leap, ctx, small_sets, opts):
    views = [leap.view_from_arrays(ms) for _, ms, _, _, _ in small_sets]
    mds = [ctx.metadata(v, ref_ant=ref) for v, (_, _, _, _, ref) in zip(views, small_sets)]
    res = ctx.calibrate_batch([(md, v, d, sl) for md, v, (_, _, d, sl, _) in zip(mds, views, small_sets)], opts)
    return res


@pytest.mark.parametrize("db,nd", [(0, 0), (32, 1), (64, 1), (128, 1), (128, 2), (256, 2)])
def test_multi_fp64_matches_oracle(leap, ctx, oracle, small_sets, db, nd):
    opts = leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db, lane_directions=nd)
    res = _run_small(leap, ctx, small_sets, opts)
    if db:
        assert ctx.last_direction_block() == db * 10 + max(nd, 1)
    for (s, ms, dirs, slice_, ref), (got, ge) in zip(small_sets, res):
        want, we = oracle.calibrate(ms, dirs, slice_, ref_ant=ref)
        assert got.shape == want.shape
        assert np.array_equal(ge, we)
        if got.size:
            assert wrapped(got, want).max() < 1e-9
    # callbacks come interval by interval, sets in order within an interval
    assert ctx.last_batch_order == sorted(ctx.last_batch_order)


def test_multi_mixed_matches_oracle(leap, ctx, oracle, small_sets):
    opts = leap.options(precision=leap.LEAP_PRECISION_MIXED, strategy=leap.LEAP_STRATEGY_BATCHED)
    res = _run_small(leap, ctx, small_sets, opts)
    worst = 0.0
    for (s, ms, dirs, slice_, ref), (got, _) in zip(small_sets, res):
        want, _ = oracle.calibrate(ms, dirs, slice_, ref_ant=ref)
        if got.size:
            worst = max(worst, wrapped(got, want).max())
    print(f"multi-set mixed: max wrapped error {worst:.2e} rad")
    assert worst < 1e-5


# complex64 sets for the NUFFT route: synth kwargs, directions, slice; the last
# one (48 channels) stays on the Horner tile under the automatic rule.
NUFFT_SETS = [
    (dict(n_antennas=16, n_timesteps=4, n_channels=200, seed=81, nan_fraction=0.002), 40, (0, 2, -1)),
    (dict(n_antennas=22, autocorrelations=0, n_timesteps=3, n_channels=130, seed=82, n_pols=2), 70, (0, -1, -1)),
    (dict(n_antennas=12, n_timesteps=5, n_channels=300, seed=83, n_pols=1), 20, (0, 1, -1)),
    (dict(n_antennas=14, n_timesteps=4, n_channels=384, seed=84), 150, (0, -1, -1)),
    (dict(n_antennas=10, n_timesteps=3, n_channels=48, seed=85), 30, (0, -1, -1)),
]


@pytest.fixture(scope="module")
def nufft_sets(leap):
    out = []
    for kw, nd, slice_ in NUFFT_SETS:
        s = leap.SynthMS(**kw)
        ms = s.host_dict(s.fill_host(np.complex64))
        out.append((ms, dirs_near(ms["phase_centre"], nd, nd + 7, spread=0.05), slice_))
    return out


@pytest.mark.parametrize("channel_sum", [0, 2], ids=["auto", "nufft"])
def test_multi_nufft_matches_oracle(leap, ctx, oracle, nufft_sets, channel_sum):
    views = [leap.view_from_arrays(ms) for ms, _, _ in nufft_sets]
    mds = [ctx.metadata(v) for v in views]
    opts = leap.options(precision=leap.LEAP_PRECISION_MIXED, strategy=leap.LEAP_STRATEGY_BATCHED,
                        channel_sum=channel_sum)
    res = ctx.calibrate_batch([(md, v, d, sl) for md, v, (_, d, sl) in zip(mds, views, nufft_sets)], opts)
    assert ctx.last_route()[0] == 2
    worst = 0.0
    for (ms, dirs, slice_), (got, ge) in zip(nufft_sets, res):
        want, we = oracle.calibrate(ms, dirs, slice_)
        assert np.array_equal(ge, we)
        worst = max(worst, wrapped(got, want).max())
    print(f"multi-set NUFFT ({'auto' if channel_sum == 0 else 'forced'}): max wrapped error {worst:.2e} rad")
    assert worst < 1e-5


@pytest.mark.parametrize("db,nd", [(64, 1), (128, 2)])
def test_multi_nufft_equals_single_set(leap, ctx, nufft_sets, db, nd):
    """Sets with more directions than one block: the single-set launch runs every
    (baseline, direction block) whole, as the multi-set launch does."""
    pick = [(ms, d, sl) for ms, d, sl in nufft_sets[:4] if len(d) > 64 * nd]
    assert pick
    views = [leap.view_from_arrays(ms) for ms, _, _ in pick]
    mds = [ctx.metadata(v) for v in views]
    opts = leap.options(precision=leap.LEAP_PRECISION_MIXED, strategy=leap.LEAP_STRATEGY_BATCHED,
                        direction_block=db, lane_directions=nd)
    res = ctx.calibrate_batch([(md, v, d, sl) for md, v, (_, d, sl) in zip(mds, views, pick)], opts)
    assert ctx.last_route()[0] == 2
    for md, v, (ms, d, sl), (got, ge) in zip(mds, views, pick, res):
        one, oe = ctx.calibrate(v, d, sl, opts=opts, md=md)
        assert ctx.last_route()[0] == 2
        assert np.array_equal(ge, oe)
        assert np.array_equal(got, one), "multi-set and single-set NUFFT results differ"


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_multi_w16_matches_oracle_and_single_set(leap, ctx, oracle, dtype):
    """FP64 multi-set calls: qualifying sets take the width-16 NUFFT (complex64 and
    complex128 cubes), within 1e-9 rad of the oracle, and bit-equal to the
    single-set launch of the same tile for sets with more directions than one block
    (the single-set launch then runs every (baseline, direction block) whole)."""
    sets = []
    for kw, nd, slice_ in NUFFT_SETS:
        s = leap.SynthMS(**kw)
        ms = s.host_dict(s.fill_host(dtype))
        sets.append((ms, dirs_near(ms["phase_centre"], nd, nd + 9, spread=0.05), slice_))
    views = [leap.view_from_arrays(ms) for ms, _, _ in sets]
    mds = [ctx.metadata(v) for v in views]
    opts = leap.options(strategy=leap.LEAP_STRATEGY_BATCHED)
    res = ctx.calibrate_batch([(md, v, d, sl) for md, v, (_, d, sl) in zip(mds, views, sets)], opts)
    assert ctx.last_route()[0] == 4
    worst = 0.0
    for (ms, dirs, slice_), (got, ge) in zip(sets, res):
        want, we = oracle.calibrate(ms, dirs, slice_)
        assert np.array_equal(ge, we)
        worst = max(worst, wrapped(got, want).max())
    print(f"multi-set width-16 NUFFT, {np.dtype(dtype).name}: max wrapped error {worst:.2e} rad")
    assert worst < 1e-9
    pick = [i for i, (_, d, _) in enumerate(sets[:4]) if len(d) > 64]
    assert pick
    tiled = leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=64, lane_directions=1)
    res = ctx.calibrate_batch([(mds[i], views[i], sets[i][1], sets[i][2]) for i in pick], tiled)
    assert ctx.last_route()[0] == 4
    for i, (got, ge) in zip(pick, res):
        one, oe = ctx.calibrate(views[i], sets[i][1], sets[i][2], opts=tiled, md=mds[i])
        assert ctx.last_route()[0] == 4
        assert np.array_equal(ge, oe)
        assert np.array_equal(got, one), "multi-set and single-set width-16 NUFFT results differ"


def test_multi_nufft_rejects_unsupported(leap, ctx, small_sets, nufft_sets):
    ms, dirs, slice_ = small_sets[0][1], small_sets[0][2], small_sets[0][3]
    v = leap.view_from_arrays(ms)  # complex128
    md = ctx.metadata(v)
    with pytest.raises(leap.LeapError, match="complex64 cube"):
        ctx.calibrate_batch([(md, v, dirs, slice_)], leap.options(precision=1, strategy=1, channel_sum=2))
    ms, dirs, slice_ = nufft_sets[0]
    v = leap.view_from_arrays(ms)
    md = ctx.metadata(v)
    with pytest.raises(leap.LeapError, match="FP64 transform"):
        ctx.calibrate_batch([(md, v, dirs, slice_)], leap.options(precision=1, strategy=1, channel_sum=3))


def test_multi_rejects_unsupported(leap, ctx, small_sets):
    s, ms, dirs, slice_, ref = small_sets[0]
    v = leap.view_from_arrays(ms)
    md = ctx.metadata(v, ref_ant=ref)
    with pytest.raises(leap.LeapError, match="batched strategy"):
        ctx.calibrate_batch([(md, v, dirs, slice_)], leap.options())
    with pytest.raises(leap.LeapError, match="polarisation 0"):
        ctx.calibrate_batch([(md, v, dirs, slice_)], leap.options(strategy=leap.LEAP_STRATEGY_BATCHED, all_pols=1))
    with pytest.raises(leap.LeapError, match="not a multi-set tile"):
        ctx.calibrate_batch([(md, v, dirs, slice_)], leap.options(strategy=leap.LEAP_STRATEGY_BATCHED,
                                                                  direction_block=256, lane_directions=4))
    assert ctx.calibrate_batch([], leap.options(strategy=leap.LEAP_STRATEGY_BATCHED)) == []


# The uneven benchmark batch: antennas, autocorrelations, timesteps, channels.
UNEVEN = [(36, 0, 32, 300), (128, 1, 14, 48), (256, 1, 64, 96)]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_uneven_batch_full_size(leap, ctx, oracle, precision):
    dt = np.complex64 if precision else np.complex128
    vd = leap.LEAP_VIS_C64 if precision else leap.LEAP_VIS_C128
    sets, ptrs = [], []
    try:
        for i, (n_ant, autos, T, C) in enumerate(UNEVEN):
            s = leap.SynthMS(n_antennas=n_ant, autocorrelations=autos, n_timesteps=T, n_channels=C,
                             freq_start_hz=50e6, freq_inc_hz=781.25e3 * 384 / C, seed=60 + i)
            nbytes = s.num_rows * s.num_channels * s.num_pols * np.dtype(dt).itemsize
            vis_p = ctx.malloc(nbytes)
            uvw_p = ctx.malloc(s.num_rows * 3 * 8)
            ptrs += [vis_p, uvw_p]
            ms = s.host_dict(None)
            host = None
            if i < 2:  # host cube for the oracle spot check, uploaded as is
                host = s.fill_host(dt)
                ctx.h2d(vis_p, host)
                ctx.h2d(uvw_p, np.ascontiguousarray(s.uvw))
            else:
                ctx.synth_fill_device(s, vd, vis_p, uvw_p)
            view = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=vd, uvw_device=uvw_p)
            dirs = dirs_near(ms["phase_centre"], 64, 70 + i, spread=0.1)
            sets.append((s, ms, host, view, ctx.metadata(view), dirs))
        batch = [(md, v, d, (0, -1, -1)) for (_, _, _, v, md, d) in sets]
        for db, nd in ((128, 2), (64, 1)):
            # the Horner tile for every set (asked for), bit-equal to the single-set kernel that
            # shares the multi-set tile code and reads the cube directly
            hopts = leap.options(precision=precision, strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db,
                                 lane_directions=nd, channel_sum=1)
            hres = ctx.calibrate_batch(batch, hopts)
            assert ctx.last_route()[0] == 1
            horner_ms = ctx.last_kernel_times()[0]
            single_ms = 0.0
            one_opts = leap.options(precision=precision, strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db,
                                    lane_directions=nd, direct_cube=1, channel_sum=1)
            for (s, ms, host, view, md, dirs), (got, _) in zip(sets, hres):
                one, _ = ctx.calibrate(view, dirs, opts=one_opts, md=md)
                single_ms += ctx.last_kernel_times()[0]
                assert np.array_equal(got, one), "multi-set and single-set results differ"
            res = hres
            line = (f"uneven batch {'mixed' if precision else 'fp64'} tile {db}x{nd}: Horner, one launch "
                    f"{horner_ms:.3f} ms, three single-set launches {single_ms:.3f} ms")
            # the automatic route: the 300- and 96-channel sets on the NUFFT (mixed precision) or the
            # width-16 NUFFT (FP64), the 48-channel set on Horner
            opts = leap.options(precision=precision, strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=db,
                                lane_directions=nd)
            res = ctx.calibrate_batch(batch, opts)
            assert ctx.last_route()[0] == (2 if precision else 4)
            nufft_ms = ctx.last_kernel_times()[0]
            assert np.array_equal(res[1][0], hres[1][0])  # the Horner set: the same tile code
            for (s, ms, host, view, md, dirs), (got, _), (hgot, _) in zip(sets, res, hres):
                one, _ = ctx.calibrate(view, dirs, opts=opts, md=md)
                # the single-set NUFFT launch splits its last round's rows into chunks
                err = wrapped(got, one).max()
                assert err < 1e-10, err
                if not precision:
                    assert wrapped(got, hgot).max() < 1e-9  # FP64: the width-16 NUFFT and Horner agree
            line += (f"; automatic route ({'NUFFT' if precision else 'width-16 NUFFT'} for 2 of 3 sets), one round "
                     f"{nufft_ms:.3f} ms")
            print(line)
        # the same batch (tile 64 x 1) captured as one HIP graph and replayed: bit-equal to the eager launch
        gopts = leap.options(precision=precision, strategy=leap.LEAP_STRATEGY_BATCHED, direction_block=64,
                             lane_directions=1, graph=1)
        for k in range(3):
            resg = ctx.calibrate_batch([(md, v, d, (0, -1, -1)) for (_, _, _, v, md, d) in sets], gopts)
            assert ctx.last_graph() == (1 if k == 0 else 2)
            for (got, ge), (want, we) in zip(resg, res):
                assert np.array_equal(got, want) and np.array_equal(ge, we), k
        for (s, ms, host, view, md, dirs), (got, _) in zip(sets[:2], res[:2]):
            pick = [0, 33, 63]
            want, _ = oracle.calibrate(dict(ms, vis=host), dirs[pick])
            assert wrapped(got[:, pick], want).max() < (1e-5 if precision else 1e-9)
    finally:
        for p in ptrs:
            ctx.free(p)

"""The packed pol-0 path: pack (pol 0 as FP64, non-finite (row, channel)s
zeroed) then the scalar-load FP64 Horner kernel, in one chunk or in chunks of
timesteps packed on a second stream.

One chunk runs the same operations in the same order as the Horner kernel that
reads the cube directly (64 lanes per row), so the two agree bit for bit;
chunked runs accumulate per chunk and are held to the oracle (avgData within
1e-12 of its largest magnitude, calibrations within 1e-9 rad wrapped).
"""
import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu


def wrapped(a, b):
    return np.abs(npr.wrap(np.asarray(a) - np.asarray(b)))


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


@pytest.mark.parametrize("tile", [641, 1282, 2564])
@pytest.mark.parametrize("dtype,n_pols,n_ch", [(np.complex128, 4, 77), (np.complex64, 4, 64), (np.complex64, 2, 33),
                                               (np.complex128, 1, 300)])
def test_packed_equals_direct_cube(leap, ctx, dtype, n_pols, n_ch, tile):
    s = leap.SynthMS(n_antennas=24, n_timesteps=7, n_channels=n_ch, n_pols=n_pols, nan_fraction=0.004, seed=91)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(dtype))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    dirs = dirs_near(ms["phase_centre"], 150, 3)
    db, nd = tile // 10, tile % 10
    # the FP64 Horner route (the automatic choice would take the width-16 NUFFT for 300 channels)
    packed = md.rotate(view, dirs, opts=leap.options(strategy=1, direction_block=db, lane_directions=nd,
                                                     channel_sum=1))
    # a host cube is packed on the host while it is staged: pol 0 only crosses the bus
    assert ctx.last_staged_bytes() == s.num_rows * n_ch * np.dtype(dtype).itemsize
    direct = md.rotate(view, dirs, opts=leap.options(strategy=1, direction_block=db, lane_directions=nd,
                                                     direct_cube=1, channel_sum=1))
    assert ctx.last_staged_bytes() == s.num_rows * n_ch * n_pols * np.dtype(dtype).itemsize
    assert np.array_equal(packed, direct)
    # the same cube resident on the device goes through the device pack kernel
    vis_p = ctx.malloc(ms["vis"].nbytes)
    try:
        ctx.h2d(vis_p, np.ascontiguousarray(ms["vis"]))
        dview = leap.view_from_arrays(ms, vis_device=vis_p,
                                      vis_dtype=leap.LEAP_VIS_C64 if dtype == np.complex64 else leap.LEAP_VIS_C128)
        dev = md.rotate(dview, dirs, opts=leap.options(strategy=1, direction_block=db, lane_directions=nd,
                                                       channel_sum=1))
        assert ctx.last_pack_times()[1] == 1
        assert np.array_equal(dev, direct)
    finally:
        ctx.free(vis_p)


@pytest.mark.parametrize("chunk", [0, 16, 7])
def test_packed_chunks_match_oracle(leap, ctx, oracle, chunk):
    s = leap.SynthMS(n_antennas=20, n_timesteps=40, n_channels=41, nan_fraction=0.002, seed=92)
    s.fill_host()
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], 70, 4)
    opts = leap.options(strategy=1, pack_timesteps=chunk)
    got = md.rotate(view, dirs, opts=opts)
    expect_chunks = 1 if chunk == 0 else -(-40 // chunk)
    assert ctx.last_kernel_times()[1] == expect_chunks  # one Horner launch per chunk
    for d in (0, 33, 69):
        want = omd.rotate(dirs[d])[:, 0]
        assert np.abs(got[d, :, 0] - want).max() <= 1e-12 * np.abs(want).max()
    cal, _ = ctx.calibrate(view, dirs, (0, 20, -1), opts=opts, md=md)
    ocal, _ = oracle.calibrate(ms, dirs, (0, 20, -1))
    assert wrapped(cal, ocal).max() < 1e-9


def test_packed_chunks_in_a_graph(leap, ctx):
    s = leap.SynthMS(n_antennas=16, n_timesteps=24, n_channels=30, seed=93)
    vis_p = ctx.malloc(s.num_rows * s.num_channels * s.num_pols * 16)
    uvw_p = ctx.malloc(s.num_rows * 24)
    try:
        ctx.synth_fill_device(s, leap.LEAP_VIS_C128, vis_p, uvw_p)
        ms = s.host_dict(None)
        view = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=leap.LEAP_VIS_C128, uvw_device=uvw_p)
        md = ctx.metadata(view)
        dirs = dirs_near(ms["phase_centre"], 40, 5)
        # an explicit packed tile: a tuned call this small would run straight from the cube
        kw = dict(strategy=1, direction_block=64, lane_directions=1, pack_timesteps=5)
        eager = ctx.calibrate(view, dirs, (0, 12, -1), opts=leap.options(**kw), md=md)
        assert ctx.last_pack_times()[1] == 6  # 2 intervals of 12 timesteps, chunks of 5
        g = leap.options(graph=1, **kw)
        ctx.calibrate(view, dirs, (0, 12, -1), opts=g, md=md)
        assert ctx.last_graph() == 1
        for rep in range(3):
            ctx.calibrate(view, dirs[:3], md=md)  # eager work in between
            got = ctx.calibrate(view, dirs, (0, 12, -1), opts=g, md=md)
            assert ctx.last_graph() == 2
            assert np.array_equal(got[0], eager[0]), f"replay {rep} differs"
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)


def test_route_follows_the_cost_model(leap, ctx):
    """Tuned calls on a device cube pack only when the FP64 work outweighs the pack pass (from
    about 800 directions on complex64 and 480 on complex128)."""
    s = leap.SynthMS(n_antennas=16, n_timesteps=4, n_channels=24, seed=94)
    for dtype, vd in ((np.complex64, leap.LEAP_VIS_C64), (np.complex128, leap.LEAP_VIS_C128)):
        host = s.fill_host(dtype)
        vis_p = ctx.malloc(host.nbytes)
        try:
            ctx.h2d(vis_p, host)
            ms = s.host_dict(host)
            view = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=vd)
            md = ctx.metadata(view)
            c = ms["phase_centre"]
            for D, packs in ((64, 0), (256, 0), (1024, 1)):
                dirs = np.column_stack([np.full(D, c[0]), np.linspace(c[1] - 0.01, c[1] + 0.01, D)])
                md.rotate(view, dirs, opts=leap.options(strategy=1))
                assert ctx.last_pack_times()[1] == packs, (dtype, D)
        finally:
            ctx.free(vis_p)

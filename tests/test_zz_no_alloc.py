"""leap_hip_calibrate_md_noalloc against leap_hip_calibrate_md (FP64, per
direction) through the C ABI: bit-identical calibrations and equal epochs, on
two calls in a row (the second on a smaller set with fewer directions, so the
scratch holds the first call's contents), and the size and residency checks
that reject a call before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def device_set(leap, ctx, vis_dtype, **kw):
    s = leap.SynthMS(**kw)
    elem = 8 if vis_dtype == leap.LEAP_VIS_C64 else 16
    vis = ctx.malloc(s.num_rows * s.num_channels * s.num_pols * elem)
    uvw = ctx.malloc(s.num_rows * 3 * 8)
    ctx.synth_fill_device(s, vis_dtype, vis, uvw)
    view = leap.view_from_arrays(s.host_dict(), vis_device=vis, vis_dtype=vis_dtype, uvw_device=uvw)
    return s, view, [vis, uvw]


def directions(centre, n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-0.02, 0.02, n), centre[1] + rng.uniform(-0.02, 0.02, n)])


def run_both(leap, ctx, view, md, dirs, slice_):
    want, want_epochs = ctx.calibrate(view, dirs, slice_, md=md)
    n_sol, D = want.shape[0], len(dirs)
    out = ctx.malloc(n_sol * D * md.N * 8)
    try:
        epochs = ctx.calibrate_noalloc(view, dirs, md, out, slice_)
        got = np.zeros((n_sol, D, md.N))
        ctx.d2h(got, out)
    finally:
        ctx.free(out)
    return got, epochs, want, want_epochs


def test_noalloc_matches_calibrate_twice(leap, ctx):
    bufs = []
    try:
        print("noalloc call 1: 40 antennas, 6 t, 70 ch, complex128, 5 directions, slice (0, 2, -1)", flush=True)
        s1, v1, b = device_set(leap, ctx, leap.LEAP_VIS_C128, n_antennas=40, n_timesteps=6, n_channels=70,
                               flag_fraction=0.03, seed=31)
        bufs += b
        md1 = ctx.metadata(v1, ref_ant=0)
        d1 = directions((v1.phase_centre_ra_rad, v1.phase_centre_dec_rad), 5, 1)
        got, ep, want, want_ep = run_both(leap, ctx, v1, md1, d1, (0, 2, -1))
        assert got.shape == (3, 5, md1.N)
        assert np.array_equal(got, want)
        assert np.array_equal(ep, want_ep)

        print("noalloc call 2: 16 antennas, 3 t, 33 ch, complex64, 2 directions, slice (0, -1, -1)", flush=True)
        s2, v2, b = device_set(leap, ctx, leap.LEAP_VIS_C64, n_antennas=16, n_timesteps=3, n_channels=33,
                               flag_fraction=0.05, seed=32)
        bufs += b
        md2 = ctx.metadata(v2)
        d2 = directions((v2.phase_centre_ra_rad, v2.phase_centre_dec_rad), 2, 2)
        got, ep, want, want_ep = run_both(leap, ctx, v2, md2, d2, (0, -1, -1))
        assert got.shape == (1, 2, md2.N)
        assert np.array_equal(got, want)
        assert np.array_equal(ep, want_ep)
    finally:
        for p in bufs:
            ctx.free(p)


def test_noalloc_rejects_before_device_work(leap, ctx):
    print("noalloc rejections: host-resident cube, working set above the scratch", flush=True)
    s = leap.SynthMS(n_antennas=12, n_timesteps=2, n_channels=8, seed=33)
    s.fill_host()
    host_view = leap.view_from_arrays(s.host_dict())
    md = ctx.metadata(host_view)
    dirs = directions(s.host_dict()["phase_centre"], 2, 3)
    out = ctx.malloc(2 * md.N * 8)
    try:
        with pytest.raises(leap.LeapError, match="device-resident"):
            ctx.calibrate_noalloc(host_view, dirs, md, out)
        _, dev_view, bufs = device_set(leap, ctx, leap.LEAP_VIS_C128, n_antennas=12, n_timesteps=2, n_channels=8,
                                       seed=33)
        try:
            many = directions(s.host_dict()["phase_centre"], 200_000, 4)  # 88 B each: above 16 MiB
            with pytest.raises(leap.LeapError, match="exceeds the no-allocation scratch"):
                ctx.calibrate_noalloc(dev_view, many, md, out)
        finally:
            for p in bufs:
                ctx.free(p)
    finally:
        ctx.free(out)

"""Parity at the benchmark's full size: 256 stations in a 40 km spiral, 32,896
baselines, 128 timesteps, 384 channels, 4 pols (the complex128 cube is
103.5 GB, the complex64 cube 51.7 GB, resident on the device).

- The first timestep as a one-interval slice, through the plain, straight-from-cube
  and packed paths, against the oracle on the same rows (<= 1e-9 rad wrapped).
- The whole interval: the batched FP64 paths against the plain per-direction
  path (two independent kernels; <= 1e-9 rad wrapped).
- Mixed precision on the complex64 cube against FP64 on the same values, on the
  NUFFT and both FP64 Horner routes (<= 1e-5 rad wrapped; max and RMS printed).
"""
import os
import sys

import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402  (the workload definition)


def wrapped(a, b):
    return np.abs(npr.wrap(np.asarray(a) - np.asarray(b)))


def device_cube(leap, ctx, synth, vd, itemsize):
    vis_p = ctx.malloc(synth.num_rows * synth.num_channels * synth.num_pols * itemsize)
    uvw_p = ctx.malloc(synth.num_rows * 24)
    ctx.synth_fill_device(synth, vd, vis_p, uvw_p)
    view = leap.view_from_arrays(synth.host_dict(None), vis_device=vis_p, vis_dtype=vd, uvw_device=uvw_p)
    return vis_p, uvw_p, view


def test_full_size_fp64(leap, ctx, oracle):
    s = leap.SynthMS(**bench.c1_config())
    nbl = s.num_baselines
    assert nbl == 32896 and s.num_rows == 32896 * 128
    vis_p, uvw_p, view = device_cube(leap, ctx, s, leap.LEAP_VIS_C128, 16)
    try:
        md = ctx.metadata(view)
        c = (view.phase_centre_ra_rad, view.phase_centre_dec_rad)
        dirs = bench.directions(c, 4, seed=11)
        plain = leap.options()
        direct = leap.options(strategy=1, direct_cube=1)
        packed = leap.options(strategy=1, direction_block=256, lane_directions=4)
        # first timestep against the oracle on those rows
        first = np.empty((nbl, s.num_channels, s.num_pols), dtype=np.complex128)
        ctx.d2h(first, vis_p)
        ms1 = s.host_dict(first)
        ms1 = dict(ms1, num_rows=nbl, time=ms1["time"][:nbl], uvw=ms1["uvw"][:nbl])
        want, we = oracle.calibrate(ms1, dirs[:2], (0, 1, 1))
        for name, o in (("plain", plain), ("direct", direct), ("packed", packed)):
            got, ge = ctx.calibrate(view, dirs[:2], (0, 1, 1), opts=o, md=md)
            assert np.array_equal(ge, we)
            err = wrapped(got, want).max()
            print(f"first timestep, {name} vs oracle: {err:.2e} rad")
            assert err < 1e-9
        # whole interval: batched paths against the plain path
        ref, _ = ctx.calibrate(view, dirs, opts=plain, md=md)
        for name, o in (("direct", direct), ("packed", packed)):
            got, _ = ctx.calibrate(view, dirs, opts=o, md=md)
            err = wrapped(got, ref).max()
            print(f"128 timesteps, {name} vs plain: {err:.2e} rad")
            assert err < 1e-9
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)


def test_full_size_mixed(leap, ctx):
    s = leap.SynthMS(**bench.c1_config())
    vis_p, uvw_p, view = device_cube(leap, ctx, s, leap.LEAP_VIS_C64, 8)
    try:
        md = ctx.metadata(view)
        c = (view.phase_centre_ra_rad, view.phase_centre_dec_rad)
        dirs = bench.directions(c, 32, seed=12)
        # FP64 (batched, one pass) on the same complex64 values
        ref, _ = ctx.calibrate(view, dirs, opts=leap.options(strategy=1), md=md)
        for name, kw in (("nufft", dict()),
                         ("horner packed", dict(channel_sum=1, direction_block=256, lane_directions=4)),
                         ("horner direct", dict(channel_sum=1, direct_cube=1))):
            got, _ = ctx.calibrate(view, dirs, opts=leap.options(precision=1, strategy=1, **kw), md=md)
            assert ctx.last_route()[0] == (2 if name == "nufft" else 1)
            err = wrapped(got, ref)
            print(f"mixed ({name}) vs FP64, 256 stations: max {err.max():.2e} rad, rms {np.sqrt((err ** 2).mean()):.2e}")
            assert err.max() < 1e-5
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)

"""Parity at the benchmark's full size: 256 stations in a 40 km spiral, 32,896
baselines, 128 timesteps, 384 channels, 4 pols (the complex128 cube is
103.5 GB, the complex64 cube 51.7 GB, resident on the device).

- The first timestep as a one-interval slice, through the plain, straight-from-cube
  and packed paths, against the oracle on the same rows (<= 1e-9 rad wrapped).
- The whole interval: the batched FP64 paths against the plain per-direction
  path (two independent kernels; <= 1e-9 rad wrapped).
- Mixed precision with 256 directions on the complex64 cube against FP64 on the
  same values, on the NUFFT and both FP64 Horner routes (<= 1e-5 rad
  wrapped; per-direction max and RMS printed).
- Batched FP64 with 64 directions on the complex128 cube under HIP graph
  capture (the automatic route: the width-16 NUFFT): replays bit-equal to the
  eager call, within 5e-9 rad of the plain path and 1e-9 rad of the FP64
  Horner route.
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


def per_direction(err):
    """max and RMS of wrapped errors [solutions][directions][antennas] for each direction."""
    return err.max(axis=(0, 2)), np.sqrt((err ** 2).mean(axis=(0, 2)))


def test_full_size_mixed_256_directions(leap, ctx):
    """The precision-plus-batching workload: 256 directions on the complex64 cube in
    one pass, against FP64 on the same values (the batched FP64 route, itself
    checked against the plain per-direction path here and, on complex128, in
    test_full_size_fp64).  Per-direction max and RMS wrapped error <= 1e-5 rad."""
    s = leap.SynthMS(**bench.c1_config())
    vis_p, uvw_p, view = device_cube(leap, ctx, s, leap.LEAP_VIS_C64, 8)
    try:
        md = ctx.metadata(view)
        c = (view.phase_centre_ra_rad, view.phase_centre_dec_rad)
        dirs = bench.directions(c, 256, seed=12)
        ref, _ = ctx.calibrate(view, dirs, opts=leap.options(strategy=1), md=md)
        pick = np.arange(0, 256, 37)
        plain, _ = ctx.calibrate(view, dirs[pick], opts=leap.options(), md=md)
        err = wrapped(ref[:, pick], plain).max()
        print(f"FP64 batched vs plain on the complex64 cube, {len(pick)} directions: {err:.2e} rad")
        assert err < 1e-9
        for name, kw, route in (("nufft", dict(), 2),
                                ("horner packed", dict(channel_sum=1, direction_block=256, lane_directions=4), 1),
                                ("horner direct", dict(channel_sum=1, direct_cube=1), 1)):
            got, _ = ctx.calibrate(view, dirs, opts=leap.options(precision=1, strategy=1, **kw), md=md)
            assert ctx.last_route()[0] == route
            mx, rms = per_direction(wrapped(got, ref))
            print(f"mixed ({name}) vs FP64, 256 stations, 256 directions: max {mx.max():.2e} rad "
                  f"(median direction {np.median(mx):.2e}), rms {rms.max():.2e} rad (worst direction)")
            assert mx.max() < 1e-5
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)


def test_full_size_batched_fp64_graph(leap, ctx):
    """Batching in FP64: 64 directions (offsets up to 0.1 rad) on the complex128
    cube in one pass, captured as a HIP graph and replayed; every replay
    bit-equal to the eager call, and within 5e-9 rad of the plain per-direction
    path.  The batched route takes the phase of channel c as sf (kappa0 + c dkappa)
    (a Horner recurrence in exp(i sf dkappa)), the plain path as sf / lambda_c with
    the rounded wavelength, as the reference does (CpuLeapCalibrator.cc:195): for
    phases of 1e3 to 1e4 rad the two FP64 roundings differ by about 1e-12 rad per
    term, growing with the offset, and a baseline whose sum nearly cancels turns
    that into a phase error the inversion spreads over the calibration (2.5e-9 rad
    measured here on the width-16 NUFFT and 2.8e-9 on the Horner route, which lie
    2.7e-10 apart; below 1e-9 with offsets up to 0.02 rad in test_full_size_fp64)."""
    s = leap.SynthMS(**bench.c1_config())
    vis_p, uvw_p, view = device_cube(leap, ctx, s, leap.LEAP_VIS_C128, 16)
    try:
        md = ctx.metadata(view)
        c = (view.phase_centre_ra_rad, view.phase_centre_dec_rad)
        dirs = bench.directions(c, 64, seed=13, spread=0.1)
        eager, ee = ctx.calibrate(view, dirs, opts=leap.options(strategy=1), md=md)
        assert ctx.last_route()[0] == leap.LEAP_CHANNEL_SUM_NUFFT_W16
        for k in range(3):
            got, ge = ctx.calibrate(view, dirs, opts=leap.options(strategy=1, graph=1), md=md)
            assert ctx.last_graph() == (1 if k == 0 else 2)
            assert np.array_equal(got, eager) and np.array_equal(ge, ee), k
        plain, _ = ctx.calibrate(view, dirs, opts=leap.options(), md=md)
        horner, _ = ctx.calibrate(view, dirs, opts=leap.options(strategy=1, channel_sum=1), md=md)
        err = wrapped(eager, plain).max()
        err_h = wrapped(eager, horner).max()
        print(f"batched FP64 (width-16 NUFFT, graph) vs plain, 64 directions on the complex128 cube: {err:.2e} rad; "
              f"vs the FP64 Horner route {err_h:.2e} rad (Horner vs plain {wrapped(horner, plain).max():.2e} rad)")
        assert err < 5e-9
        assert err_h < 1e-9
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)

"""Host-resident visibilities: staged through the pinned double buffer (every
path) or packed to pol 0 on the host (the packed path).

The device receives the same values either way, so results equal those of the
same cube resident on the device bit for bit; the packed path sends only pol 0
(a quarter of a 4-pol cube) across the bus.  The cube is large enough to need
several pinned chunks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


@pytest.fixture(scope="module", params=["c64", "c128"])
def big(request, leap, ctx):
    """128 antennas (8256 baselines) x 80 t x 48 ch x 4 pols with NaNs: 1.27 GB c64 / 2.5 GB c128 on the host."""
    c64 = request.param == "c64"
    vd = leap.LEAP_VIS_C64 if c64 else leap.LEAP_VIS_C128
    dt = np.complex64 if c64 else np.complex128
    s = leap.SynthMS(n_antennas=128, n_timesteps=80, n_channels=48, nan_fraction=1e-4, seed=95)
    nbytes = s.num_rows * s.num_channels * s.num_pols * np.dtype(dt).itemsize
    vis_p = ctx.malloc(nbytes)
    uvw_p = ctx.malloc(s.num_rows * 24)
    ctx.synth_fill_device(s, vd, vis_p, uvw_p)
    host = np.empty((s.num_rows, s.num_channels, s.num_pols), dtype=dt)
    ctx.d2h(host, vis_p)
    ms = s.host_dict(host)
    dev_view = leap.view_from_arrays(ms, vis_device=vis_p, vis_dtype=vd, uvw_device=uvw_p)
    host_view = leap.view_from_arrays(ms)
    md = ctx.metadata(host_view)
    yield s, ms, md, host_view, dev_view, np.dtype(dt).itemsize
    ctx.free(vis_p)
    ctx.free(uvw_p)


# Packed cases name their tile: a tuned call on the device cube would be routed straight from the cube
# at this direction count, while a host cube always takes the packed route.
PK = dict(strategy=1, direction_block=64, lane_directions=1)


@pytest.mark.parametrize("opts_kw", [PK, dict(PK, pack_timesteps=30), dict(PK, precision=1),
                                     dict(strategy=1, direct_cube=1), dict(strategy=0)],
                         ids=["packed", "packed-chunks", "packed-mixed", "direct-cube", "per-direction"])
def test_host_cube_equals_device_cube(leap, ctx, big, opts_kw):
    s, ms, md, host_view, dev_view, elem = big
    if opts_kw.get("precision") == 1 and elem == 16:
        pytest.skip("mixed precision is run on the complex64 cube")
    dirs = dirs_near(ms["phase_centre"], 3 if opts_kw["strategy"] == 0 else 40, 6)
    opts = leap.options(**opts_kw)
    got_host, eh = ctx.calibrate(host_view, dirs, (0, 40, -1), opts=opts, md=md)
    staged = ctx.last_staged_bytes()
    got_dev, ed = ctx.calibrate(dev_view, dirs, (0, 40, -1), opts=opts, md=md)
    assert ctx.last_staged_bytes() == 0
    assert np.array_equal(got_host, got_dev) and np.array_equal(eh, ed)
    rc = s.num_rows * s.num_channels
    if opts_kw["strategy"] == 1 and not opts_kw.get("direct_cube"):
        assert staged == rc * 2 * (elem // 2)  # pol 0 only, in the source precision
    else:
        assert staged == rc * s.num_pols * elem  # whole rows

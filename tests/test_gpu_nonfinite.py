"""Infinite visibilities, pinned per route.

The reference skips a (row, channel) when a *rotated* pol is NaN (CpuLeapCalibrator.cc:201-206),
the rotation being a std::complex product.  The per-direction FP64 path forms the same product
(a cs - b sn, a sn + b cs) and tests it: an infinite pol whose product stays infinite in both
parts is summed, so that baseline's avgData is infinite as in the oracle (which multiplies
std::complex values); one whose product has a NaN part is skipped, as in the oracle.  The batched
routes never form the rotated value: they skip every (row, channel) with a non-finite input, so an
infinity there gives the same bits as a NaN in its place (and the baseline stays finite, where the
reference's would be infinite).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF_BL, INF_T, INF_CH = 3, 1, 5  # pol 0 = +inf + 0.25i: both product parts infinite, summed
NAN_BL, NAN_T, NAN_CH = 5, 2, 7  # pol 1 = inf + inf i: one product part NaN, skipped


def dirs_near(centre, n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    return np.column_stack([centre[0] + rng.uniform(-spread, spread, n), centre[1] + rng.uniform(-spread, spread, n)])


def planted(leap, dtype, n_pols=4, n_ch=40, fill=complex(np.inf, 0.25), fill2=complex(np.inf, np.inf)):
    s = leap.SynthMS(n_antennas=10, n_timesteps=4, n_channels=n_ch, n_pols=n_pols, seed=77)
    s.fill_host(np.complex128)
    ms = s.host_dict(s.vis.astype(dtype))
    nbl = ms["num_baselines"]
    vis = np.array(ms["vis"]).reshape(ms["num_rows"], ms["num_channels"], ms["num_pols"])
    vis[INF_T * nbl + INF_BL, INF_CH, 0] = fill
    vis[NAN_T * nbl + NAN_BL, NAN_CH, min(1, n_pols - 1)] = fill2
    return s.host_dict(vis.reshape(ms["vis"].shape))


def test_per_direction_path_follows_the_rotated_product(leap, ctx, oracle):
    ms = planted(leap, np.complex128)
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    omd = oracle.metadata(ms)
    dirs = dirs_near(ms["phase_centre"], 6, 3)
    got = md.rotate(view, dirs)
    for d in range(len(dirs)):
        want = omd.rotate(dirs[d])
        # the summed infinity: pol 0 of that baseline infinite in both parts, with the oracle's signs
        g, w = got[d, INF_BL, 0], want[INF_BL, 0]
        assert np.isinf(g.real) and np.isinf(g.imag), g
        assert np.isinf(w.real) and np.isinf(w.imag), w
        assert np.sign(g.real) == np.sign(w.real) and np.sign(g.imag) == np.sign(w.imag)
        # everything else finite and equal to the oracle (the NaN-part product was skipped by both)
        mask = np.ones(got.shape[1:], dtype=bool)
        mask[INF_BL, 0] = False
        assert np.all(np.isfinite(got[d][mask])) and np.all(np.isfinite(want[mask]))
        scale = np.abs(want[mask]).max()
        assert np.abs(got[d][mask] - want[mask]).max() <= 1e-12 * scale


def nan_twin(leap, dtype, n_pols=4, n_ch=40):
    """The planted cube and the same cube with NaN where it holds infinities."""
    inf = planted(leap, dtype, n_pols, n_ch)
    nan = planted(leap, dtype, n_pols, n_ch, fill=complex(np.nan, 0.25), fill2=complex(np.nan, np.nan))
    return inf, nan


def device_view(leap, ctx, ms, vd, offset, held):
    raw = np.ascontiguousarray(ms["vis"]).view(np.uint8)
    p = ctx.malloc(raw.nbytes + 64)
    held.append(p)
    ctx.h2d(p + offset, raw)
    return leap.view_from_arrays(ms, vis_device=p + offset, vis_dtype=vd)


ROUTES = {
    # FP64 Horner: host cube (packed pol 0), device cube (copied into LDS), device cube at an
    # address off 16 bytes (staged through registers)
    "packed": dict(dtype=np.complex128, device=False, offset=0,
                   opts=dict(precision=0, strategy=1)),
    "cube": dict(dtype=np.complex128, device=True, offset=0,
                 opts=dict(precision=0, strategy=1, direction_block=64, lane_directions=2)),
    "cube-registers": dict(dtype=np.complex128, device=True, offset=8,
                           opts=dict(precision=0, strategy=1, direction_block=64, lane_directions=2)),
    "cube-complex64": dict(dtype=np.complex64, device=True, offset=0,
                           opts=dict(precision=0, strategy=1, direction_block=128, lane_directions=2)),
    # mixed precision, NUFFT channel sums from the complex64 cube
    "nufft": dict(dtype=np.complex64, device=True, offset=0, n_ch=100,
                  opts=dict(precision=1, strategy=1, channel_sum=2)),
}


@pytest.mark.parametrize("route", list(ROUTES))
def test_batched_routes_skip_infinities_as_nans(leap, ctx, oracle, route):
    r = ROUTES[route]
    ms_inf, ms_nan = nan_twin(leap, r["dtype"], n_ch=r.get("n_ch", 40))
    vd = leap.LEAP_VIS_C64 if r["dtype"] == np.complex64 else leap.LEAP_VIS_C128
    dirs = dirs_near(ms_inf["phase_centre"], 70, 9)
    opts = leap.options(**r["opts"])
    held = []
    try:
        got = []
        for ms in (ms_inf, ms_nan):
            view = device_view(leap, ctx, ms, vd, r["offset"], held) if r["device"] else leap.view_from_arrays(ms)
            md = ctx.metadata(view)
            got.append(md.rotate(view, dirs, opts=opts))
    finally:
        for p in held:
            ctx.free(p)
    assert np.array_equal(got[0], got[1])
    assert np.all(np.isfinite(got[0]))
    if r["opts"]["precision"] == 0:
        # and both are the oracle's sums over the finite (row, channel)s, pol 0
        omd = oracle.metadata(ms_nan)
        for d in (0, 33, 69):
            want = omd.rotate(dirs[d])
            assert np.abs(got[0][d, :, 0] - want[:, 0]).max() <= 1e-12 * np.abs(want[:, 0]).max()

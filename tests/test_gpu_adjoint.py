"""Differentiation: the device adjoint of rotate -> average -> arg -> Ad * dphi.

Checked against the FP64 oracle's gradient (explicit sums) and against central
finite differences (h = 1e-6) of the device loss; the torch.autograd bridge is
checked the same way.  Bars: loss within 1e-11 relative of the oracle's,
gradients within 1e-7 of the largest gradient (oracle) and 1e-5 (differences).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(leap, n_ant, T, C, D, seed, autos=1, dtype=np.complex128):
    s = leap.SynthMS(n_antennas=n_ant, autocorrelations=autos, n_timesteps=T, n_channels=C, seed=seed)
    s.fill_host(dtype)
    ms = s.host_dict()
    c0 = ms["phase_centre"]
    rng = np.random.default_rng(seed)
    dirs = np.column_stack([c0[0] + rng.uniform(-0.02, 0.02, D), c0[1] + rng.uniform(-0.02, 0.02, D)])
    off = rng.normal(0, 1e-4, (D, 2))
    dp = rng.normal(0, 0.01, (n_ant, 3))
    return s, ms, dirs, off, dp


@pytest.fixture
def adjoint_mode(ctx):
    """Sets the context's adjoint mode for one test and restores the automatic one."""
    def set_mode(mode):
        ctx.set_adjoint_mode(mode)
    yield set_mode
    ctx.set_adjoint_mode(0)


MODES = [0, 1, 2]
MODE_IDS = ["auto", "recompute", "row-sums"]


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n_dirs,n_ch", [(5, 40), (16, 300), (20, 300), (70, 130)])
def test_adjoint_matches_oracle_gradient(leap, ctx, oracle, adjoint_mode, dtype, n_dirs, n_ch, mode):
    # direction lanes per row 16 / 16 / 32 / 128; several LDS channel chunks at 130 and 300 channels;
    # the automatic mode keeps row sums for all of these but 70 directions on 130 complex64 channels
    print(f"adjoint vs oracle: {n_dirs} directions, {n_ch} channels, {np.dtype(dtype).name}, mode {mode}",
          flush=True)
    adjoint_mode(mode)
    s, ms, dirs, off, dp = problem(leap, 12, 4, n_ch, n_dirs, 21, dtype=dtype)
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    L, g_off, g_ant = md.loss_grad(view, dirs, s.enu_to_uvw, off, dp)
    Lo, go_off, go_ant = oracle.loss_grad(ms, dirs, s.enu_to_uvw, off, dp)
    print(f"loss {L:.12e} vs oracle {Lo:.12e}")
    assert abs(L - Lo) <= 1e-11 * abs(Lo)
    assert np.abs(g_off - go_off).max() <= 1e-7 * np.abs(go_off).max()
    assert np.abs(g_ant - go_ant).max() <= 1e-7 * np.abs(go_ant).max()


@pytest.mark.parametrize("mode", [1, 2], ids=["recompute", "row-sums"])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_adjoint_zero_avgdata_matches_oracle(leap, ctx, oracle, adjoint_mode, dtype, mode):
    """Baselines whose every (row, channel) holds a NaN have zero avgData; the loss and gradients
    still follow the oracle's arg(exp(-i theta) avg) for them (a signed-zero phase).  The NaNs
    sit in pol 1, so the whole-channel finiteness test of the adjoint's staging is exercised."""
    adjoint_mode(mode)
    s, ms, dirs, off, dp = problem(leap, 12, 4, 40, 6, 23, dtype=dtype)
    vis = ms["vis"]
    nbl = ms["num_baselines"]
    for b in (3, 17, nbl - 2):
        vis[b::nbl, :, 1] = np.nan  # pol 1 of every channel of every row of baseline b
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    L, g_off, g_ant = md.loss_grad(view, dirs, s.enu_to_uvw, off, dp)
    Lo, go_off, go_ant = oracle.loss_grad(ms, dirs, s.enu_to_uvw, off, dp)
    print(f"loss {L:.12e} vs oracle {Lo:.12e}")
    assert abs(L - Lo) <= 1e-11 * abs(Lo)
    assert np.abs(g_off - go_off).max() <= 1e-7 * np.abs(go_off).max()
    assert np.abs(g_ant - go_ant).max() <= 1e-7 * np.abs(go_ant).max()


def test_adjoint_matches_oracle_gradient_128_stations(leap, ctx, oracle, adjoint_mode):
    """The differentiation workload's array and channel count (128 stations without
    autocorrelations, 8128 baselines, 192 channels, complex64, 16 directions) on a
    4-timestep slice, against the oracle's explicit FP64 gradient, with the adjoint
    from the forward pass's row sums and with the recomputing adjoint."""
    print("adjoint vs oracle: 128 stations, 4 timesteps, 192 channels, 16 directions", flush=True)
    s = leap.SynthMS(n_antennas=128, autocorrelations=0, n_timesteps=4, n_channels=192, seed=4)
    s.fill_host(np.complex64)
    ms = s.host_dict()
    c0 = ms["phase_centre"]
    rng = np.random.default_rng(4)
    dirs = np.column_stack([c0[0] + rng.uniform(-0.02, 0.02, 16), c0[1] + rng.uniform(-0.02, 0.02, 16)])
    off = rng.normal(0, 1e-4, (16, 2))
    dp = rng.normal(0, 0.01, (128, 3))
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    Lo, go_off, go_ant = oracle.loss_grad(ms, dirs, s.enu_to_uvw, off, dp)
    for mode, name in ((2, "row sums"), (1, "recompute")):
        adjoint_mode(mode)
        L, g_off, g_ant = md.loss_grad(view, dirs, s.enu_to_uvw, off, dp)
        e_off = np.abs(g_off - go_off).max() / np.abs(go_off).max()
        e_ant = np.abs(g_ant - go_ant).max() / np.abs(go_ant).max()
        print(f"{name}: loss {L:.12e} vs oracle {Lo:.12e}; gradient errors (of the largest) offsets "
              f"{e_off:.1e}, antennas {e_ant:.1e}")
        assert abs(L - Lo) <= 1e-11 * abs(Lo)
        assert e_off <= 1e-7 and e_ant <= 1e-7


def test_adjoint_matches_central_differences_128_stations(leap, ctx):
    """128 stations without autocorrelations (8128 baselines), 64 t, 192 channels, complex64 generated on the
    device, 16 directions, antenna perturbations N(0, 0.01 m): device gradient vs central differences."""
    s = leap.SynthMS(n_antennas=128, autocorrelations=0, n_timesteps=64, n_channels=192, seed=4)
    assert s.num_baselines == 8128
    vis_p = ctx.malloc(s.num_rows * 192 * 4 * 8)
    uvw_p = ctx.malloc(s.num_rows * 3 * 8)
    try:
        ctx.synth_fill_device(s, leap.LEAP_VIS_C64, vis_p, uvw_p)
        view = leap.view_from_arrays(s.host_dict(), vis_device=vis_p, vis_dtype=leap.LEAP_VIS_C64, uvw_device=uvw_p)
        c0 = (view.phase_centre_ra_rad, view.phase_centre_dec_rad)
        rng = np.random.default_rng(4)
        dirs = np.column_stack([c0[0] + rng.uniform(-0.02, 0.02, 16), c0[1] + rng.uniform(-0.02, 0.02, 16)])
        off = rng.normal(0, 1e-4, (16, 2))
        dp = rng.normal(0, 0.01, (128, 3))
        md = ctx.metadata(view)
        M = s.enu_to_uvw
        L, g_off, g_ant = md.loss_grad(view, dirs, M, off, dp)

        def loss(o, a):
            return md.loss_grad(view, dirs, M, o, a, want_grad=False)[0]

        # The loss is smooth only between branch cuts: arg() in phi1 and dInt, and the wrap of the
        # residuals.  A direction offset of h rad moves the ~1.3e5 baseline phases by ~2 pi |uvw| h /
        # lambda (5e-3 rad at h = 1e-6), which crosses many cuts, so offsets are differenced with
        # h = 1e-9 rad (5e-6 rad of phase); antenna positions with h = 1e-6 m (4e-6 rad of phase).
        worst_off = worst_ant = 0.0
        for d, j in [(0, 0), (0, 1), (7, 0), (15, 1)]:
            h = 1e-9
            e = np.zeros_like(off)
            e[d, j] = h
            fd = (loss(off + e, dp) - loss(off - e, dp)) / (2 * h)
            print(f"offset[{d},{j}]: adjoint {g_off[d, j]:.9e} central difference {fd:.9e}")
            worst_off = max(worst_off, abs(fd - g_off[d, j]) / np.abs(g_off).max())
        for a, k in [(0, 0), (5, 1), (64, 2), (127, 0)]:
            h = 1e-6
            e = np.zeros_like(dp)
            e[a, k] = h
            fd = (loss(off, dp + e) - loss(off, dp - e)) / (2 * h)
            print(f"antenna[{a},{k}]: adjoint {g_ant[a, k]:.9e} central difference {fd:.9e}")
            worst_ant = max(worst_ant, abs(fd - g_ant[a, k]) / np.abs(g_ant).max())
        print(f"loss {L:.6e}; worst |fd - adjoint| / max|grad|: offsets {worst_off:.2e}, antennas {worst_ant:.2e}")
        assert worst_off < 1e-5 and worst_ant < 1e-5
    finally:
        ctx.free(vis_p)
        ctx.free(uvw_p)


def test_torch_autograd_bridge(leap, ctx):
    torch = pytest.importorskip("torch")
    import leap_torch

    s, ms, dirs, off, dp = problem(leap, 10, 3, 24, 3, 9)
    view = leap.view_from_arrays(ms)
    md = ctx.metadata(view)
    prob = leap_torch.CalibrationProblem(md=md, view=view, directions=dirs, enu_to_uvw=s.enu_to_uvw)
    o = torch.tensor(off, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(dp, dtype=torch.float64, requires_grad=True)
    loss = 3.0 * leap_torch.residual_loss(o, a, prob)
    loss.backward()
    _, g_off, g_ant = md.loss_grad(view, dirs, s.enu_to_uvw, off, dp)
    assert np.allclose(o.grad.numpy(), 3.0 * g_off, rtol=1e-12, atol=0)
    assert np.allclose(a.grad.numpy(), 3.0 * g_ant, rtol=1e-12, atol=0)
    # one central difference through the bridge (h small enough to stay between the loss's branch cuts)
    h = 1e-9
    e = torch.zeros_like(o)
    e[1, 0] = h
    with torch.no_grad():
        fd = (leap_torch.residual_loss(o + e, a, prob) - leap_torch.residual_loss(o - e, a, prob)) / (2 * h)
    assert abs(fd.item() - g_off[1, 0]) <= 1e-5 * np.abs(g_off).max()

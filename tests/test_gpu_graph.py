"""Calibration captured as a HIP graph and replayed.

Each replay runs after eager work of the program's own (other calibrations,
synchronisations) and must equal the eager result bit for bit; the graph reads
the current contents of the device buffers, so new visibilities and new
directions of the same shape replay the same graph.  One case is also checked
against the oracle (FP64 <= 1e-9 rad wrapped).
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


class DeviceSet:
    """A synthetic set whose visibilities and uvw live on the device."""

    def __init__(self, leap, ctx, dtype=np.complex128, host_fill=False, **kw):
        self.ctx = ctx
        self.s = leap.SynthMS(**kw)
        self.vd = leap.LEAP_VIS_C64 if dtype == np.complex64 else leap.LEAP_VIS_C128
        self.dtype = dtype
        nbytes = self.s.num_rows * self.s.num_channels * self.s.num_pols * np.dtype(dtype).itemsize
        self.vis_p = ctx.malloc(nbytes)
        self.uvw_p = ctx.malloc(self.s.num_rows * 24)
        self.ms = self.s.host_dict(None)
        if host_fill:
            self.host = self.s.fill_host(dtype)
            ctx.h2d(self.vis_p, self.host)
            ctx.h2d(self.uvw_p, np.ascontiguousarray(self.s.uvw))
        else:
            ctx.synth_fill_device(self.s, self.vd, self.vis_p, self.uvw_p)
        self.view = leap.view_from_arrays(self.ms, vis_device=self.vis_p, vis_dtype=self.vd, uvw_device=self.uvw_p)
        self.md = ctx.metadata(self.view)

    def refill(self, leap, seed):
        """New visibilities of the same shape into the same device buffer."""
        cfg = dict(n_antennas=self.s.cfg.n_antennas, autocorrelations=self.s.cfg.autocorrelations,
                   n_timesteps=self.s.cfg.n_timesteps, n_channels=self.s.cfg.n_channels, seed=seed)
        other = leap.SynthMS(**cfg)
        self.ctx.synth_fill_device(other, self.vd, self.vis_p, 0)

    def free(self):
        self.ctx.free(self.vis_p)
        self.ctx.free(self.uvw_p)


CASES = [  # precision, strategy, direction block, channel sums
    (0, 0, 0, 0), (0, 1, 0, 0), (1, 1, 0, 0), (1, 1, 0, 2), (1, 1, 0, 3), (0, 1, 64, 0),
]


@pytest.mark.parametrize("precision,strategy,db,cs", CASES,
                         ids=["fp64-per-dir", "fp64-batched", "mixed", "mixed-nufft", "mixed-nufft-fp32", "fp64-db64"])
def test_graph_replays_equal_eager(leap, ctx, precision, strategy, db, cs):
    a = DeviceSet(leap, ctx, np.complex64 if precision else np.complex128, n_antennas=24, n_timesteps=9,
                  n_channels=50, seed=81, nan_fraction=0.002)
    b = DeviceSet(leap, ctx, n_antennas=10, n_timesteps=3, n_channels=7, seed=82)
    try:
        dirs = dirs_near(a.ms["phase_centre"], 12, 3)
        slice_ = (0, 3, -1)
        eager = ctx.calibrate(a.view, dirs, slice_, opts=leap.options(precision, strategy, db, channel_sum=cs), md=a.md)
        g = leap.options(precision, strategy, db, graph=1, channel_sum=cs)
        first = ctx.calibrate(a.view, dirs, slice_, opts=g, md=a.md)
        assert ctx.last_graph() == 1
        assert np.array_equal(first[0], eager[0]) and np.array_equal(first[1], eager[1])
        for rep in range(3):
            # eager work of the program's own between replays
            ctx.calibrate(b.view, dirs_near(b.ms["phase_centre"], 2, rep), md=b.md)
            again = ctx.calibrate(a.view, dirs, slice_, opts=g, md=a.md)
            assert ctx.last_graph() == 2
            assert np.array_equal(again[0], eager[0]), f"replay {rep} differs"
        # new contents of the same buffers and new directions replay the same graph
        a.refill(leap, 99)
        dirs2 = dirs_near(a.ms["phase_centre"], 12, 4)
        eager2 = ctx.calibrate(a.view, dirs2, slice_, opts=leap.options(precision, strategy, db, channel_sum=cs), md=a.md)
        replay2 = ctx.calibrate(a.view, dirs2, slice_, opts=g, md=a.md)
        assert ctx.last_graph() == 2
        assert np.array_equal(replay2[0], eager2[0])
        assert not np.array_equal(eager2[0], eager[0])
        # another direction count is another graph
        ctx.calibrate(a.view, dirs2[:5], slice_, opts=g, md=a.md)
        assert ctx.last_graph() == 1
    finally:
        a.free()
        b.free()


@pytest.mark.parametrize("precision,strategy,db,cs", CASES,
                         ids=["fp64-per-dir", "fp64-batched", "mixed", "mixed-nufft", "mixed-nufft-fp32", "fp64-db64"])
def test_graph_capture_first_on_a_fresh_context(leap, precision, strategy, db, cs):
    """A captured calibration as the first call of a new context: every buffer the captured
    sequence uses (the NUFFT's chunked-round scratch included) is reserved before the capture,
    and the result equals the eager one bit for bit."""
    c2 = leap.Context(0)
    a = DeviceSet(leap, c2, np.complex64 if precision else np.complex128, n_antennas=24, n_timesteps=9,
                  n_channels=130, seed=83)
    try:
        dirs = dirs_near(a.ms["phase_centre"], 40, 5)
        slice_ = (0, 3, -1)
        first = c2.calibrate(a.view, dirs, slice_, opts=leap.options(precision, strategy, db, graph=1, channel_sum=cs),
                             md=a.md)
        assert c2.last_graph() == 1
        eager = c2.calibrate(a.view, dirs, slice_, opts=leap.options(precision, strategy, db, channel_sum=cs), md=a.md)
        assert np.array_equal(first[0], eager[0]) and np.array_equal(first[1], eager[1])
    finally:
        a.free()
        c2.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_set_graph_capture_first_on_a_fresh_context(leap, precision):
    c3 = leap.Context(0)
    sets = [DeviceSet(leap, c3, n_antennas=12, autocorrelations=0, n_timesteps=6, n_channels=40, seed=88),
            DeviceSet(leap, c3, np.complex64, n_antennas=20, n_timesteps=4, n_channels=9, seed=89)]
    try:
        arg = [(st.md, st.view, dirs_near(st.ms["phase_centre"], n, n), sl)
               for st, n, sl in zip(sets, (5, 33), ((0, 2, -1), (0, -1, -1)))]
        got = c3.calibrate_batch(arg, leap.options(precision, 1, graph=1))
        assert c3.last_graph() == 1
        eager = c3.calibrate_batch(arg, leap.options(precision, 1))
        for (gc, ge), (ec, ee) in zip(got, eager):
            assert np.array_equal(gc, ec) and np.array_equal(ge, ee)
    finally:
        for st in sets:
            st.free()
        c3.close()


def test_graph_matches_oracle(leap, ctx, oracle):
    a = DeviceSet(leap, ctx, host_fill=True, n_antennas=20, n_timesteps=6, n_channels=33, seed=83)
    try:
        dirs = dirs_near(a.ms["phase_centre"], 9, 5)
        for _ in range(2):
            got, ge = ctx.calibrate(a.view, dirs, (0, 2, -1), opts=leap.options(strategy=1, graph=1), md=a.md)
        assert ctx.last_graph() == 2
        want, we = oracle.calibrate(dict(a.ms, vis=a.host), dirs, (0, 2, -1))
        assert np.array_equal(ge, we)
        assert wrapped(got, want).max() < 1e-9
    finally:
        a.free()


def test_graph_needs_device_resident_data(leap, ctx):
    s = leap.SynthMS(n_antennas=8, n_timesteps=2, n_channels=4, seed=84)
    s.fill_host()
    view = leap.view_from_arrays(s.host_dict())
    with pytest.raises(leap.LeapError, match="device-resident"):
        ctx.calibrate(view, [(0.5, -0.47)], opts=leap.options(strategy=1, graph=1))


def test_multi_set_graph_replays_equal_eager(leap, ctx):
    sets = [DeviceSet(leap, ctx, n_antennas=12, autocorrelations=0, n_timesteps=6, n_channels=40, seed=85),
            DeviceSet(leap, ctx, np.complex64, n_antennas=20, n_timesteps=4, n_channels=9, seed=86),
            DeviceSet(leap, ctx, n_antennas=9, n_timesteps=5, n_channels=300, seed=87)]
    try:
        spec = [(st, dirs_near(st.ms["phase_centre"], n, n), sl)
                for st, n, sl in zip(sets, (5, 70, 1), ((0, 2, -1), (0, -1, -1), (0, 1, -1)))]
        arg = [(st.md, st.view, d, sl) for st, d, sl in spec]
        for precision in (0, 1):
            eager = ctx.calibrate_batch(arg, leap.options(precision, 1))
            ctx.calibrate_batch(arg, leap.options(precision, 1, graph=1))
            assert ctx.last_graph() == 1
            for rep in range(3):
                ctx.calibrate(sets[0].view, spec[0][1], md=sets[0].md)  # eager work in between
                got = ctx.calibrate_batch(arg, leap.options(precision, 1, graph=1))
                assert ctx.last_graph() == 2
                for (gc, ge), (ec, ee) in zip(got, eager):
                    assert np.array_equal(gc, ec) and np.array_equal(ge, ee), f"replay {rep} differs"
            assert ctx.last_batch_order == sorted(ctx.last_batch_order)
    finally:
        for st in sets:
            st.free()

"""The host schedule of the multi-set persistent launch, without a GPU.

Pinned by what any correct static schedule must satisfy: every (set,
baseline, direction block) item appears exactly once, the reported makespan is
the heaviest workgroup's load recomputed from the plan, and greedy list
scheduling's bound (Graham 1966) holds: makespan <= total / m + (1 - 1/m) *
largest item, with makespan >= max(total / m, largest item) for any schedule.
"""
import numpy as np
import pytest

SHAPES = [  # (baselines, timesteps, channels, directions)
    (630, 32, 300, 64),
    (8256, 14, 48, 64),
    (2056, 64, 96, 40),
    (10, 3, 5, 0),  # no directions: no items
    (45, 1, 1, 3),
]
TILES = [321, 641, 1281, 1282, 2562]


def _loads(leap, shapes, tile, off, iset):
    costs = np.array([leap.multi_item_cost(T, C, tile) for (_, T, C, _) in shapes])
    owner = np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))
    return costs, np.bincount(owner, weights=costs[iset], minlength=len(off) - 1)


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("n_blocks", [1, 7, 512, 100000])
def test_schedule_covers_every_item_once(leap, tile, n_blocks):
    off, iset, iloc, makespan = leap.multi_schedule(SHAPES, tile, n_blocks)
    db = tile // 10
    assert off[0] == 0 and off[-1] == len(iset)
    assert np.all(np.diff(off.astype(np.int64)) >= 0)
    for s, (nbl, _, _, D) in enumerate(SHAPES):
        n = nbl * (-(-D // db)) if D else 0
        assert np.array_equal(np.sort(iloc[iset == s]), np.arange(n)), s
    costs, loads = _loads(leap, SHAPES, tile, off, iset)
    assert loads.max() == pytest.approx(makespan, rel=1e-12)
    total = loads.sum()
    biggest = max(costs[s] for s, sh in enumerate(SHAPES) if sh[3] > 0)
    m = n_blocks
    assert makespan <= total / m + (1 - 1 / m) * biggest + 1e-6
    assert makespan >= max(total / m, biggest) - 1e-6


def test_schedule_is_deterministic_and_balanced(leap):
    a = leap.multi_schedule(SHAPES, 1282, 1024)
    b = leap.multi_schedule(SHAPES, 1282, 1024)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    costs, loads = _loads(leap, SHAPES, 1282, a[0], a[1])
    # many items per workgroup: within one largest item of perfect balance
    assert loads.max() - loads.min() <= costs.max() + 1e-9


def test_item_cost_model(leap):
    # more timesteps or channels never cost less; unsupported tiles are rejected
    for tile in TILES:
        c = [leap.multi_item_cost(T, C, tile) for T in (1, 8, 64) for C in (1, 96, 384)]
        grid = np.array(c).reshape(3, 3)
        assert np.all(np.diff(grid, axis=0) >= 0) and np.all(np.diff(grid, axis=1) >= 0)
    assert leap.multi_item_cost(8, 8, 642) < 0
    with pytest.raises(leap.LeapError, match="unsupported multi-set tile"):
        leap.multi_schedule(SHAPES, 642, 4)
    with pytest.raises(leap.LeapError, match="at least one workgroup"):
        leap.multi_schedule(SHAPES, 641, 0)

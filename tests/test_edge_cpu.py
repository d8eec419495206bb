"""Edge cases of the host half (no GPU): a single body, bodies fewer than a
group, sizes straddling group boundaries, many coincident bodies (the tree
stops at depth 21 and a leaf holds more than group_max bodies), and the input
checks (non-positive or non-finite masses, non-finite positions)."""
import numpy as np
import pytest

import listoracle
from pgrav import batch
from pgrav._lib import PgravError


def _cover(H, n):
    st = H.stats()
    for g in range(st["ngroups"]):
        assert listoracle.interactions_cover_all(H, g, n)
    first, cnt, _ = H.groups()
    assert cnt.sum() == n and first[0] == 0 and np.all(first[1:] == first[:-1] + cnt[:-1])
    assert cnt.max() <= 64
    return st


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129])
def test_small_and_ragged_sizes(n):
    rng = np.random.default_rng(n)
    pos = rng.uniform(size=(n, 3))
    H = batch.HostPlan(pos, np.full(n, 1.0 / n), theta=0.6, order=4, eps=1e-3)
    st = _cover(H, n)
    if n == 1:
        assert st["ngroups"] == 1 and st["entries_body"] == 1


def test_coincident_bodies():
    rng = np.random.default_rng(7)
    pos = np.concatenate([np.full((300, 3), 0.25), rng.uniform(size=(40, 3))])
    mass = np.full(pos.shape[0], 1.0 / pos.shape[0])
    H = batch.HostPlan(pos, mass, theta=0.6, order=4, eps=1e-3, leaf_max=16)
    _cover(H, pos.shape[0])


@pytest.mark.parametrize("bad", ["zero_mass", "negative_mass", "nan_mass", "nan_pos", "inf_pos"])
def test_bad_bodies_are_rejected(bad):
    rng = np.random.default_rng(1)
    pos = rng.uniform(size=(50, 3))
    mass = np.full(50, 0.02)
    if bad == "zero_mass":
        mass[3] = 0.0
    elif bad == "negative_mass":
        mass[3] = -1.0
    elif bad == "nan_mass":
        mass[3] = np.nan
    elif bad == "nan_pos":
        pos[3, 1] = np.nan
    else:
        pos[3, 2] = np.inf
    with pytest.raises(PgravError):
        batch.HostPlan(pos, mass)

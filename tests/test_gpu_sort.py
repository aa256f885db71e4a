"""The spatial order on the device (sort.hip): a hand-written stable LSD radix sort of the
cells' Morton keys and a prefix scan numbering the occupied cells.

Checked on the host: the order is a permutation, rank is its inverse, the keys are the
particles' classes (sites, fixed ions, moving ions) over the Morton codes of their 8 A cells
(10 bits per axis, interleaved x, y, z) and non-decreasing, equal keys keep index order (the order equals numpy's stable argsort of the
same keys), and cell_start marks exactly the slots where the key changes.  Sizes from one
particle to the large translocation system, with particles on and outside the box edges
(clamped to the edge cells) and many particles sharing cells.
"""
import numpy as np
import pytest

from bromoc_amd import configs
from bromoc_amd.engine import Engine

from helpers import plain, pore_small, with_fixed

pytestmark = pytest.mark.gpu


def morton(pos, box, nsites=0, nfix=0, cell=8.0):
    lo = np.array([box[0], box[2], box[4]])
    hi = np.array([box[1], box[3], box[5]])
    n = np.clip(np.ceil((hi - lo) / cell).astype(int), 1, 1024)
    c = (pos - lo) / cell
    k = np.where(c > 0.0, np.floor(c), 0).astype(np.int64)
    k = np.minimum(k, n - 1)

    def spread(v):
        out = np.zeros_like(v)
        for b in range(10):
            out |= ((v >> b) & 1) << (3 * b)
        return out
    i = np.arange(len(pos))
    cls = np.where(i < nsites, 0, np.where(i < nsites + nfix, 1, 2)).astype(np.int64)
    return ((cls << 30) | spread(k[:, 0]) | (spread(k[:, 1]) << 1) |
            (spread(k[:, 2]) << 2)).astype(np.uint32)


def check(e, s, pos):
    keys, order, rank, cs = e.spatial_order()
    n = len(pos)
    assert np.array_equal(np.sort(order), np.arange(n))
    assert np.array_equal(rank[order], np.arange(n))
    want = morton(pos, s.box, s.nsites, s.nfix)
    assert np.array_equal(keys, want[order])
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert np.array_equal(order, np.argsort(want, kind="stable"))
    change = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
    assert np.array_equal(cs[:-1], change) and cs[-1] == n
    return len(change)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4096, 4097])
def test_random_positions(n):
    s = plain(seed=1, n_per_type=max(1, (n + 1) // 2))
    e = Engine(s)
    rng = np.random.default_rng(n)
    half = 0.5 * (s.box[1] - s.box[0])
    pos = rng.uniform(-1.2 * half, 1.2 * half, (s.ntot, 3))     # some outside the box
    pos[: s.ntot // 4] = np.round(pos[: s.ntot // 4] / 8.0) * 8.0   # on cell edges
    pos[s.ntot // 4: s.ntot // 2] = pos[0]                         # many in one cell
    e.set_positions(pos)
    ncell = check(e, s, pos)
    assert ncell <= s.ntot


def test_pore_dna_and_translocation():
    for s in (pore_small(seed=3), with_fixed(pore_small(seed=4), 12), configs.pore_dna(seed=1),
              configs.translocation()):
        e = Engine(s)
        pos = np.stack([s.x, s.y, s.z], 1)[: s.ntot]
        ncell = check(e, s, pos)
        print(f"{s.ntot} particles in {ncell} occupied cells")
        e.close()

"""EFPOT table lookups on knot boundaries (getyd, polynom.f90:225-226) on the HIP path.

Two ions are placed exactly on knot distances r_k and one ulp either side.  The table jumps
by ~1e-9 kcal/mol at every knot (fixdy, polynom.f90:91-118), so a lookup that takes the
neighbouring segment -- as an index formed from a 1/sqrt that is one ulp off would -- misses
the oracle's pair energy by far more than the 1e-12 relative allowed here; the C oracle
itself is bit-identical to the reference on the same placements (test_oracle_vs_reference).
Checked through every launch shape of the step kernels and in the batch kernel.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Batch, Engine
from bromoc_amd.system import Replica, replica_system
from oracle.oracle import Oracle

from helpers import force_err

pytestmark = pytest.mark.gpu


def knot_positions(s, every=17, rmin=2.5):
    r = configs.knot_grid()
    out = []
    for k in range(1, len(r) - 1, every):
        if r[k] < rmin:
            continue
        for x in (r[k], np.nextafter(r[k], 0.0), np.nextafter(r[k], 1e9)):
            out.append(np.array([[0.0, 0.0, 0.0], [float(x), 0.0, 0.0]]))
    return out


@pytest.mark.parametrize("tpb", [1, 2, 4, 8])
def test_knot_boundaries(tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = configs.plain_ions(seed=3, n_per_type=1, box=40.0)
    g, o = Engine(s), Oracle(s)
    bad = []
    for pos in knot_positions(s):
        g.set_positions(pos)
        o.set_positions(pos)
        eg, eo = g.energy(), o.energy()
        if abs(eg["eefpot"] - eo["eefpot"]) > 1e-12 * abs(eo["eefpot"]):
            bad.append((pos[1, 0], eg["eefpot"], eo["eefpot"]))
        assert force_err(g.forces(), o.forces()) <= 1e-12
    assert not bad, bad[:5]


def test_knot_boundaries_batch():
    """The batch step kernel's table lookup on the same placements, one replica per
    placement, energies after ENERGY-only evaluation."""
    s = configs.plain_ions(seed=3, n_per_type=1, box=40.0)
    sh, _ = configs.replica_batch(seed=1, nrep=1, n_range=(2, 2), box=(36.0, 36.0, 72.0),
                                  dcel=1.5, rpore=9.0, half_slab=10.0, gcmc=False)
    places = knot_positions(s, every=41)
    reps = []
    for pos in places:
        pos = pos + np.array([0.0, 0.0, 20.0])      # in the bath, clear of the pore walls
        reps.append(Replica(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
                            typei=np.array([sh.nold + 1, sh.nold + 2], np.int32), nfix=0,
                            capacity=8, phix=np.zeros(sh.phix.data.shape, np.float32)))
    b = Batch(sh, reps)
    eb = b.energy()
    bad = []
    for r, rep in enumerate(reps):
        eo = Oracle(replica_system(sh, rep)).energy()
        if abs(eb[r]["eefpot"] - eo["eefpot"]) > 1e-12 * abs(eo["eefpot"]):
            bad.append((r, eb[r]["eefpot"], eo["eefpot"]))
    assert not bad, bad[:5]

"""Replica batches on the GPU against the C oracle run one replica at a time.

Replica r of a batch is the single system of its own particles and static field over the
shared tables, with Philox stream r for its Gaussians and its GCMC uniforms drawn as
oracle.batch_uniforms describes.  Bar: positions 1e-8 A after tens of steps, energies and
forces 1e-10 relative, integer state (ntot, nat, prob, ibuffer, flux) exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Batch, Engine
from bromoc_amd.system import replica_system
from oracle.oracle import Oracle, batch_uniforms

from helpers import energy_errs, force_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def small_batch(seed=1, nrep=5, n_range=(16, 300), gcmc=False, **kw):
    return configs.replica_batch(seed=seed, nrep=nrep, n_range=n_range, box=(36.0, 36.0, 72.0),
                                 dcel=1.5, rpore=9.0, half_slab=10.0, gcmc=gcmc, **kw)


def _check_energy(eg, eo):
    errs = energy_errs(eg, eo)
    bad = {k: v for k, v in errs.items() if v > TOL}
    assert not bad, bad


def test_batch_energy_matches_single_systems():
    sh, reps = small_batch(seed=2)
    reps[1].nfix = 5
    b = Batch(sh, reps)
    eb = b.energy()
    for r, rep in enumerate(reps):
        o = Oracle(replica_system(sh, rep))
        eo = o.energy()
        _check_energy(eb[r], eo)
        assert force_err(b.forces(r)[rep.nfix:], o.forces()[rep.nfix:]) <= TOL, r


@pytest.mark.parametrize("shape", ["auto", "0/64", "512/64", "0/16", "256/32"])
def test_batch_steps_match_single_systems(shape, monkeypatch):
    if shape != "auto":
        blocks, lanes = shape.split("/")
        monkeypatch.setenv("BROMOC_BATCH_BLOCKS", blocks)
        monkeypatch.setenv("BROMOC_BATCH_LANES", lanes)
    sh, reps = small_batch(seed=3)
    reps[2].nfix = 4
    b = Batch(sh, reps)
    print("batch launch shape (blocks, lanes):", b.tuning())
    nsteps, seed, c0 = 40, 17, 100
    eb = b.run(nsteps, seed=seed, counter0=c0, energies=True)
    for r, rep in enumerate(reps):
        o = Oracle(replica_system(sh, rep), stream=r)
        eo = o.step(nsteps, abi.NOISE_PHILOX, seed=seed, counter0=c0)
        pos, typ, n, _, _, _ = b.get(r)
        err = float(np.max(np.abs(pos - o.positions()[0])))
        assert n == rep.ntot and err <= 1e-8, (r, err)
        _check_energy(eb[r], eo)
        for a, c in zip(b.flux(r), o.flux()):
            assert np.array_equal(a, c), r


def test_single_replica_equals_plain_path():
    sh, reps = small_batch(seed=4, nrep=1, n_range=(150, 150))
    b = Batch(sh, reps)
    e = Engine(replica_system(sh, reps[0]))
    b.run(37, seed=5, counter0=9)
    e.step(37, abi.NOISE_PHILOX, seed=5, counter0=9)
    err = float(np.max(np.abs(b.get(0)[0] - e.positions()[0])))
    assert err <= 1e-9, err


def test_batch_gcmc_matches_oracle():
    sh, reps = small_batch(seed=6, nrep=3, n_range=(40, 160), gcmc=True, conc_range=(0.5, 2.0))
    b = Batch(sh, reps)
    seed, c0, cyc0, every, ngcmc, ncyc = 11, 3, 2, 5, 12, 4
    b.run(every * ncyc, seed=seed, counter0=c0, gcmc_every=every, ngcmc=ngcmc, cycle0=cyc0)
    changed = 0
    for r, rep in enumerate(reps):
        s = replica_system(sh, rep)
        o = Oracle(s, stream=r)
        o.set_gcmc(rep.gcmc, np.concatenate([rep.ibuffer, np.zeros(s.capacity - rep.ntot, np.int32)]))
        nat = rep.nat.copy()
        prob = np.zeros((rep.gcmc.nbuffer, s.capacity + 1), np.int32)
        ener = 0.0
        for c in range(ncyc):
            u = batch_uniforms(seed, cyc0 + c, r, 6 * ngcmc + 8)
            _, ener = o.grand(ngcmc, u, nat, prob, cyc0 + c + 1, ener)
            o.step(every, abi.NOISE_PHILOX, seed=seed, counter0=c0 + every * c)
        pos, typ, n, ib, natg, eg = b.get(r)
        po, to = o.positions()
        changed += n != rep.ntot
        assert n == o.ntot, (r, n, o.ntot)
        assert np.array_equal(natg, nat), r
        assert np.array_equal(b.prob(r), prob), r
        assert np.array_equal(typ, to[:n]) and np.array_equal(ib, o.ibuffer()[:n]), r
        err = float(np.max(np.abs(pos - po[:n])))
        assert err <= 1e-8, (r, err)
        assert abs(eg - ener) <= TOL * max(abs(ener), 1.0), (r, eg, ener)
    assert changed > 0      # GRAND did insert or delete somewhere


def test_batch_capacity_error():
    sh, reps = small_batch(seed=7, nrep=2, n_range=(20, 30), gcmc=True, conc_range=(5.0, 6.0),
                           mu=(5.0, 5.0), cap_factor=1.0)
    for rep in reps:
        rep.capacity = rep.ntot + 1
    b = Batch(sh, reps)
    with pytest.raises(abi.BromocError) as ei:
        b.run(20, seed=1, gcmc_every=5, ngcmc=50)
    assert ei.value.code == abi.E_CAPACITY

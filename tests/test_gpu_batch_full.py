"""The replica-batch workload at full size: 256 ion + pore replicas of log-uniform 100..4000
ions (concentration 0.1..2 M, voltage -0.5..0.5 V), each with its own 129 x 129 x 257 static
field at 0.75 A over the shared repulsion map, GCMC bath buffers at both ends; 20 BD steps
with GRAND (20 attempts) every 10, through the persistent work-queue kernel and captured
graphs the autotuner picks.

A stratified sample of replicas -- the smallest, the largest and six between -- is replayed
one at a time by the C oracle with the same Philox stream and GCMC uniforms.  Bar: positions
1e-8 A, energies 1e-10 relative, particle count, nat, prob, ibuffer, types and flux exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Batch
from bromoc_amd.system import replica_system
from oracle.oracle import Oracle, batch_uniforms

from helpers import energy_errs

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def batch_run():
    sh, reps = configs.replica_batch(seed=3)
    assert len(reps) == 256 and sh.phix.data.shape == (129, 129, 257)
    b = Batch(sh, reps)
    seed, c0, cyc0, every, ngcmc, nsteps = 7, 0, 0, 10, 20, 20
    eb = b.run(nsteps, seed=seed, counter0=c0, gcmc_every=every, ngcmc=ngcmc, cycle0=cyc0,
               energies=True)
    print("launch shape (blocks, lanes):", b.tuning(), flush=True)
    yield sh, reps, b, eb, (seed, c0, cyc0, every, ngcmc, nsteps)
    b.close()


def sample(reps):
    order = np.argsort([r.ntot for r in reps])
    picks = [int(order[0]), int(order[-1])]
    picks += [int(order[k]) for k in np.linspace(0, len(reps) - 1, 8)[1:-1].round().astype(int)]
    return sorted(set(picks))


def test_full_batch_sample_matches_oracle(batch_run):
    sh, reps, b, eb, (seed, c0, cyc0, every, ngcmc, nsteps) = batch_run
    changed = 0
    for r in sample(reps):
        rep = reps[r]
        s = replica_system(sh, rep)
        o = Oracle(s, stream=r)
        o.set_gcmc(rep.gcmc, np.concatenate([rep.ibuffer, np.zeros(s.capacity - rep.ntot, np.int32)]))
        nat = rep.nat.copy()
        prob = np.zeros((rep.gcmc.nbuffer, s.capacity + 1), np.int32)
        ener, eo = 0.0, None
        for c in range(nsteps // every):
            u = batch_uniforms(seed, cyc0 + c, r, 6 * ngcmc + 8)
            _, ener = o.grand(ngcmc, u, nat, prob, cyc0 + c + 1, ener)
            eo = o.step(every, abi.NOISE_PHILOX, seed=seed, counter0=c0 + every * c)
        pos, typ, n, ib, natg, eg = b.get(r)
        po, to = o.positions()
        print(f"replica {r}: {rep.ntot} -> {n} ions", flush=True)
        changed += n != rep.ntot
        assert n == o.ntot, (r, n, o.ntot)
        assert np.array_equal(natg, nat), r
        assert np.array_equal(b.prob(r), prob), r
        assert np.array_equal(typ, to[:n]) and np.array_equal(ib, o.ibuffer()[:n]), r
        err = float(np.max(np.abs(pos - po[:n])))
        assert err <= 1e-8, (r, err)
        assert abs(eg - ener) <= TOL * max(abs(ener), 1.0), (r, eg, ener)
        bad = {k: v for k, v in energy_errs(eb[r], eo).items() if v > TOL}
        assert not bad, (r, bad)
        for a, c in zip(b.flux(r), o.flux()):
            assert np.array_equal(a, c), r
    assert changed > 0

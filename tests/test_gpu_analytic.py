"""The analytic pair branch on the HIP path against the C oracle (which is pinned against the
reference routines and closed forms in test_oracle_analytic.py): Coulomb, 12-6 LJ and SRPMF
pair types without EFPOT tables (energy.mono.f90:135-178,623-668; interact.f90:182-204,
365-404), all-analytic and mixed with tabulated types, ion-ion and site-ion, through every
launch shape of the step kernels (BROMOC_TPB 1, 2, 4 targets per block; 8, one target per
wavefront), in BD trajectories (5 steps launched directly, 40 steps replayed from captured
graphs), in INTERACT and GRAND, and in replica batches (the batch step kernel's analytic
path and its device GRAND).

Also on the GPU here: the CG-DNA options the other GPU tests leave at their defaults -- base
pairing scaled by scalepairing != 1 (energy.mono.f90:372-376; INTERACT leaves it out,
interact.f90:288-298), plain Debye screening of the P-P pairs (QDEBY,
energy.mono.f90:439-447) and per-type repulsion scales (QSVDW, vdwgd1trln.f90:139).

Bar: energies and norm-wise forces 1e-10 relative, positions 1e-9 A after 5 steps and
1e-8 A after 40, INTERACT 1e-10 relative, integer state exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Batch, Engine
from bromoc_amd.system import replica_system
from oracle.oracle import Oracle, batch_uniforms

from helpers import energy_errs, force_err, near_wall, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def dense_plain(seed, n=40, box=24.0):
    return configs.plain_ions(seed=seed, n_per_type=n, box=box)


def variant(name):
    if name == "all":
        return with_fixed(configs.with_analytic(dense_plain(1), seed=11), 6)
    if name == "mixed":
        return configs.with_analytic(dense_plain(2), types=[1, 3], seed=12)
    if name == "lj":
        return configs.with_analytic(dense_plain(3), seed=13, srpmf=False, coulomb=False)
    s = pore_small(seed=5, n_bp=6, n_ions=150)
    if name == "site":
        t = configs.with_analytic(s, types=configs.analytic_pair_types(s, "site"), seed=15)
        return with_fixed(t, 5, nonfree=(0, 4, 9))
    if name == "pore_all":
        return with_fixed(configs.with_analytic(s, seed=16), 3)
    raise KeyError(name)


def _check_energy(eg, eo):
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
    assert not bad, bad


@pytest.mark.parametrize("tpb", [1, 2, 4, 8])
@pytest.mark.parametrize("name", ["all", "mixed", "lj", "site", "pore_all"])
def test_energy_forces(name, tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = variant(name)
    g, o = Engine(s), Oracle(s)
    assert g.tuning() == tpb
    eg, eo = g.energy(), o.energy()
    assert eo["esrpmf"] != 0.0 or eo["esrpmfmx"] != 0.0 or name == "lj"
    _check_energy(eg, eo)
    assert force_err(g.forces(), o.forces()) <= TOL


@pytest.mark.parametrize("tpb", [1, 2, 4, 8])
@pytest.mark.parametrize("name", ["all", "mixed", "site"])
def test_steps_and_graphs(name, tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = variant(name)
    noise = np.random.default_rng(3).standard_normal((40, 3 * s.ntot))
    for nsteps, tol in ((5, 1e-9), (40, 1e-8)):
        g, o = Engine(s), Oracle(s)
        g.upload_noise(noise)
        o.upload_noise(noise)
        eg = g.step(nsteps, abi.NOISE_INJECTED)
        eo = o.step(nsteps, abi.NOISE_INJECTED)
        err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
        print(f"{name} tpb={tpb} {nsteps} steps: position err {err:.3e}")
        assert err <= tol
        _check_energy(eg, eo)
        for a, b in zip(g.flux(), o.flux()):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["all", "mixed", "site", "pore_all"])
def test_interact(name):
    s = variant(name)
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    cases = [(pos[j] + 0.25, abs(int(typ[j])) if j >= s.nsites else 0, j)
             for j in range(0, s.ntot, 5)]
    cases += [((0.5, -1.0, 2.0), jt, s.ntot) for jt in range(s.nold + 1, s.ntype + 1)]
    for xyz, jt, j in cases:
        a, b = g.interact(xyz, jt, j, True), o.interact(xyz, jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (name, j, a, b)
    assert np.array_equal(g.warn(), o.warn())


def test_grand_analytic():
    """GRAND (grand.f90:40-118) weighs every trial by INTERACT's analytic branch."""
    s = pore_small(seed=64, n_bp=0, n_ions=120, dna=False)
    s = configs.with_analytic(s, seed=17)
    s.capacity = s.ntot + 200
    gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    g, o = Engine(s), Oracle(s)
    g.set_gcmc(gc, ib)
    o.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(20000)
    natg, nato = nat.copy(), nat.copy()
    probg = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    probo = probg.copy()
    ug, eg = g.grand(200, u, natg, probg, 1, 0.0)
    uo, eo = o.grand(200, u, nato, probo, 1, 0.0)
    assert ug == uo and g.ntot == o.ntot
    assert np.array_equal(natg, nato) and np.array_equal(probg, probo)
    assert abs(eg - eo) <= TOL * max(abs(eo), 1.0)
    assert np.array_equal(g.positions()[0], o.positions()[0])


def _analytic_batch(seed, nrep=4, gcmc=False, **kw):
    sh, reps = configs.replica_batch(seed=seed, nrep=nrep, n_range=(30, 260),
                                     box=(36.0, 36.0, 72.0), dcel=1.5, rpore=9.0,
                                     half_slab=10.0, gcmc=gcmc, **kw)
    return configs.with_analytic(sh, types=[configs.analytic_pair_types(sh, "ion")[1]],
                                 seed=seed), reps


@pytest.mark.parametrize("shape", ["auto", "0/64", "256/16"])
def test_batch_analytic_steps(shape, monkeypatch):
    if shape != "auto":
        blocks, lanes = shape.split("/")
        monkeypatch.setenv("BROMOC_BATCH_BLOCKS", blocks)
        monkeypatch.setenv("BROMOC_BATCH_LANES", lanes)
    sh, reps = _analytic_batch(31)
    reps[1].nfix = 3
    b = Batch(sh, reps)
    nsteps, seed, c0 = 30, 19, 50
    eb = b.run(nsteps, seed=seed, counter0=c0, energies=True)
    for r, rep in enumerate(reps):
        o = Oracle(replica_system(sh, rep), stream=r)
        eo = o.step(nsteps, abi.NOISE_PHILOX, seed=seed, counter0=c0)
        pos, _, n, _, _, _ = b.get(r)
        err = float(np.max(np.abs(pos - o.positions()[0])))
        assert n == rep.ntot and err <= 1e-8, (r, err)
        _check_energy(eb[r], eo)


def test_batch_analytic_gcmc():
    sh, reps = _analytic_batch(32, nrep=3, gcmc=True, conc_range=(0.5, 2.0))
    b = Batch(sh, reps)
    seed, c0, cyc0, every, ngcmc, ncyc = 13, 3, 2, 5, 12, 3
    b.run(every * ncyc, seed=seed, counter0=c0, gcmc_every=every, ngcmc=ngcmc, cycle0=cyc0)
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
        assert n == o.ntot and np.array_equal(natg, nat) and np.array_equal(b.prob(r), prob)
        assert np.array_equal(typ, to[:n])
        assert float(np.max(np.abs(pos - po[:n]))) <= 1e-8
        assert abs(eg - ener) <= TOL * max(abs(ener), 1.0)


# ------------------------------------------------- CG-DNA options off their defaults
def dna_options(seed):
    s = with_fixed(near_wall(pore_small(seed=seed, n_bp=8, n_ions=150), 20, seed=seed), 4,
                   nonfree=(2,))
    s.flags |= abi.QDEBY | abi.QSVDW
    s.kappa, s.ikappa = 9.6, 1.0 / 9.6
    s.scal = np.linspace(0.5, 1.5, s.nttyp + 1)
    s.scalepairing = 1.7
    return s


@pytest.mark.parametrize("tpb", [1, 2, 4, 8])
def test_dna_options_energy_forces(tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = dna_options(71)
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    assert eo["ebp"] != 0.0 and eo["eqq"] != 0.0 and eo["evdwgd"] != 0.0
    _check_energy(eg, eo)
    assert force_err(g.forces(), o.forces()) <= TOL


def test_dna_options_steps_and_interact():
    s = dna_options(72)
    noise = np.random.default_rng(8).standard_normal((40, 3 * s.ntot))
    for nsteps, tol in ((5, 1e-9), (40, 1e-8)):
        g, o = Engine(s), Oracle(s)
        g.upload_noise(noise)
        o.upload_noise(noise)
        g.step(nsteps, abi.NOISE_INJECTED)
        o.step(nsteps, abi.NOISE_INJECTED)
        assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= tol
        assert np.array_equal(g.warn(), o.warn())
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    for j in list(range(0, s.nsites, 2)) + list(range(s.nsites, s.ntot, 13)):
        jt = abs(int(typ[j])) if j >= s.nsites else 0
        a, b = g.interact(pos[j], jt, j, True), o.interact(pos[j], jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)

"""Spherical and elliptic-cylinder systems on the device against the C oracle: FIXCOOR after
each BD step (k_curved, every launch shape, hipGraph chunks included), OUTBOX in the hybrid
Debye screening (also inside the one-target step kernel's CG-DNA terms), INSERT's redraws
in the device GRAND and MOVE's test in METROPOLIS.

Bar: positions 1e-9 A after BD steps, energies 1e-10 relative (norm-wise below 1e-6 of the
largest term), GCMC and METROPOLIS decisions, counts and buffers exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import energy_errs, force_err, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _plain():
    return configs.plain_ions(seed=4, n_per_type=150)


def _pore(seed=7):
    return configs.pore_dna(seed=seed, n_bp=6, n_ions=150, box=(40.0, 40.0, 60.0), dcel=0.5,
                            rpore=12.0, half_slab=10.0, dna=True)


def _shaped(which, base):
    if which == "sphere":
        return configs.with_sphere(base, 0.5 * (base.box[1] - base.box[0]), seed=2)
    return configs.with_ecyl(base, seed=2)


def _inside(s, pos):
    c = np.array(s.center)
    mv = pos[s.nsites + s.nfix:]
    if s.flags & abi.QSPHERE:
        return bool(np.all(np.linalg.norm(mv - c, axis=1) <= s.rsphe * (1 + 1e-15)))
    return bool(np.all(np.hypot((mv[:, 0] - c[0]) * s.iecx, (mv[:, 1] - c[1]) * s.iecy) <= 1.0))


@pytest.mark.parametrize("tpb", [1, 2, 4, 8, 17])
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
@pytest.mark.parametrize("base", ["plain", "pore", "pore_notrans"])
def test_steps(base, which, tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = _shaped(which, _plain() if base == "plain" else with_fixed(_pore(), 5, nonfree=(2,)))
    if base == "pore_notrans":   # the centroid reset runs after the curved FIXCOOR
        s = configs.with_notrans(s, axes=7)
    nsteps = 6
    noise = 3.0 * np.random.default_rng(3).standard_normal((nsteps, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    eg = g.step(nsteps, abi.NOISE_INJECTED)
    eo = o.step(nsteps, abi.NOISE_INJECTED)
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
    assert not bad, bad
    pg = g.positions()[0]
    err = float(np.max(np.abs(pg - o.positions()[0])))
    print(f"{base} {which} tpb={tpb}: {nsteps}-step position err {err:.3e}")
    assert err <= 1e-9
    assert _inside(s, pg)
    for a, b in zip(g.flux(), o.flux()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_graph_chunks_philox(which):
    """80 steps in hipGraph chunks with device Philox noise, against the oracle's Philox."""
    s = _shaped(which, _pore(seed=8))
    g, o = Engine(s), Oracle(s)
    eg = g.step(80, abi.NOISE_PHILOX, 0, seed=99)
    eo = o.step(80, abi.NOISE_PHILOX, 0, seed=99)
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > 1e-9}
    assert not bad, bad
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    print(f"{which}: 80 graph steps, position err {err:.3e}")
    assert err <= 1e-8
    assert _inside(s, g.positions()[0])


@pytest.mark.parametrize("tpb", [1, 8, 17])
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_debyhyb_outbox(which, tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = configs.with_debyhyb(_shaped(which, _pore(seed=9)))
    s.x, s.y, s.z = s.x.copy(), s.y.copy(), s.z.copy()
    pk = np.flatnonzero(np.asarray(s.namsite) == abi.SITE_P)
    for k, i in enumerate(pk):
        s.x[i], s.y[i], s.z[i] = 18.0 - 0.05 * k, 18.0, -14.0 + 28.0 * k / len(pk)
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
    assert not bad, bad
    assert force_err(g.forces(), o.forces()) <= TOL
    # then steps: the one-target kernel evaluates the P-P terms inside its site blocks
    noise = np.random.default_rng(5).standard_normal((4, 3 * s.ntot))
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(4, abi.NOISE_INJECTED)
    o.step(4, abi.NOISE_INJECTED)
    assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-9


@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_grand(which):
    base = configs.pore_dna(seed=10, n_bp=0, n_ions=120, box=(40.0, 40.0, 60.0), dcel=0.5,
                            rpore=12.0, half_slab=10.0, dna=False)
    s = _shaped(which, base).copy()
    s.capacity = s.ntot + 200
    if which == "sphere":
        gc = configs.sphere_buffers(s, conc=1.0, shell_len=6.0, mu=(-0.5, -0.4))
    else:
        gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    g, o = Engine(s), Oracle(s)
    g.set_gcmc(gc, ib)
    o.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(40000)
    natg, nato = nat.copy(), nat.copy()
    probg = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    probo = probg.copy()
    ug, eg = g.grand(200, u, natg, probg, 1, 0.0)
    uo, eo = o.grand(200, u, nato, probo, 1, 0.0)
    assert ug == uo and ug > 6 * 200 * 0.5
    assert np.array_equal(natg, nato) and np.array_equal(probg, probo)
    assert g.ntot == o.ntot
    assert abs(eg - eo) <= TOL * max(abs(eo), 1.0)
    pg, tg = g.positions()
    po, to = o.positions()
    assert np.array_equal(pg, po) and np.array_equal(tg, to)
    assert np.array_equal(g.ibuffer(), o.ibuffer())
    # BD steps after the insertions and deletions
    noise = np.random.default_rng(8).standard_normal((10, 3 * g.ntot))
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(10, abi.NOISE_INJECTED)
    o.step(10, abi.NOISE_INJECTED)
    assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-9


@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_metropolis(which):
    s = _shaped(which, with_fixed(_pore(seed=36), 6))
    g, o = Engine(s), Oracle(s)
    u = np.random.default_rng(9).random(40000)
    ikbt = 1.0 / (0.0019872 * 300.0)
    ug, dg, eg = g.metropolis(300, 6.0, ikbt, u, 1.5)
    uo, do, eo = o.metropolis(300, 6.0, ikbt, u, 1.5)
    assert dg == do == 300 and ug == uo
    assert abs(eg - eo) <= TOL * max(abs(eo), 1.0)
    assert np.array_equal(g.positions()[0], o.positions()[0])

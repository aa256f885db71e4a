"""GRAND of one system on the device (grand.hip: trial, FIND, multi-workgroup INTERACT and
the acceptance test run back to back on the stream, one host round trip per call) against
the C oracle's GRAND (grand.f90:40-118), itself pinned to the reference's routine.

Bar: the same uniforms consumed, the same acceptance sequence (nat, prob, particle count,
positions, types and buffers of every slot exact), GRAND's energy 1e-10 relative, and BD
steps on the resulting particle set 1e-9 A.  Systems: ions in a pore, with CG-DNA (site
partners), with the reaction-field grids (RFPARIONJ and its refreshed srfe / reff after every
accepted move), with MEMBRANE, with analytic pair types, with biased buffers; an exhausted
uniform stream and a full capacity are reported as errors.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import plain, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _gcmc(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4), bias=False):
    gc = configs.gcmc_buffers(s, conc=conc, buffer_len=buffer_len, mu=mu)
    if bias:
        gc.bufferbias[1:] = 1
        gc.kb[1:] = np.linspace(0.2, 0.6, gc.nbuffer)
        gc.ntotat[1:] = gc.ntotat[1:] * 3 + 7
    return gc


def _systems():
    return {
        "pore": pore_small(seed=64, n_bp=0, n_ions=200, dna=False),
        "pore_dna": with_fixed(pore_small(seed=65, n_bp=6, n_ions=200), 4),
        "rfpar": configs.with_rfpar(pore_small(seed=66, n_bp=0, n_ions=200, dna=False), seed=3,
                                    shape=(37, 37, 55), dcel=1.0, wall_radius=12.0,
                                    zero_frac=0.05),
        "membrane": configs.with_membrane(pore_small(seed=67, n_bp=0, n_ions=200, dna=False),
                                          thick=16.0, pore=(10.0, 6.0, 2.0)),
        "analytic": configs.with_analytic(pore_small(seed=68, n_bp=6, n_ions=200), seed=4),
    }


@pytest.mark.parametrize("bias", [False, True])
@pytest.mark.parametrize("name", ["pore", "pore_dna", "rfpar", "membrane", "analytic"])
def test_grand_matches_oracle(name, bias):
    s = _systems()[name].copy()
    s.capacity = s.ntot + 400
    gc = _gcmc(s, bias=bias)
    ib, nat = configs.assign_buffers(s, gc)
    g, o = Engine(s), Oracle(s)
    g.set_gcmc(gc, ib)
    o.set_gcmc(gc, ib)
    rng = np.random.default_rng(7)
    natg, nato = nat.copy(), nat.copy()
    probg = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    probo = probg.copy()
    eg = eo = 0.0
    changed = 0
    for cyc in range(3):
        u = rng.random(6 * 120 + 8)
        ug, eg = g.grand(120, u, natg, probg, cyc + 1, eg)
        uo, eo = o.grand(120, u, nato, probo, cyc + 1, eo)
        assert ug == uo, (cyc, ug, uo)
        assert g.ntot == o.ntot
        assert np.array_equal(natg, nato) and np.array_equal(probg, probo)
        pg, tg = g.positions()
        po, to = o.positions()
        # exact before any BD step; afterwards the positions carry the steps' rounding
        assert np.max(np.abs(pg - po)) <= (0.0 if cyc == 0 else 1e-9)
        assert np.array_equal(tg, to)
        assert np.array_equal(g.ibuffer(), o.ibuffer())
        assert abs(eg - eo) <= TOL * max(abs(eo), 1.0), (eg, eo)
        changed += g.ntot != s.ntot
        g.step(7, abi.NOISE_PHILOX, seed=3, counter0=7 * cyc)
        o.step(7, abi.NOISE_PHILOX, seed=3, counter0=7 * cyc)
        assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-9
    assert changed > 0
    assert np.array_equal(g.warn(), o.warn())


def test_uniforms_exhausted():
    s = pore_small(seed=69, n_bp=0, n_ions=120, dna=False).copy()
    s.capacity = s.ntot + 100
    gc = _gcmc(s)
    ib, nat = configs.assign_buffers(s, gc)
    g = Engine(s)
    g.set_gcmc(gc, ib)
    prob = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    with pytest.raises(abi.BromocError) as ei:
        g.grand(50, np.full(20, 0.3), nat.copy(), prob, 1, 0.0)
    assert ei.value.code == abi.E_ARG


def test_capacity_error():
    s = plain(seed=2, n_per_type=10)
    g = Engine(s)
    gc = configs.gcmc_buffers(s, conc=5.0, buffer_len=20.0)
    ib, nat = configs.assign_buffers(s, gc)
    g.set_gcmc(gc, ib)
    prob = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    with pytest.raises(abi.BromocError) as ei:
        g.grand(1, np.full(10, 0.01), nat.copy(), prob, 1, 0.0)
    assert ei.value.code == abi.E_CAPACITY

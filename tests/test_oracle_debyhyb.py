"""Hybrid Debye screening of the CG-DNA's P-P Coulomb pairs (DEBYHYB: energy.mono.f90:
439-478, interact.f90:318-331): a pair is Debye-screened when both of its sites lie outside
the box (OUTBOX, grandmod.f90:74-97), plain Coulomb otherwise.

Pins: the P-P energy recomputed in numpy pair by pair from the sites' positions (screened
where both are outside), the reference's own routines (oracle/_ref; skipped where not
built) for energies, forces, BD steps and INTERACT.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from oracle.oracle import Oracle, RefFortran, ref_available

from helpers import energy_errs, force_err, pore_small

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def _hybrid(seed=91, nout=6):
    s0 = pore_small(seed=seed, n_ions=80)
    p_sites = [i for i in range(s0.nsites) if s0.namsite[i] == abi.SITE_P]
    return configs.with_debyhyb(s0, debye=9.6, outside=p_sites[:nout])


def test_qq_energy_by_pairs():
    s = _hybrid()
    o = Oracle(s)
    e = o.energy()
    pos = o.positions()[0]
    qq = s.lists["qq"][0].reshape(2, -1)
    b = s.box
    out = ((pos[:, 0] < b[0]) | (pos[:, 0] > b[1]) | (pos[:, 1] < b[2]) | (pos[:, 1] > b[3]) |
           (pos[:, 2] < b[4]) | (pos[:, 2] > b[5]))
    d = np.linalg.norm(pos[qq[1]] - pos[qq[0]], axis=1)
    both = out[qq[0]] & out[qq[1]]
    want = np.sum(np.where(both, s.fctn / d * np.exp(-d * s.ikappa), s.fctn / d))
    assert both.sum() >= 3 and (~both).sum() > 0
    assert abs(e["eqq"] - want) <= 1e-12 * abs(want)


@needs_ref
def test_hybrid_matches_reference():
    s = _hybrid(seed=92)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14
    pos = o.positions()[0]
    for j in range(0, s.nsites, 5):
        a, b = o.interact(pos[j], 0, j, True), r.interact(pos[j], 0, j, True)
        assert abs(a - b) <= 1e-13 * max(abs(b), 1.0), (j, a, b)
    noise = np.random.default_rng(4).standard_normal((4, 3 * s.ntot))
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(4):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    assert float(np.max(np.abs(o.positions()[0] - r.positions()[0]))) <= 1e-12

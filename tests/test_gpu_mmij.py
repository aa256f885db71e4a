"""The MMIJ reaction field of a rectangular region (rect_rf1.f90 in ENERGY, rect_rf0.f90
in INTERACT) on the HIP path against the C oracle.

Bar as the other parity tests: energies and norm-wise forces 1e-10 relative, positions
after BD steps 1e-9 A (1e-8 after 40 graph-replayed steps), INTERACT 1e-10 relative.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import energy_errs, force_err, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _sys(seed=51, npol=(3, 4, 5), n_ions=160, region=(30.0, 30.0, 44.0), rfscal=0.9):
    s = with_fixed(pore_small(seed=seed, n_ions=n_ions), 3, nonfree=(1,))
    return configs.with_mmij(s, npol=npol, rfscal=rfscal, region=region, seed=seed)


@pytest.mark.parametrize("npol", [(3, 3, 3), (3, 4, 5), (6, 5, 7)])
def test_energy_forces(npol):
    s = _sys(npol=npol)
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
    assert not bad, bad
    assert eo["egsbpb"] != 0.0
    assert force_err(g.forces(), o.forces()) <= TOL


def test_steps_and_graphs():
    s = _sys(seed=52)
    noise = np.random.default_rng(1).standard_normal((45, 3 * s.ntot))
    for nsteps, tol in ((5, 1e-9), (45, 1e-8)):
        g, o = Engine(s), Oracle(s)
        g.upload_noise(noise)
        o.upload_noise(noise)
        eg = g.step(nsteps, abi.NOISE_INJECTED)
        eo = o.step(nsteps, abi.NOISE_INJECTED)
        err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
        print(f"{nsteps} steps: position err {err:.3e}")
        assert err <= tol
        bad = {k: v for k, v in energy_errs(eg, eo).items() if v > 1e-9}
        assert not bad, bad


def test_interact():
    s = _sys(seed=53)
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    ion0 = s.nsites + s.nfix
    cases = [(pos[j] + 0.2, abs(typ[j]), j) for j in (ion0, ion0 + 9, s.ntot - 1)]
    cases += [((1.0, -2.0, 3.0), s.nold + 1, s.ntot), ((-9.0, 4.0, -12.0), s.nold + 2, s.ntot),
              ((0.0, 0.0, 25.0), s.nold + 1, s.ntot),        # outside the region: no term
              (pos[2], 0, 2), (pos[s.nsites], abs(typ[s.nsites]), s.nsites)]
    for xyz, jt, j in cases:
        a, b = g.interact(xyz, jt, j, True), o.interact(xyz, jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)


def test_large_basis():
    """640 basis functions and 2000 ions (the pore workload's ion count)."""
    s = _sys(seed=54, npol=(8, 8, 10), n_ions=2000, region=(36.0, 36.0, 56.0))
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    assert abs(eg["egsbpb"] - eo["egsbpb"]) <= TOL * abs(eo["egsbpb"])
    assert force_err(g.forces(), o.forces()) <= TOL

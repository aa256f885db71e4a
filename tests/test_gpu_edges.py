"""Edge cases of the HIP path against the C oracle: a DNA duplex without ions (with and
without QPAR), a single moving ion among fixed ones, a system whose ions are all fixed
(nothing moves), and one target per launch shape boundary (ntargets just above a multiple
of the block's targets).

Bar as the other parity tests: energies and norm-wise forces 1e-10 relative, positions
after BD steps 1e-9 A (1e-8 after 40 graph-replayed steps).
"""
import numpy as np
import pytest

from bromoc_amd import abi
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import energy_errs, force_err, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _energy(s):
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
    assert not bad, bad
    fo = o.forces()
    if np.max(np.abs(fo)) > 0.0:
        assert force_err(g.forces(), fo) <= TOL
    else:
        assert np.max(np.abs(g.forces())) == 0.0
    return eg, eo


def _steps(s, nsteps, tol, seed=2):
    noise = np.random.default_rng(seed).standard_normal((nsteps, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(nsteps, abi.NOISE_INJECTED)
    o.step(nsteps, abi.NOISE_INJECTED)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    assert err <= tol, err
    for a, b in zip(g.flux(), o.flux()):
        assert np.array_equal(a, b)
    return g, o


@pytest.mark.parametrize("qpar", [True, False])
def test_dna_without_ions(qpar):
    s = pore_small(seed=81, n_ions=0)
    if not qpar:
        s.flags &= ~abi.QPAR
    _energy(s)
    _steps(s, 5, 1e-9)
    _steps(s, 40, 1e-8)


def test_single_moving_ion():
    s = with_fixed(pore_small(seed=82, n_ions=6), 5)
    assert s.ntot - s.nsites - s.nfix == 1
    _energy(s)
    _steps(s, 5, 1e-9)
    _steps(s, 40, 1e-8)


def test_all_ions_fixed():
    s = with_fixed(pore_small(seed=83, n_ions=12, dna=False), 12)
    _energy(s)
    g, o = _steps(s, 3, 0.0)
    assert np.array_equal(g.positions()[0], np.stack([s.x, s.y, s.z], axis=1))


@pytest.mark.parametrize("n_ions", [129, 257, 513])
def test_target_counts_past_block_multiples(n_ions):
    """ntargets one past a multiple of the targets per block and of the block size."""
    s = pore_small(seed=84, n_bp=0, n_ions=n_ions, dna=False)
    _energy(s)
    _steps(s, 4, 1e-9)

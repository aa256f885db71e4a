"""The large translocation workload at full size (56,000 particles: 50,000 ions and a
1000-bp CG-DNA of 6000 sites with 21 M list terms, 321x321x641 repulsion / static-field /
reaction-field grids, PROFILE diffusion) against the C oracle: ENERGY (every term and the
forces) and one BD step with injected noise.  Bar: 1e-10 relative, positions 1e-9 A.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import energy_errs, force_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def large():
    s = configs.translocation()
    print(f"translocation system: {s.ntot} particles, {s.nsites} sites", flush=True)
    return s


def test_large_energy_and_step(large):
    s = large
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    errs = energy_errs(eg, eo)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= TOL
    assert force_err(g.forces(), o.forces()) <= TOL
    noise = np.random.default_rng(8).standard_normal((1, 3 * s.ntot))
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(1, abi.NOISE_INJECTED, 0)
    o.step(1, abi.NOISE_INJECTED, 0)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    print(f"1-step position err {err:.3e}", flush=True)
    assert err <= 1e-9
    for a, b in zip(g.flux(), o.flux()):
        assert np.array_equal(a, b)
    g.close()

"""EFPOT tables rebuilt on the device (bromoc_update_tables) and the table Jacobian's
vector product on the device (bromoc_table_vjp_device).

Bar: the device-built tables give the energies and forces of a context created with the
same tables built on the host (1e-13 relative: identical knots, only the autotuned launch
shape of the two contexts may order the pair sums differently), and identical BD
trajectories to 1e-12 A; the device VJP is bit-identical to the host's.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine, table_vjp

from helpers import energy_errs, force_err

pytestmark = pytest.mark.gpu


def _with_new_knots(s, seed):
    """The fitted pair types' samples replaced by a second seed's curves: the system with
    those tables built on the host, and the samples for the device update."""
    other = configs.adjoint_system(seed=s.meta["seed"], n_per_type=(s.ntot - 1) // 2,
                                   box=2 * s.box[1], curve_seed=seed)
    new_u = {is_: other.meta["params"][is_][0] for is_ in s.meta["params"]}
    t = configs.with_knots(s, new_u)
    return t, {is_: (s.meta["rgrid"], u) for is_, u in new_u.items()}


def test_update_tables_matches_host_built():
    s = configs.adjoint_system(seed=3, n_per_type=150, box=50.0)
    t, tables = _with_new_knots(s, 44)
    r0 = next(iter(tables.values()))[0]
    assert len(r0) > 100
    g = Engine(s)
    e_old = g.energy()
    g.update_tables(tables)
    e_dev, f_dev = g.energy(), g.forces()
    h = Engine(t)
    e_host, f_host = h.energy(), h.forces()
    bad = {k: v for k, v in energy_errs(e_dev, e_host).items() if v > 1e-13}
    assert not bad, bad
    assert force_err(f_dev, f_host) <= 1e-13
    assert abs(e_old["eefpot"] - e_host["eefpot"]) > 1e-6 * abs(e_host["eefpot"])
    noise = np.random.default_rng(5).standard_normal((40, 3 * s.ntot))
    for e in (g, h):
        e.upload_noise(noise)
        e.step(40, abi.NOISE_INJECTED)
    assert float(np.max(np.abs(g.positions()[0] - h.positions()[0]))) <= 1e-12


def test_update_tables_rejects_bad_input():
    s = configs.adjoint_system(seed=4, n_per_type=20, box=30.0)
    g = Engine(s)
    is_ = next(iter(s.meta["params"]))
    r, u = s.meta["rgrid"], s.meta["params"][is_][0]
    with pytest.raises(abi.BromocError):
        g.update_tables({is_: (r[:-1], u[:-1])})          # knot count differs
    with pytest.raises(abi.BromocError):
        g.update_tables({999: (r, u)})                     # no such pair type


def test_table_vjp_device_bitexact():
    rng = np.random.default_rng(6)
    s = configs.adjoint_system(seed=5, n_per_type=10, box=30.0)
    g = Engine(s)
    for n, lo, res in ((1851, 1.5, 0.01), (300, 2.0, 0.05), (2, 3.0, 0.1)):
        r = lo + res * np.arange(n)
        gep = rng.standard_normal(3 * n)
        gsc = rng.standard_normal(3)
        a = g.table_vjp(r, gep, gsc)
        b = table_vjp(r, gep, gsc)
        assert np.array_equal(a, b), np.max(np.abs(a - b))

"""HIP path vs the C oracle for the B-spline repulsion map (vdwgd1spln.f90, vdwgd0spln.f90),
APFOR and CONTRANS (energy.mono.f90:509-580), NOTRANS (simul1.f90:382-386), PROXDIFF /
DYNAMICS3 (energy.mono.f90:603-620, dynamics3.mono.f90), METROPOLIS (metropolis.f90,
move.f90) and MEMBRANE (membrane.f90, interact.f90:58-119).

Tolerances as the other parity tests: energies and norm-wise forces 1e-10 relative,
positions after BD steps 1e-9 A, integer outputs (warnings, flux, MC decisions) exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import energy_errs, force_err, near_wall, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _check_energy(eg, eo):
    errs = energy_errs(eg, eo)
    bad = {k: v for k, v in errs.items() if v > TOL}
    assert not bad, bad


def _energy_parity(s):
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    _check_energy(eg, eo)
    assert force_err(g.forces(), o.forces()) <= TOL
    assert np.array_equal(g.warn(), o.warn())
    return g, o, eg, eo


def _steps_parity(s, nsteps=6, seed=1, tol_pos=1e-9, graphs=False):
    noise = np.random.default_rng(seed).standard_normal((nsteps, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    if graphs:
        g.step(nsteps, abi.NOISE_INJECTED)
        o.step(nsteps, abi.NOISE_INJECTED)
    else:
        for k in range(nsteps):
            eg = g.step(1, abi.NOISE_INJECTED, k)
            eo = o.step(1, abi.NOISE_INJECTED, k)
        _check_energy(eg, eo)
        assert force_err(g.forces(), o.forces()) <= TOL
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    print(f"{nsteps}-step position err {err:.3e}")
    assert err <= tol_pos
    for a, b in zip(g.flux(), o.flux()):
        assert np.array_equal(a, b)
    assert np.array_equal(g.warn(), o.warn())
    return g, o


def _bspline(seed, thold27):
    s = near_wall(with_fixed(pore_small(seed=seed), 5, nonfree=(2,)), 16, seed=seed)
    return configs.with_bspline(s, thold27)


@pytest.mark.parametrize("thold27", [27.0, 3.0])
def test_bspline_energy_forces(thold27):
    g, o, eg, eo = _energy_parity(_bspline(61, thold27))
    assert eo["evdwgd"] != 0.0
    print("warn", o.warn())


def test_bspline_steps_and_graphs():
    s = _bspline(62, 5.0)
    _steps_parity(s, nsteps=5)
    _steps_parity(s, nsteps=40, graphs=True, tol_pos=1e-8)


def test_bspline_interact():
    s = _bspline(63, 3.0)
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 3, s.nsites - 1, s.nsites + 5, s.ntot - 1, s.ntot - 3)]
    cases += [((11.5, 0.0, 2.0), s.nold + 1, s.ntot), ((13.0, 1.0, -3.0), s.nold + 2, s.ntot),
              ((19.9, -19.9, 29.0), s.nold + 1, s.ntot)]
    for xyz, jt, j in cases:
        a, b = g.interact(xyz, jt, j, True), o.interact(xyz, jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)
    assert np.array_equal(g.warn(), o.warn())


def _restrained(seed, unsplit):
    s = with_fixed(pore_small(seed=seed), 3, nonfree=(1, 4))
    s = configs.with_apfor(s, seed=seed, n=5)
    s = configs.with_contrans(s, [(0, (2.0, 0.0, 3.5)), (4, (1.5, 2.5, 0.0)),
                                  (2, (0.0, 0.0, 4.0))], unsplit=unsplit)
    rng = np.random.default_rng(seed)
    s.x[:s.nsites] += rng.normal(0, 0.3, s.nsites)
    s.z[:s.nsites] += rng.normal(0, 0.3, s.nsites)
    return s


@pytest.mark.parametrize("unsplit", [False, True])
def test_apfor_contrans_energy_forces(unsplit):
    g, o, eg, eo = _energy_parity(_restrained(64, unsplit))
    assert eo["econ"] > 0.0


@pytest.mark.parametrize("unsplit", [False, True])
def test_apfor_contrans_steps(unsplit):
    s = _restrained(65, unsplit)
    _steps_parity(s, nsteps=5)
    _steps_parity(s, nsteps=40, graphs=True, tol_pos=1e-8)


@pytest.mark.parametrize("axes", [7, 5])
def test_notrans_steps(axes):
    s = configs.with_notrans(pore_small(seed=66), axes)
    g, _ = _steps_parity(s, nsteps=5)
    c = np.array(s.notrans[1])
    pc = g.positions()[0][:s.nsites].mean(axis=0)
    for a in range(3):
        if axes & (1 << a):
            assert abs(pc[a] - c[a]) <= 1e-10
    _steps_parity(s, nsteps=40, graphs=True, tol_pos=1e-8)


def test_notrans_not_in_dynamics():
    """NOTRANS belongs to the step loop (simul1.f90:382-386), not to DYNAMICS0NUC: one
    bromoc_dynamics pass leaves the centroid where the update puts it."""
    s = configs.with_notrans(pore_small(seed=67), 7)
    noise = np.random.default_rng(2).standard_normal((1, 3 * s.ntot))
    g = Engine(s)
    g.upload_noise(noise)
    g.energy()
    g.dynamics(abi.NOISE_INJECTED, 0)
    pc = g.positions()[0][:s.nsites].mean(axis=0)
    assert np.max(np.abs(pc - np.array(s.notrans[1]))) > 1e-6


@pytest.mark.parametrize("cut", [None, 6.0])
def test_proxdiff_steps(cut):
    s = configs.with_proxdiff(with_fixed(pore_small(seed=68, n_ions=200), 4), beta=2.5,
                              diff0=0.05, diffcutoff=cut)
    _steps_parity(s, nsteps=5)
    _steps_parity(s, nsteps=40, graphs=True, tol_pos=1e-8)


def test_proxdiff_philox():
    s = configs.with_proxdiff(pore_small(seed=69), beta=2.93, diff0=0.101)
    g, o = Engine(s), Oracle(s)
    g.step(37, abi.NOISE_PHILOX, seed=5, counter0=11)
    o.step(37, abi.NOISE_PHILOX, seed=5, counter0=11)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    assert err <= 1e-8, err


def test_metropolis_matches():
    s = with_fixed(pore_small(seed=70, n_bp=5, n_ions=120), 6)
    g, o = Engine(s), Oracle(s)
    u = np.random.default_rng(9).random(20000)
    ikbt = 1.0 / (0.0019872 * 300.0)
    ug, dg, eg = g.metropolis(300, 2.0, ikbt, u, 1.5)
    uo, do, eo = o.metropolis(300, 2.0, ikbt, u, 1.5)
    assert (ug, dg) == (uo, do) and dg == 300
    assert abs(eg - eo) <= TOL * max(abs(eo), 1.0)
    assert np.array_equal(g.positions()[0], o.positions()[0])
    assert np.array_equal(g.warn(), o.warn())
    # the device state after the moves feeds the next BD steps
    noise = np.random.default_rng(3).standard_normal((3, 3 * s.ntot))
    g.upload_noise(noise)
    o.upload_noise(noise)
    _check_energy(g.step(3, abi.NOISE_INJECTED), o.step(3, abi.NOISE_INJECTED))
    assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-9


def test_metropolis_short_stream():
    s = pore_small(seed=71, n_bp=4, n_ions=60)
    g, o = Engine(s), Oracle(s)
    u = np.random.default_rng(4).random(37)
    ikbt = 1.0 / (0.0019872 * 300.0)
    rg = g.metropolis(20, 2.0, ikbt, u, 0.0)
    ro = o.metropolis(20, 2.0, ikbt, u, 0.0)
    assert rg[:2] == ro[:2] and rg[1] < 20
    assert np.array_equal(g.positions()[0], o.positions()[0])


def test_all_options_together():
    """B-spline map, APFOR, CONTRANS, NOTRANS and PROXDIFF on one pore + DNA system."""
    s = _restrained(72, True)
    s = configs.with_bspline(near_wall(s, 12, seed=3), 4.0)
    s = configs.with_notrans(s, 3)
    s = configs.with_proxdiff(s, beta=2.0, diff0=0.08)
    _steps_parity(s, nsteps=4)
    _steps_parity(s, nsteps=36, graphs=True, tol_pos=1e-8)


def _membrane(seed, pore, nfix=4):
    s = with_fixed(pore_small(seed=seed, n_ions=180), nfix)
    s.flags &= ~(abi.QPHIV | abi.QPHIX)
    s.phiv = s.phix = None
    s = configs.with_membrane(s, thick=16.0, zmemb=1.0, volts=0.15, debye=9.0,
                              pore=(11.0, 5.0, 2.5) if pore else None)
    rng = np.random.default_rng(seed)
    idx = np.arange(s.ntot - 12, s.ntot)
    s.z[idx] = rng.uniform(-10.0, 12.0, 12)
    r = rng.uniform(9.0, 15.0, 12)
    th = rng.uniform(0, 2 * np.pi, 12)
    s.x[idx], s.y[idx] = r * np.cos(th), r * np.sin(th)
    return s


@pytest.mark.parametrize("pore", [False, True])
def test_membrane_energy_forces_steps(pore):
    s = _membrane(73, pore)
    g, o, eg, eo = _energy_parity(s)
    assert eo["ememb"] != 0.0
    _steps_parity(s, nsteps=5)
    _steps_parity(s, nsteps=40, graphs=True, tol_pos=1e-8)


def test_membrane_interact():
    s = _membrane(74, True)
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 1, s.nsites + 2, s.ntot - 1, s.ntot - 4, s.ntot - 7)]
    cases += [((0.5, 0.5, 3.0), s.nold + 1, s.ntot), ((12.0, 1.0, -8.5), s.nold + 2, s.ntot),
              ((14.0, 0.0, 9.5), s.nold + 1, s.ntot), ((3.0, 2.0, 25.0), s.nold + 2, s.ntot)]
    for xyz, jt, j in cases:
        a, b = g.interact(xyz, jt, j, True), o.interact(xyz, jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)
    assert np.array_equal(g.warn(), o.warn())


def test_debyhyb_energy_forces_steps():
    """Hybrid Debye (energy.mono.f90:439-478): P-P pairs of two sites outside the box
    screened, the rest plain Coulomb; ENERGY, INTERACT and BD steps."""
    s0 = pore_small(seed=75, n_ions=120)
    p_sites = [i for i in range(s0.nsites) if s0.namsite[i] == abi.SITE_P]
    s = configs.with_debyhyb(s0, debye=9.6, outside=p_sites[:6])
    g, o, eg, eo = _energy_parity(s)
    pos, typ = o.positions()
    for j in p_sites[:8]:
        a, b = g.interact(pos[j], 0, j, True), o.interact(pos[j], 0, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)
    _steps_parity(s, nsteps=5)
    _steps_parity(s, nsteps=40, graphs=True, tol_pos=1e-8)

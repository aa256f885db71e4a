"""HIP path vs the C oracle (and vs golden vectors of the reference) on identical inputs.

Tolerances: energies and norm-wise forces 1e-10 relative (the FP64 target); positions
after BD steps 1e-9 A absolute; integer outputs (flux counters, warnings, GCMC decisions,
Philox words) exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine, philox4x64
from oracle.oracle import Oracle, oracle_lib, _p

from helpers import energy_errs, force_err, plain, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _energy_forces(s):
    g, o = Engine(s), Oracle(s)
    eg, eo = g.energy(), o.energy()
    return g, o, eg, eo, g.forces(), o.forces()


def _check_energy(eg, eo, tol=TOL):
    errs = energy_errs(eg, eo)
    bad = {k: v for k, v in errs.items() if v > tol}
    assert not bad, bad


@pytest.mark.parametrize("n", [1, 2, 7, 64, 200])
def test_plain_ions_energy_forces(n):
    s = plain(seed=n, n_per_type=n)
    g, o, eg, eo, fg, fo = _energy_forces(s)
    print(f"plain N={s.ntot}: ener {eg['ener']!r} vs {eo['ener']!r}")
    _check_energy(eg, eo)
    if s.ntot > 1:
        assert force_err(fg, fo) <= TOL


def test_plain_ions_fixed_ions():
    s = with_fixed(plain(seed=3, n_per_type=50), 17)
    g, o, eg, eo, fg, fo = _energy_forces(s)
    _check_energy(eg, eo)
    assert force_err(fg, fo) <= TOL
    assert np.all(fg[:17] == 0.0)


def test_pore_dna_energy_forces():
    s = pore_small()
    g, o, eg, eo, fg, fo = _energy_forces(s)
    print({k: (eg[k], eo[k]) for k in eo if eo[k]})
    _check_energy(eg, eo)
    assert force_err(fg, fo) <= TOL
    assert np.array_equal(g.warn(), o.warn())


def test_pore_dna_fixed_and_nonfree_sites():
    s = with_fixed(pore_small(seed=9, n_bp=6, n_ions=120), 10, nonfree=(0, 3, 7, 20))
    g, o, eg, eo, fg, fo = _energy_forces(s)
    _check_energy(eg, eo)
    assert force_err(fg, fo) <= TOL


def test_pore_ions_no_dna():
    s = pore_small(seed=4, n_ions=300, dna=False)
    g, o, eg, eo, fg, fo = _energy_forces(s)
    _check_energy(eg, eo)
    assert force_err(fg, fo) <= TOL


@pytest.mark.parametrize("which", ["plain", "pore"])
def test_steps_injected_noise(which):
    s = plain(seed=21, n_per_type=100) if which == "plain" else pore_small(seed=6)
    nsteps = 6
    noise = np.random.default_rng(1).standard_normal((nsteps, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    eg = g.step(nsteps, abi.NOISE_INJECTED)
    eo = o.step(nsteps, abi.NOISE_INJECTED)
    _check_energy(eg, eo)
    pg, po = g.positions()[0], o.positions()[0]
    err = float(np.max(np.abs(pg - po)))
    print(f"{which}: {nsteps}-step position err {err:.3e}")
    assert err <= 1e-9
    assert force_err(g.forces(), o.forces()) <= TOL
    if len(s.zcont):
        for a, b in zip(g.flux(), o.flux()):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("nsteps", [33, 70])
def test_steps_graph_replays(nsteps):
    """Long step calls replay captured 32-step graphs (one per buffer parity); the result
    must equal step-by-step execution in the oracle."""
    s = plain(seed=31, n_per_type=40)
    noise = np.random.default_rng(7).standard_normal((nsteps + 1, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(1, abi.NOISE_INJECTED, 0)          # odd start: the other buffer parity
    o.step(1, abi.NOISE_INJECTED, 0)
    g.step(nsteps, abi.NOISE_INJECTED, 1)
    o.step(nsteps, abi.NOISE_INJECTED, 1)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    print(f"{nsteps} steps through graphs: position err {err:.3e}")
    assert err <= 1e-8


def test_steps_philox():
    s = pore_small(seed=7, n_bp=6, n_ions=150)
    g, o = Engine(s), Oracle(s)
    g.step(4, abi.NOISE_PHILOX, seed=1234, counter0=77)
    o.step(4, abi.NOISE_PHILOX, seed=1234, counter0=77)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    assert err <= 1e-9


def test_dynamics_after_energy_matches():
    s = plain(seed=5, n_per_type=80)
    noise = np.random.default_rng(2).standard_normal((1, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.energy()
    o.energy()
    g.dynamics(abi.NOISE_INJECTED, 0)
    o.dynamics(abi.NOISE_INJECTED, 0)
    assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-9


def test_philox_raw_bitexact():
    rng = np.random.default_rng(3)
    ctr = rng.integers(0, 2**63, (64, 4), dtype=np.uint64)
    key = rng.integers(0, 2**63, (64, 2), dtype=np.uint64)
    got = philox4x64(ctr, key)
    L = oracle_lib()
    import ctypes as C
    for k in range(64):
        c = np.ascontiguousarray(ctr[k])
        kk = np.ascontiguousarray(key[k])
        out = np.zeros(4, np.uint64)
        L.orc_philox4x64(_p(c, C.c_uint64), _p(kk, C.c_uint64), _p(out, C.c_uint64))
        assert np.array_equal(out, got[k])


def test_interact_matches():
    s = pore_small(seed=8, n_bp=6, n_ions=150)
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j) for j in
             (0, 1, 5, s.nsites - 1, s.nsites, s.nsites + 3, s.ntot - 1)]
    cases.append(((1.0, 2.0, -25.0), s.nold + 1, s.ntot))      # insertion trial
    cases.append(((15.0, 0.0, 0.0), s.nold + 2, s.ntot))      # trial inside the wall
    for xyz, jt, j in cases:
        a = g.interact(xyz, jt, j, alert=True)
        b = o.interact(xyz, jt, j, alert=True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)
    assert np.array_equal(g.warn(), o.warn())


def test_grand_matches():
    s = pore_small(seed=10, n_bp=0, n_ions=120, dna=False)
    s = s.copy()
    s.capacity = s.ntot + 200
    gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    g, o = Engine(s), Oracle(s)
    g.set_gcmc(gc, ib)
    o.set_gcmc(gc, ib)
    rng = np.random.default_rng(4)
    u = rng.random(20000)
    natg, nato = nat.copy(), nat.copy()
    probg = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    probo = probg.copy()
    ug, eg = g.grand(200, u, natg, probg, 1, 0.0)
    uo, eo = o.grand(200, u, nato, probo, 1, 0.0)
    assert ug == uo
    assert np.array_equal(natg, nato) and np.array_equal(probg, probo)
    assert g.ntot == o.ntot
    assert abs(eg - eo) <= TOL * max(abs(eo), 1.0)
    pg, tg = g.positions()
    po, to = o.positions()
    assert np.array_equal(pg, po) and np.array_equal(tg, to)
    assert np.array_equal(g.ibuffer(), o.ibuffer())


def test_steps_after_grand():
    """GCMC inserts/deletes change ntot and the particle order; the following BD steps
    (new partner order, new slots) must still match the oracle."""
    s = pore_small(seed=12, n_bp=0, n_ions=100, dna=False).copy()
    s.capacity = s.ntot + 200
    gc = configs.gcmc_buffers(s, conc=1.5, buffer_len=8.0, mu=(-0.3, -0.3))
    ib, nat = configs.assign_buffers(s, gc)
    g, o = Engine(s), Oracle(s)
    g.set_gcmc(gc, ib)
    o.set_gcmc(gc, ib)
    u = np.random.default_rng(6).random(20000)
    prob = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    natg, nato = nat.copy(), nat.copy()
    g.grand(150, u, natg, prob.copy(), 1, 0.0)
    o.grand(150, u, nato, prob.copy(), 1, 0.0)
    assert g.ntot == o.ntot and g.ntot != s.ntot
    noise = np.random.default_rng(8).standard_normal((40, 3 * g.ntot))
    g.upload_noise(noise)
    o.upload_noise(noise)
    eg = g.step(40, abi.NOISE_INJECTED, 0)
    eo = o.step(40, abi.NOISE_INJECTED, 0)
    _check_energy(eg, eo)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    assert err <= 1e-8, err


# --------------------------------- position-dependent diffusion and reaction field
def _steps_parity(s, nsteps=6, seed=1, tol_pos=1e-9):
    noise = np.random.default_rng(seed).standard_normal((nsteps, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    eg = g.step(nsteps, abi.NOISE_INJECTED)
    eo = o.step(nsteps, abi.NOISE_INJECTED)
    _check_energy(eg, eo)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    print(f"{nsteps}-step position err {err:.3e}")
    assert err <= tol_pos
    assert force_err(g.forces(), o.forces()) <= TOL
    for a, b in zip(g.flux(), o.flux()):
        assert np.array_equal(a, b)
    return g, o


@pytest.mark.parametrize("which", ["plain", "pore"])
def test_switch_diffusion_steps(which):
    if which == "plain":
        s = configs.with_switch_diffusion(plain(seed=41, n_per_type=80), porelength=20.0,
                                          pcenter=3.0, lowfraction=(0.3, 0.6),
                                          switchlength=(6.0, 9.0))
    else:
        s = configs.with_switch_diffusion(pore_small(seed=42), porelength=20.0,
                                          lowfraction=0.4, switchlength=5.0)
    _steps_parity(s)


@pytest.mark.parametrize("desc", [False, True])
def test_profile_diffusion_steps(desc):
    zk, fk = configs.tanh_profile(-24.0, 24.0, 41, 8.0)
    if desc:
        zk, fk = zk[::-1].copy(), fk[::-1].copy()
    s = configs.with_profile_diffusion(plain(seed=43, n_per_type=80), zk, fk)
    _steps_parity(s)


def test_profile_diffusion_philox_graphs():
    zk, fk = configs.tanh_profile(-15.0, 15.0, 25, 6.0, low=0.3)
    s = configs.with_profile_diffusion(pore_small(seed=44, n_bp=6, n_ions=150), zk, fk)
    g, o = Engine(s), Oracle(s)
    g.step(70, abi.NOISE_PHILOX, seed=99, counter0=5)
    o.step(70, abi.NOISE_PHILOX, seed=99, counter0=5)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    print(f"70 Philox steps with a D(z) profile: position err {err:.3e}")
    assert err <= 1e-8


def _rf_plain(seed, n, nfix=0, **kw):
    s = with_fixed(plain(seed=seed, n_per_type=n), nfix)
    return configs.with_rfpar(s, seed=seed, shape=(45, 45, 45), dcel=1.2, **kw)


@pytest.mark.parametrize("kw", [dict(), dict(single=True), dict(wall_radius=14.0, zero_frac=0.1)])
def test_rfpar_energy_forces(kw):
    s = _rf_plain(51, 60, nfix=7, **kw)
    g, o, eg, eo, fg, fo = _energy_forces(s)
    print(f"erfpar {eg['erfpar']!r} vs {eo['erfpar']!r}")
    assert eo["erfpar"] != 0.0
    _check_energy(eg, eo)
    assert force_err(fg, fo) <= TOL


def test_rfpar_pore_dna_steps():
    s = configs.with_rfpar(pore_small(seed=52), seed=3, shape=(37, 37, 55), dcel=1.0,
                           wall_radius=12.0, zero_frac=0.05)
    _steps_parity(s, nsteps=5)


def test_rfpar_graph_replays():
    s = _rf_plain(53, 40, nfix=3, zero_frac=0.05)
    noise = np.random.default_rng(9).standard_normal((40, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(40, abi.NOISE_INJECTED, 0)
    o.step(40, abi.NOISE_INJECTED, 0)
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    assert err <= 1e-8, err


def test_rfpar_interact():
    s = _rf_plain(54, 40, nfix=4, wall_radius=16.0, zero_frac=0.1)
    g, o = Engine(s), Oracle(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]), j) for j in (0, 3, 5, 17, s.ntot - 1)]
    cases += [((1.0, 2.0, -5.0), 1, s.ntot), ((4.0, -3.0, 9.0), 2, s.ntot),
              ((29.0, 0.0, 0.0), 1, s.ntot)]
    for xyz, jt, j in cases:
        a, b = g.interact(xyz, jt, j, True), o.interact(xyz, jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)


def test_capacity_error():
    s = plain(seed=2, n_per_type=10)
    g = Engine(s)
    gc = configs.gcmc_buffers(s, conc=5.0, buffer_len=20.0)
    ib, nat = configs.assign_buffers(s, gc)
    g.set_gcmc(gc, ib)
    prob = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    u = np.full(10, 0.01)       # insert into buffer 1, always accepted
    with pytest.raises(abi.BromocError) as ei:
        g.grand(1, u, nat.copy(), prob, 1, 0.0)
    assert ei.value.code == abi.E_CAPACITY


def test_full_size_pore_dna_config():
    """The 2000-ion + 300-site pore workload at full size: ENERGY and 2 BD steps."""
    s = configs.pore_dna(seed=1)
    noise = np.random.default_rng(5).standard_normal((2, 3 * s.ntot))
    g, o = Engine(s), Oracle(s)
    _check_energy(g.energy(), o.energy())
    assert force_err(g.forces(), o.forces()) <= TOL
    g.upload_noise(noise)
    o.upload_noise(noise)
    g.step(2)
    o.step(2)
    assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-9

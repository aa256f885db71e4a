"""The Fortran front-end (bromoc_amd/fortran/bromoc_gpu.f90) on the GPU: the reference's own
routines and harness with ENERGY, DYNAMICS0/1/2/3, DYNAMICS0NUC, INTERACT, GRAND and
METROPOLIS taken from the front-end, linked against the HIP library
(oracle/_ref/libbromoc_linkcheck.so), against the unmodified reference routines
(oracle/_ref/libbromoc_ref.so) on the same module state.
The front-end runs what a driver linked this way runs: it marshals the module globals into
bromoc_create, calls the device and writes energies, forces, positions, flux counters,
prob, nat and ibuffer back where the driver reads them.

Bar: energies 1e-10 relative (norm-wise for a term below 1e-6 of the largest, such as the
stacking energy at native geometry, which is rounding noise around zero), forces 1e-10
norm-wise (the device against the reference FP64 Fortran), positions 1e-9 A after BD
steps, INTERACT 1e-10, GRAND's counts, prob, particle count and buffers exact and its
energy 1e-10.  Skipped where oracle/_ref was not built (it is built from the reference
sources, with the HIP library already built).
"""
import os

import numpy as np
import pytest

from bromoc_amd import abi, configs
from oracle.oracle import FRONT_GPU_SO, REF_SO, RefFortran, oracle_table_builder

from helpers import energy_errs, force_err, with_fixed

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(FRONT_GPU_SO) and os.path.exists(REF_SO)),
                                 reason="oracle/_ref not built")]
TOL = 1e-10


def _pore(seed=5, n_bp=8, n_ions=120, dna=True):
    return configs.pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=(40.0, 40.0, 60.0),
                            dcel=0.5, rpore=12.0, half_slab=10.0, dna=dna,
                            table_builder=oracle_table_builder)


def _pair(s):
    # one live instance per library: the module state is global in each
    return RefFortran(s, lib=FRONT_GPU_SO), RefFortran(s)


@pytest.mark.parametrize("which", ["plain", "pore_dna"])
def test_energy_forces(which):
    if which == "plain":
        s = with_fixed(configs.plain_ions(seed=1, n_per_type=40, table_builder=oracle_table_builder), 9)
    else:
        s = with_fixed(_pore(), 7, nonfree=(0, 5, 11))
    f, r = _pair(s)
    ef, er = f.energy(), r.energy()
    bad = {k: v for k, v in energy_errs(ef, er).items() if v > TOL}
    assert not bad, bad
    assert force_err(f.forces(), r.forces()) <= TOL
    assert np.array_equal(f.warn(), r.warn())


def test_bd_steps():
    s = _pore(seed=7)
    seed, c0 = 1234, 17
    f, r = _pair(s)
    f.L.bromoc_gpu_set_seed(seed, c0)
    for k in range(4):
        f.energy()
        f.dynamics(abi.NOISE_PHILOX)        # the front-end draws from its own seed/counter
        r.energy()
        r.dynamics(abi.NOISE_PHILOX, 0, seed, c0 + k)
    dp = float(np.max(np.abs(f.positions()[0] - r.positions()[0])))
    print(f"4 BD steps through the front-end on the GPU: position difference {dp:.2e} A")
    assert dp <= 1e-9
    for a, b in zip(f.flux(), r.flux()):
        assert np.array_equal(a, b)


def test_interact():
    s = _pore(seed=8, n_bp=6, n_ions=100)
    f, r = _pair(s)
    pos, typ = r.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 3, s.nsites - 1, s.nsites, s.ntot - 1)]
    cases += [((1.0, 2.0, -25.0), s.nold + 1, s.ntot)]
    for xyz, jt, j in cases:
        a, b = f.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= TOL * max(abs(b), 1.0), (j, a, b)


def test_grand():
    s = _pore(seed=10, n_bp=0, n_ions=80, dna=False).copy()
    s.capacity = s.ntot + 150
    gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    f, r = _pair(s)
    f.set_gcmc(gc, ib)
    r.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(10000)
    pa = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    pb = pa.copy()
    na, nb = nat.copy(), nat.copy()
    _, ea = f.grand(100, u, na, pa, 1, 0.0)
    _, eb = r.grand(100, u, nb, pb, 1, 0.0)
    assert np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert abs(ea - eb) <= TOL * max(abs(eb), 1.0)
    assert f.ntot == r.ntot
    assert np.array_equal(f.positions()[0], r.positions()[0])
    assert np.array_equal(f.ibuffer(), r.ibuffer())


def _steps(s, seed, c0, n):
    f, r = _pair(s)
    f.L.bromoc_gpu_set_seed(seed, c0)
    for k in range(n):
        ef = f.energy()
        f.dynamics(abi.NOISE_PHILOX)
        er = r.energy()
        r.dynamics(abi.NOISE_PHILOX, 0, seed, c0 + k)
        bad = {k2: v for k2, v in energy_errs(ef, er).items() if v > TOL}
        assert not bad, (k, bad)
    dp = float(np.max(np.abs(f.positions()[0] - r.positions()[0])))
    assert dp <= 1e-9, dp
    assert np.array_equal(f.warn(), r.warn())
    return f, r


@pytest.mark.parametrize("which", ["switch", "profile", "rfpar"])
def test_diffusion_and_rfpar(which):
    from oracle.oracle import oracle_spline
    base = with_fixed(configs.plain_ions(seed=31, n_per_type=40, table_builder=oracle_table_builder), 5)
    if which == "switch":
        s = configs.with_switch_diffusion(base, porelength=20.0, pcenter=2.0,
                                          lowfraction=(0.3, 0.6), switchlength=(6.0, 9.0))
    elif which == "profile":
        zk, fk = configs.tanh_profile(-24.0, 24.0, 33, 8.0)
        s = configs.with_profile_diffusion(base, zk, fk, oracle_spline)
    else:
        s = configs.with_rfpar(base, seed=4, shape=(45, 45, 45), dcel=1.2, zero_frac=0.1)
    _steps(s, 77, 3, 3)


@pytest.mark.parametrize("which", ["bspline", "apfor_contrans", "proxdiff", "membrane", "mmij",
                                   "nmcden"])
def test_dna_options(which):
    """The module state of each option (B-spline repulsion map, APFOR + CONTRANS, PROXDIFF,
    MEMBRANE, MMIJ reaction field, per-species repulsion maps) handed over by the front-end
    to the device: energies and three BD steps."""
    base = with_fixed(_pore(seed=41), 4, nonfree=(3,))
    if which == "bspline":
        s = configs.with_bspline(base, thold27=3.0)
    elif which == "apfor_contrans":
        s = configs.with_contrans(configs.with_apfor(base, seed=5, n=4),
                                  [(0, (2.0, 1.0, 3.0)), (6, (1.0, 0.0, 2.0))], unsplit=True)
        s.x[:s.nsites] += 0.3
    elif which == "proxdiff":
        s = configs.with_proxdiff(base, beta=2.5, diff0=0.05)
    elif which == "membrane":
        s = configs.with_membrane(base, thick=16.0, volts=0.15, pore=(11.0, 5.0, 2.5))
    elif which == "mmij":
        s = configs.with_mmij(base, npol=(3, 4, 4), rfscal=0.8, region=(30.0, 30.0, 40.0))
    else:
        s = configs.with_nmcden(base, seed=7, density=0.2)
    _steps(s, 91, 4, 3)


def test_metropolis():
    s = with_fixed(_pore(seed=42, n_bp=5, n_ions=90), 6)
    f, r = _pair(s)
    u = np.random.default_rng(12).random(30000)
    ikbt = 1.0 / (0.0019872 * 300.0)
    _, _, ea = f.metropolis(250, 2.0, ikbt, u, 0.75)
    _, _, eb = r.metropolis(250, 2.0, ikbt, u, 0.75)
    assert abs(ea - eb) <= TOL * max(abs(eb), 1.0)
    assert np.array_equal(f.positions()[0], r.positions()[0])


@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_curved_systems(which):
    """Spherical and elliptic-cylinder systems (Qsphere, Qecyl with cx, cy, cz, rsphe,
    rsphe2, iecx, iecy and the buffers' Rmin/Rmax): BD steps whose noise carries ions past
    the wall (FIXCOOR), then GRAND (INSERT's redraws)."""
    base = configs.pore_dna(seed=10, n_bp=0, n_ions=80, box=(40.0, 40.0, 60.0), dcel=0.5,
                            rpore=12.0, half_slab=10.0, dna=False,
                            table_builder=oracle_table_builder)
    if which == "sphere":
        s = configs.with_sphere(base, 20.0, seed=2).copy()
    else:
        s = configs.with_ecyl(base, seed=2).copy()
    s.capacity = s.ntot + 150
    seed, c0 = 55, 9
    f, r = _pair(s)
    f.L.bromoc_gpu_set_seed(seed, c0)
    for k in range(3):
        f.energy()
        f.dynamics(abi.NOISE_PHILOX)
        r.energy()
        r.dynamics(abi.NOISE_PHILOX, 0, seed, c0 + k)
    assert float(np.max(np.abs(f.positions()[0] - r.positions()[0]))) <= 1e-9
    s2 = s.copy()
    s2.x, s2.y, s2.z = (np.ascontiguousarray(v) for v in r.positions()[0].T)
    if which == "sphere":
        gc = configs.sphere_buffers(s2, conc=1.0, shell_len=6.0, mu=(-0.5, -0.4))
    else:
        gc = configs.gcmc_buffers(s2, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s2, gc)
    f.set_gcmc(gc, ib)
    r.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(20000)
    pa = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    pb = pa.copy()
    na, nb = nat.copy(), nat.copy()
    _, ea = f.grand(60, u, na, pa, 1, 0.0)
    _, eb = r.grand(60, u, nb, pb, 1, 0.0)
    assert np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert abs(ea - eb) <= TOL * max(abs(eb), 1.0)
    assert f.ntot == r.ntot
    assert float(np.max(np.abs(f.positions()[0] - r.positions()[0]))) <= 1e-9

"""The Fortran front-end (bromoc_amd/fortran/bromoc_gpu.f90) linked into the reference's own
routines in place of ENERGY, DYNAMICS0/1/2, DYNAMICS0NUC, INTERACT and GRAND, with its C-ABI
calls served by the C oracle (oracle/shim_adapter.c), against the unmodified reference
routines: shows that the front-end hands the driver's module state (flags, pair tables,
knot ranges, DNA lists, grids, GCMC buffers) to the C ABI and writes the results back
where the driver reads them.  CPU only; skipped where oracle/_ref is not built.
"""
import os

import numpy as np
import pytest

from bromoc_amd import abi, configs
from oracle.oracle import SHIM_SO, RefFortran, oracle_table_builder, ref_available

from helpers import energy_errs, force_err, with_fixed

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def small_plain(seed, n):
    return configs.plain_ions(seed=seed, n_per_type=n, table_builder=oracle_table_builder)


def small_pore(seed=5, n_bp=8, n_ions=120, dna=True):
    return configs.pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=(40.0, 40.0, 60.0),
                            dcel=0.5, rpore=12.0, half_slab=10.0, dna=dna,
                            table_builder=oracle_table_builder)


def _pair(s):
    # one live instance per library: module state is global in each
    return RefFortran(s, lib=SHIM_SO), RefFortran(s)


@pytest.mark.parametrize("which", ["plain", "pore", "pore_deby_svdw"])
def test_energy_through_front_end(which):
    if which == "plain":
        s = with_fixed(small_plain(1, 40), 9)
    elif which == "pore":
        s = with_fixed(small_pore(), 7, nonfree=(0, 5, 11))
    else:
        s = small_pore(seed=6, n_bp=6, n_ions=80).copy()
        s.flags |= abi.QDEBY | abi.QSVDW
        s.kappa, s.ikappa = 9.6, 1.0 / 9.6
        s.scal = np.linspace(0.5, 1.5, s.nttyp + 1)
    f, r = _pair(s)
    ef, er = f.energy(), r.energy()
    assert max(energy_errs(ef, er, floor=1e-12).values()) <= 1e-14
    assert force_err(f.forces(), r.forces()) <= 1e-14
    assert np.array_equal(f.warn(), r.warn())
    # the energy-only call leaves the energies the same
    ef2 = f.energy(forces=False)
    assert abs(ef2["ener"] - ef["ener"]) <= 1e-14 * abs(ef["ener"])


@pytest.mark.parametrize("which", ["plain", "pore"])
def test_bd_steps_through_front_end(which):
    s = small_plain(4, 50) if which == "plain" else small_pore(seed=7)
    seed, c0 = 1234, 17
    f, r = _pair(s)
    f.L.bromoc_gpu_set_seed(seed, c0)
    for k in range(4):
        f.energy()
        f.dynamics(abi.NOISE_PHILOX)        # the front-end draws from its own seed/counter
        r.energy()
        r.dynamics(abi.NOISE_PHILOX, 0, seed, c0 + k)
    pf, pr = f.positions()[0], r.positions()[0]
    assert float(np.max(np.abs(pf - pr))) <= 1e-12
    for a, b in zip(f.flux(), r.flux()):
        assert np.array_equal(a, b)


def test_interact_through_front_end():
    s = small_pore(seed=8, n_bp=6, n_ions=100)
    f, r = _pair(s)
    pos, typ = r.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 3, s.nsites - 1, s.nsites, s.ntot - 1)]
    cases += [((1.0, 2.0, -25.0), s.nold + 1, s.ntot)]
    for xyz, jt, j in cases:
        a, b = f.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1.0), (j, a, b)


def test_grand_through_front_end():
    s = small_pore(seed=10, n_bp=0, n_ions=80, dna=False).copy()
    s.capacity = s.ntot + 150
    gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    f, r = _pair(s)
    f.set_gcmc(gc, ib)
    r.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(10000)
    # the driver's prob(nbuffer, 0:datom) is handed over whole: compare its first columns
    pa = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    pb = pa.copy()
    na, nb = nat.copy(), nat.copy()
    _, ea = f.grand(100, u, na, pa, 1, 0.0)
    _, eb = r.grand(100, u, nb, pb, 1, 0.0)
    assert np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert f.ntot == r.ntot
    assert np.array_equal(f.positions()[0], r.positions()[0])
    assert np.array_equal(f.ibuffer(), r.ibuffer())


@pytest.mark.parametrize("which", ["switch", "profile", "rfpar"])
def test_diffusion_and_rfpar_through_front_end(which):
    from oracle.oracle import oracle_spline
    base = with_fixed(small_plain(31, 40), 5)
    if which == "switch":
        s = configs.with_switch_diffusion(base, porelength=20.0, pcenter=2.0,
                                          lowfraction=(0.3, 0.6), switchlength=(6.0, 9.0))
    elif which == "profile":
        zk, fk = configs.tanh_profile(-24.0, 24.0, 33, 8.0)
        s = configs.with_profile_diffusion(base, zk, fk, oracle_spline)
    else:
        s = configs.with_rfpar(base, seed=4, shape=(45, 45, 45), dcel=1.2, zero_frac=0.1)
    seed, c0 = 77, 3
    f, r = _pair(s)
    f.L.bromoc_gpu_set_seed(seed, c0)
    for k in range(3):
        ef = f.energy()
        f.dynamics(abi.NOISE_PHILOX)
        er = r.energy()
        r.dynamics(abi.NOISE_PHILOX, 0, seed, c0 + k)
        assert max(energy_errs(ef, er, floor=1e-12).values()) <= 1e-14
    pf, pr = f.positions()[0], r.positions()[0]
    assert float(np.max(np.abs(pf - pr))) <= 1e-12
    if which == "rfpar":
        pos, typ = r.positions()
        for j in (5, 9, s.ntot - 1):
            a = f.interact(pos[j], abs(typ[j]), j, True)
            b = r.interact(pos[j], abs(typ[j]), j, True)
            assert abs(a - b) <= 1e-14 * max(abs(b), 1.0)


@pytest.mark.parametrize("which", ["bspline", "apfor_contrans", "proxdiff", "membrane", "mmij",
                                   "nmcden"])
def test_dna_options_through_front_end(which):
    """B-spline repulsion map (Qtrln off), APFOR + CONTRANS (apfmod, nucleotmod), PROXDIFF
    (DYNAMICS3), MEMBRANE, the MMIJ reaction field and the per-species repulsion maps
    (gsbpmod) handed over from the modules: energies, forces and BD steps."""
    base = with_fixed(small_pore(seed=41), 4, nonfree=(3,))
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
    seed, c0 = 91, 4
    f, r = _pair(s)
    f.L.bromoc_gpu_set_seed(seed, c0)
    for k in range(3):
        ef = f.energy()
        f.dynamics(abi.NOISE_PHILOX)
        er = r.energy()
        r.dynamics(abi.NOISE_PHILOX, 0, seed, c0 + k)
        assert max(energy_errs(ef, er, floor=1e-12).values()) <= 1e-14
    pf, pr = f.positions()[0], r.positions()[0]
    assert float(np.max(np.abs(pf - pr))) <= 1e-12
    assert np.array_equal(f.warn(), r.warn())


def test_metropolis_through_front_end():
    s = with_fixed(small_pore(seed=42, n_bp=5, n_ions=90), 6)
    f, r = _pair(s)
    u = np.random.default_rng(12).random(30000)
    ikbt = 1.0 / (0.0019872 * 300.0)
    _, _, ea = f.metropolis(250, 2.0, ikbt, u, 0.75)
    _, _, eb = r.metropolis(250, 2.0, ikbt, u, 0.75)
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert np.array_equal(f.positions()[0], r.positions()[0])


def test_repwalls_through_front_end():
    """REPWALLS replaced by the front-end's external (bromoc_repwalls on the module map)."""
    from oracle.oracle import ref_lib
    from test_maps import _body_map
    import ctypes as C
    phiv = _body_map()
    outs = []
    for path in (SHIM_SO, None):
        L = ref_lib(path) if path else ref_lib()
        m = phiv.copy()
        L.ref_repwalls(*m.shape, m.ctypes.data_as(C.POINTER(C.c_int8)), 126)
        outs.append(m)
    assert np.array_equal(outs[0], outs[1]) and outs[0].max() > 1


FRONT_END = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         "bromoc_amd", "fortran", "bromoc_gpu.f90")


def test_front_end_entry_points_are_exported():
    """Every C entry point the Fortran front-end binds (an interface block's
    bind(C, name=...)) is exported by the HIP library."""
    import re
    src = open(FRONT_END).read()
    iface = src[src.index("interface"):src.index("end interface")]
    names = set(re.findall(r"bind\(C,\s*name='(\w+)'\)", iface))
    assert {"bromoc_create", "bromoc_step", "bromoc_grand", "bromoc_interact"} <= names
    lib = abi.load_library()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not ref_available(), reason="reference sources not present here")
def test_front_end_links_against_the_library():
    """The front-end and the reference objects it is linked with (a driver's build) link
    against libbromoc_gpu.so with no undefined symbol; the linked library loads."""
    import ctypes
    import subprocess
    oracle_dir = os.path.join(os.path.dirname(FRONT_END), "..", "..", "oracle")
    r = subprocess.run(["make", "-s", "-C", oracle_dir, "linkcheck"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    ctypes.CDLL(os.path.join(oracle_dir, "_ref", "libbromoc_linkcheck.so"))


@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_curved_systems_through_front_end(which):
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
    assert float(np.max(np.abs(f.positions()[0] - r.positions()[0]))) <= 1e-12
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
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert f.ntot == r.ntot
    assert float(np.max(np.abs(f.positions()[0] - r.positions()[0]))) <= 1e-12

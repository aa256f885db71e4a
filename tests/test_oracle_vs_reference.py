"""The C oracle against the reference's own Fortran routines (oracle/_ref, built from the
reference sources by oracle/Makefile).  Skipped where that build is absent (the GPU box).

Bar: bit-identical where both evaluate the same expression sequence; <= 1e-14 relative
where the Fortran compiler's evaluation order of a long expression can differ by an ulp.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from oracle.oracle import (Oracle, RefFortran, oracle_table_builder, ref_available,
                           ref_table_builder)

from helpers import energy_errs, force_err, near_wall, with_fixed

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def small_plain(seed, n, **kw):
    return configs.plain_ions(seed=seed, n_per_type=n, table_builder=oracle_table_builder, **kw)


def small_pore(seed=5, n_bp=8, n_ions=120, dna=True):
    return configs.pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=(40.0, 40.0, 60.0),
                            dcel=0.5, rpore=12.0, half_slab=10.0, dna=dna,
                            table_builder=oracle_table_builder)


@pytest.mark.parametrize("q,lo,hi,res", [(-1.0, 1.5, 20.0, 0.01), (1.0, 2.0, 12.0, 0.05),
                                         (0.0, 1.5, 15.0, 0.1)])
def test_table_build_bitexact(q, lo, hi, res):
    r = configs.knot_grid(lo, hi, res)
    u = configs.efpot_curve(r, 3.7, 0.25, q, 80.0)
    fct = configs.CELEC * q / 80.0
    chr_ = 1 if q else 0
    e1, p1 = oracle_table_builder(len(r), r, u, chr_, fct)
    e2, p2 = ref_table_builder(len(r), r, u, chr_, fct)
    e3, p3 = configs.lib_table_builder(len(r), r, u, chr_, fct)
    assert np.array_equal(e1, e2) and np.array_equal(e3, e2)
    for f in ("sc1", "sc2", "sc3", "dmi", "dm2lo", "dm2hi"):
        assert getattr(p1, f) == getattr(p2, f) == getattr(p3, f), f


@pytest.mark.parametrize("seed,n,nfix", [(0, 1, 0), (1, 30, 0), (2, 100, 0), (3, 60, 13)])
def test_energy_plain(seed, n, nfix):
    s = with_fixed(small_plain(seed, n), nfix)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert np.array_equal(o.forces(), r.forces())


def test_energy_pore_dna():
    s = with_fixed(small_pore(), 7, nonfree=(0, 5, 11))
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14
    assert np.array_equal(o.warn(), r.warn())


def test_energy_debye_and_svdw():
    s = small_pore(seed=6, n_bp=6, n_ions=80).copy()
    s.flags |= abi.QDEBY | abi.QSVDW
    s.kappa, s.ikappa = 9.6, 1.0 / 9.6
    s.scal = np.linspace(0.5, 1.5, s.nttyp + 1)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14


@pytest.mark.parametrize("which", ["plain", "pore"])
def test_dynamics_bitexact(which):
    s = small_plain(4, 50) if which == "plain" else small_pore(seed=7)
    noise = np.random.default_rng(3).standard_normal((5, 3 * s.ntot))
    o, r = Oracle(s), RefFortran(s)
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(5):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    po, pr = o.positions()[0], r.positions()[0]
    assert float(np.max(np.abs(po - pr))) <= 1e-12
    for a, b in zip(o.flux(), r.flux()):
        assert np.array_equal(a, b)


def test_interact():
    s = small_pore(seed=8, n_bp=6, n_ions=100)
    o, r = Oracle(s), RefFortran(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 1, 2, 7, s.nsites - 1, s.nsites, s.ntot - 1)]
    cases += [((1.0, 2.0, -25.0), s.nold + 1, s.ntot), ((15.0, 0.0, 0.0), s.nold + 2, s.ntot)]
    for xyz, jt, j in cases:
        a, b = o.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1.0), (j, a, b)
    assert np.array_equal(o.warn(), r.warn())


def test_grand():
    s = small_pore(seed=10, n_bp=0, n_ions=80, dna=False).copy()
    s.capacity = s.ntot + 150
    gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    o, r = Oracle(s), RefFortran(s)
    o.set_gcmc(gc, ib)
    r.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(10000)
    pa = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    pb = pa.copy()
    na, nb = nat.copy(), nat.copy()
    ua, ea = o.grand(100, u, na, pa, 1, 0.0)
    ub, eb = r.grand(100, u, nb, pb, 1, 0.0)
    assert ua == ub and np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert o.ntot == r.ntot
    assert np.array_equal(o.positions()[0], r.positions()[0])
    assert np.array_equal(o.ibuffer(), r.ibuffer())


# ------------------------------------------------ DYNAMICS1 / DYNAMICS2 / SPLINE / RFPAR
from oracle.oracle import oracle_spline, ref_spline  # noqa: E402


@pytest.mark.parametrize("n,desc", [(2, False), (3, False), (4, False), (11, False),
                                    (60, False), (25, True)])
def test_spline_bitexact(n, desc):
    rng = np.random.default_rng(n)
    x = np.cumsum(rng.uniform(0.3, 2.0, n)) - 10.0
    if desc:
        x = x[::-1].copy()
    y = rng.uniform(0.2, 1.0, n)
    b1, c1, d1 = oracle_spline(x, y)
    b2, c2, d2 = ref_spline(x, y)
    b3, c3, d3 = configs.lib_spline(x, y)
    for a, b, c in ((b1, b2, b3), (c1, c2, c3), (d1, d2, d3)):
        assert np.array_equal(a, b) and np.array_equal(c, b)


def _steps_match(s, nsteps=5, seed=3):
    noise = np.random.default_rng(seed).standard_normal((nsteps, 3 * s.ntot))
    o, r = Oracle(s), RefFortran(s)
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(nsteps):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    po, pr = o.positions()[0], r.positions()[0]
    assert float(np.max(np.abs(po - pr))) <= 1e-12
    for a, b in zip(o.flux(), r.flux()):
        assert np.array_equal(a, b)
    return po


@pytest.mark.parametrize("which", ["plain", "pore"])
def test_dynamics1_bitexact(which):
    if which == "plain":
        s = configs.with_switch_diffusion(small_plain(11, 60), porelength=20.0, pcenter=3.0,
                                          lowfraction=(0.3, 0.6), switchlength=(6.0, 9.0))
    else:
        s = configs.with_switch_diffusion(small_pore(seed=12), porelength=20.0,
                                          lowfraction=0.4, switchlength=5.0)
    _steps_match(s)


@pytest.mark.parametrize("desc", [False, True])
def test_dynamics2_bitexact(desc):
    zk, fk = configs.tanh_profile(-24.0, 24.0, 33, 8.0)
    if desc:
        zk, fk = zk[::-1].copy(), fk[::-1].copy()
    s = configs.with_profile_diffusion(small_plain(13, 60), zk, fk, oracle_spline)
    _steps_match(s)


def _rf_plain(seed, n, nfix=0, **kw):
    s = with_fixed(small_plain(seed, n), nfix)
    return configs.with_rfpar(s, seed=seed, shape=(45, 45, 45), dcel=1.2, **kw)


@pytest.mark.parametrize("kw", [dict(), dict(single=True), dict(wall_radius=14.0),
                                dict(zero_frac=0.15)])
def test_rfpar_energy(kw):
    s = _rf_plain(21, 50, nfix=6, **kw)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert er["erfpar"] != 0.0 and er["egsbpb"] == er["erfpar"]
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14


def test_rfpar_pore_dna_steps():
    s = configs.with_rfpar(small_pore(seed=22), seed=3, shape=(37, 37, 55), dcel=1.0,
                           wall_radius=12.0)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14
    _steps_match(s, nsteps=3)


def test_rfpar_interact():
    s = _rf_plain(23, 40, nfix=4, wall_radius=16.0, zero_frac=0.15)
    o, r = Oracle(s), RefFortran(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]), j) for j in (0, 3, 5, 17, s.ntot - 1)]
    cases += [((1.0, 2.0, -5.0), 1, s.ntot), ((4.0, -3.0, 9.0), 2, s.ntot),
              ((29.0, 0.0, 0.0), 1, s.ntot)]       # outside the grid box
    for xyz, jt, j in cases:
        a, b = o.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1.0), (j, a, b)


# ---------------------- VDWGD1SPLN / APFOR / CONTRANS / NOTRANS / PROXDIFF / METROPOLIS
def _energy_match(s):
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14
    assert np.array_equal(o.warn(), r.warn())
    return eo, er


@pytest.mark.parametrize("thold27", [27.0, 3.0])
def test_bspline_energy_steps_interact(thold27):
    """27-point B-spline repulsion (vdwgd1spln.f90, vdwgd0spln.f90); a low threshold
    exercises the wall warning and INTERACT's 1e10 wall energy."""
    s = near_wall(with_fixed(small_pore(seed=31), 5, nonfree=(2,)), 12)
    s = configs.with_bspline(s, thold27)
    eo, er = _energy_match(s)
    assert er["evdwgd"] != 0.0
    _steps_match(s, nsteps=3)
    o, r = Oracle(s), RefFortran(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 3, s.nsites - 1, s.nsites + 5, s.ntot - 1)]
    cases += [((11.5, 0.0, 2.0), s.nold + 1, s.ntot), ((13.0, 1.0, -3.0), s.nold + 2, s.ntot),
              ((19.9, -19.9, 29.0), s.nold + 1, s.ntot)]
    for xyz, jt, j in cases:
        a, b = o.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1.0), (j, a, b)
    assert np.array_equal(o.warn(), r.warn())


def test_apfor_contrans_energy():
    s = with_fixed(small_pore(seed=32), 3, nonfree=(1, 4))
    s = configs.with_apfor(s, seed=2, n=5)
    s = configs.with_contrans(s, [(0, (2.0, 0.0, 3.5)), (4, (1.5, 2.5, 0.0)),
                                  (2, (0.0, 0.0, 4.0))])
    rng = np.random.default_rng(1)           # move the sites off the restraint centres
    s.x[:s.nsites] += rng.normal(0, 0.3, s.nsites)
    s.z[:s.nsites] += rng.normal(0, 0.3, s.nsites)
    eo, er = _energy_match(s)
    assert er["econ"] > 0.0
    _steps_match(s, nsteps=3)


def test_contrans_unsplit():
    s = configs.with_contrans(small_pore(seed=33), [(0, (1.0, 2.0, 3.0)), (7, (0.5, 0.0, 0.0))],
                              unsplit=True)
    s.y[:s.nsites] += 0.4
    eo, er = _energy_match(s)
    assert er["econ"] > 0.0
    _steps_match(s, nsteps=3)


@pytest.mark.parametrize("axes", [7, 4, 3])
def test_notrans_steps(axes):
    s = configs.with_notrans(small_pore(seed=34), axes)
    po = _steps_match(s, nsteps=4)
    c = np.array([s.x[:s.nsites].mean(), s.y[:s.nsites].mean(), s.z[:s.nsites].mean()])
    pc = po[:s.nsites].mean(axis=0)
    for a in range(3):
        if axes & (1 << a):
            assert abs(pc[a] - c[a]) <= 1e-12
        else:
            assert abs(pc[a] - c[a]) > 1e-6


def test_proxdiff_records():
    """ENERGY's nearest-site records (energy.mono.f90:583,603-620): the oracle's equal the
    reference's, and the configuration has ions with a record and ions whose first site is
    outside its table range while other sites are inside (the reference then leaves the
    record at zero)."""
    s = configs.with_proxdiff(with_fixed(small_pore(seed=35, n_ions=160), 4))
    o, r = Oracle(s), RefFortran(s)
    o.energy()
    r.energy()
    do, dr = o.dids(), r.dids()
    assert np.array_equal(do, dr)
    pos = o.positions()[0]
    ions = np.arange(s.nsites, s.ntot)
    d0 = np.linalg.norm(pos[ions] - pos[0], axis=1)
    dmin = np.min(np.linalg.norm(pos[ions, None, :] - pos[None, :s.nsites, :], axis=2), axis=1)
    far0 = (d0 > 20.0) & (dmin < 15.0)
    assert far0.sum() > 5 and np.all(dr[ions[far0]] == 0.0)
    assert np.sum(dr[ions, 0] > 0.0) > 20


@pytest.mark.parametrize("cut", [None, 6.0])
def test_proxdiff_steps(cut):
    base = with_fixed(small_pore(seed=35, n_ions=160), 4)
    s = configs.with_proxdiff(base, beta=2.5, diff0=0.05, diffcutoff=cut)
    p3 = _steps_match(s, nsteps=4)
    p0 = _steps_match(base, nsteps=4)
    assert float(np.max(np.abs(p3 - p0))) > 1e-3     # DYNAMICS3 moved the ions differently


def test_metropolis():
    s = with_fixed(small_pore(seed=36, n_bp=5, n_ions=90), 6)
    o, r = Oracle(s), RefFortran(s)
    u = np.random.default_rng(9).random(20000)
    ikbt = 1.0 / (0.0019872 * 300.0)
    ua, da, ea = o.metropolis(300, 2.0, ikbt, u, 1.5)
    ub, db, eb = r.metropolis(300, 2.0, ikbt, u, 1.5)
    assert da == 300 and ua == ub
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert np.array_equal(o.positions()[0], r.positions()[0])
    assert np.array_equal(o.warn(), r.warn())


# ---------------------------------------------------------------- MEMBRANE
def _membrane(seed, pore, nfix=4):
    """Plain ions (and a DNA duplex) across a membrane slab; some ions inside the membrane
    (the reference warns and, in INTERACT, charges 1e10 there) and in the walls."""
    s = with_fixed(small_pore(seed=seed, n_ions=140), nfix)
    s.flags &= ~(abi.QPHIV | abi.QPHIX)
    s.phiv = s.phix = None
    s = configs.with_membrane(s, thick=16.0, zmemb=1.0, volts=0.15, debye=9.0,
                              pore=(11.0, 5.0, 2.5) if pore else None)
    rng = np.random.default_rng(seed)
    k = 10
    idx = np.arange(s.ntot - k, s.ntot)
    s.z[idx] = rng.uniform(-10.0, 12.0, k)           # in and beside the membrane
    r = rng.uniform(9.0, 15.0, k)
    th = rng.uniform(0, 2 * np.pi, k)
    s.x[idx], s.y[idx] = r * np.cos(th), r * np.sin(th)
    return s


@pytest.mark.parametrize("pore", [False, True])
def test_membrane_energy_steps_interact(pore):
    s = _membrane(81, pore)
    eo, er = _energy_match(s)
    assert er["ememb"] != 0.0
    _steps_match(s, nsteps=3)
    o, r = Oracle(s), RefFortran(s)
    pos, typ = o.positions()
    cases = [(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
             for j in (0, 1, s.nsites + 2, s.ntot - 1, s.ntot - 4, s.ntot - 7)]
    cases += [((0.5, 0.5, 3.0), s.nold + 1, s.ntot), ((12.0, 1.0, -8.5), s.nold + 2, s.ntot),
              ((14.0, 0.0, 9.5), s.nold + 1, s.ntot), ((3.0, 2.0, 25.0), s.nold + 2, s.ntot)]
    for xyz, jt, j in cases:
        a, b = o.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1.0), (j, a, b)
    assert np.array_equal(o.warn(), r.warn())


def test_scalepairing_vs_reference_and_interact_quirk():
    """Base pairing scaled by scalepairing (energy.mono.f90:372-376) against the reference;
    the energy is linear in it (ebp at 1.7 is 1.7 times ebp at 1.0), while INTERACT's base
    pair term leaves it out (interact.f90:288-298): a site's INTERACT is the same at both."""
    base = small_pore(seed=9, n_bp=6, n_ions=60)
    sys_ = {}
    for sp in (1.0, 1.7):
        s = base.copy()
        s.scalepairing = sp
        sys_[sp] = s
        o, r = Oracle(s), RefFortran(s)
        eo, er = o.energy(), r.energy()
        assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
        assert force_err(o.forces(), r.forces()) <= 1e-14
        pos, _ = o.positions()
        for j in range(0, s.nsites, 5):
            a, b = o.interact(pos[j], 0, j, False), r.interact(pos[j], 0, j, False)
            assert abs(a - b) <= 1e-14 * max(abs(b), 1.0)
    e1, e17 = Oracle(sys_[1.0]).energy(), Oracle(sys_[1.7]).energy()
    assert e1["ebp"] != 0.0
    assert e17["ebp"] == pytest.approx(1.7 * e1["ebp"], rel=1e-13)
    o1, o17 = Oracle(sys_[1.0]), Oracle(sys_[1.7])
    pos, _ = o1.positions()
    for j in range(0, base.nsites, 3):
        assert o1.interact(pos[j], 0, j, False) == o17.interact(pos[j], 0, j, False)


def knot_cases(s, every=23, rmin=2.5):
    """Two-ion separations exactly on table knots r_k (as EFPOT files list them) and one ulp
    either side, with the energy jump the table has there (fixdy, polynom.f90:91-118):
    a lookup that picks the neighbouring segment is off by that jump."""
    nt1 = s.nttyp + 1
    p = s.pairs[int(s.pairidx[1 * nt1 + 2])]
    r = configs.knot_grid()
    out = []
    for k in range(1, len(r) - 1, every):
        if r[k] < rmin:
            continue
        for x in (r[k], np.nextafter(r[k], 0.0), np.nextafter(r[k], 1e9)):
            out.append(float(x))
    return p, out


def test_knot_boundaries_vs_reference():
    """Pairs placed on and one ulp around knot boundaries select the reference's segment:
    ENERGY's table energy and force equal the reference's bit for bit (getyd's index from
    the correctly rounded sqrt, polynom.f90:225-226)."""
    s = small_plain(3, 1)
    p, xs = knot_cases(s)
    o, r = Oracle(s), RefFortran(s)
    ep = s.ep.reshape(-1, 3)
    jumps = 0
    for x in xs:
        pos = np.array([[0.0, 0.0, 0.0], [x, 0.0, 0.0]])
        o.set_positions(pos)
        r.set_positions(pos)
        eo, er = o.energy(), r.energy()
        assert eo["eefpot"] == er["eefpot"], x
        assert np.array_equal(o.forces(), r.forces()), x
        k = int((x - p["dmi"]) * (1.0 / s.res))
        a, b = ep[p["nxi"] + k], ep[p["nxi"] + k - 1]
        jumps += abs((a[0] + x * (a[1] + x * a[2])) - (b[0] + x * (b[1] + x * b[2]))) > 1e-12 * abs(eo["eefpot"])
    assert jumps > len(xs) // 2     # the neighbouring segment would be seen

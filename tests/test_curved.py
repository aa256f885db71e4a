"""Spherical and elliptic-cylinder systems: the C oracle against the reference's own FIXCOOR
(fixcoor.f90:24-46), OUTBOX (grandmod.f90:74-97, through DEBYHYB), MOVE (move.f90:32-47)
and INSERT (insert.f90:29-41), compiled from its sources (oracle/_ref); and closed forms
of FIXCOOR: a particle leaving the sphere at distance r > R comes back along its radius
at 2R - r, one leaving the ellipse lands on it scaled by 0.999999999999999 with z clamped.
Bars as for the box: positions 1e-12 A, energies 1e-14 relative, GCMC decisions exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from oracle.oracle import Oracle, RefFortran, oracle_table_builder, ref_available

from helpers import energy_errs, force_err, with_fixed

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def plain(seed=4, n=50):
    return configs.plain_ions(seed=seed, n_per_type=n, table_builder=oracle_table_builder)


def pore(seed=7, n_bp=6, n_ions=100):
    return configs.pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=(40.0, 40.0, 60.0),
                            dcel=0.5, rpore=12.0, half_slab=10.0, dna=True,
                            table_builder=oracle_table_builder)


def shaped(which, base):
    if which == "sphere":
        lx2m, lx2p = base.box[0], base.box[1]
        return configs.with_sphere(base, 0.5 * (lx2p - lx2m), center=(0.0, 0.0, 0.0), seed=2)
    return configs.with_ecyl(base, seed=2)


def _single_ion(which, start, g):
    """One moving ion (no pair partner, no drift): one step with injected noise g."""
    s = plain(seed=1, n=1)
    s = s.copy()
    s.ntot = s.capacity = 1
    s.x, s.y, s.z = np.array([start[0]]), np.array([start[1]]), np.array([start[2]])
    s.typei = s.typei[:1].copy()
    s.fact1a = np.zeros_like(s.fact1a)
    s.fact2a = np.ones_like(s.fact2a)
    s.flags &= ~abi.QCOUNTION
    s = shaped(which, s) if which == "sphere" else configs.with_ecyl(s)
    s.x, s.y, s.z = np.array([start[0]]), np.array([start[1]]), np.array([start[2]])
    o = Oracle(s)
    o.upload_noise(np.array([g], float))
    o.step(1, abi.NOISE_INJECTED, 0)
    return s, o.positions()[0][0]


@pytest.mark.parametrize("g", [(40.0, 3.0, -2.0), (12.0, 30.0, 5.0), (-31.0, 0.0, 0.5)])
def test_sphere_reflection_closed_form(g):
    s, p = _single_ion("sphere", (1.0, -2.0, 0.5), g)
    q = np.array([1.0, -2.0, 0.5]) + np.array(g)
    c = np.array(s.center)
    r = np.linalg.norm(q - c)
    assert r > s.rsphe
    want = c + (2.0 * s.rsphe / r - 1.0) * (q - c)
    assert np.allclose(p, want, rtol=0, atol=1e-12)
    assert abs(np.linalg.norm(p - c) - (2.0 * s.rsphe - r)) <= 1e-12


@pytest.mark.parametrize("g", [(40.0, 3.0, -2.0), (1.0, 45.0, 80.0), (-50.0, -1.0, -90.0)])
def test_ecyl_wall_closed_form(g):
    s, p = _single_ion("ecyl", (1.0, -2.0, 0.5), g)
    q = np.array([1.0, -2.0, 0.5]) + np.array(g)
    cx, cy, _ = s.center
    e = np.hypot((p[0] - cx) * s.iecx, (p[1] - cy) * s.iecy)
    assert abs(e - 0.999999999999999) <= 1e-15
    # the ray from the axis keeps its direction in x-y
    d0, d1 = np.array([q[0] - cx, q[1] - cy]), np.array([p[0] - cx, p[1] - cy])
    assert abs(d0[0] * d1[1] - d0[1] * d1[0]) <= 1e-10 * np.linalg.norm(d0) * np.linalg.norm(d1)
    assert p[2] == min(max(q[2], s.box[4]), s.box[5])


@needs_ref
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
@pytest.mark.parametrize("base", ["plain", "pore"])
def test_steps_vs_reference(which, base):
    s = shaped(which, plain() if base == "plain" else with_fixed(pore(), 5, nonfree=(2,)))
    noise = 3.0 * np.random.default_rng(3).standard_normal((6, 3 * s.ntot))
    o, r = Oracle(s), RefFortran(s)
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(6):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    po, pr = o.positions()[0], r.positions()[0]
    assert float(np.max(np.abs(po - pr))) <= 1e-12
    for a, b in zip(o.flux(), r.flux()):
        assert np.array_equal(a, b)
    # the noise carried some particles out: every moving one is back inside
    c = np.array(s.center)
    mv = po[s.nsites + s.nfix:]
    if which == "sphere":
        assert np.all(np.linalg.norm(mv - c, axis=1) <= s.rsphe * (1 + 1e-15))
    else:
        assert np.all(np.hypot((mv[:, 0] - c[0]) * s.iecx, (mv[:, 1] - c[1]) * s.iecy) <= 1.0)


@needs_ref
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_notrans_after_fixcoor_vs_reference(which):
    """NOTRANS resets the DNA centroid after DYNAMICS0NUC's FIXCOOR (simul1.f90:369-386):
    the curved wall acts on the sites first."""
    s = configs.with_notrans(shaped(which, pore(seed=12)), axes=7)
    noise = 3.0 * np.random.default_rng(13).standard_normal((5, 3 * s.ntot))
    o, r = Oracle(s), RefFortran(s)
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(5):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    assert float(np.max(np.abs(o.positions()[0] - r.positions()[0]))) <= 1e-12


@needs_ref
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_debyhyb_outbox_vs_reference(which):
    base = shaped(which, pore(seed=9))
    s = configs.with_debyhyb(base, outside=())
    # every P site beyond the curved wall, inside the bounding box's corner column
    s.x, s.y, s.z = s.x.copy(), s.y.copy(), s.z.copy()
    pk = np.flatnonzero(np.asarray(s.namsite) == abi.SITE_P)
    assert len(pk) >= 4
    for k, i in enumerate(pk):
        s.x[i], s.y[i], s.z[i] = 18.0 - 0.05 * k, 18.0, -14.0 + 28.0 * k / len(pk)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14
    # those sites are inside the bounding box: as a box system they would not be screened
    b = s.copy()
    b.flags &= ~(abi.QSPHERE | abi.QECYL)
    assert abs(Oracle(b).energy()["eqq"] - eo["eqq"]) > 1e-6 * abs(eo["eqq"])


@needs_ref
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_grand_vs_reference(which):
    s = shaped(which, configs.pore_dna(seed=10, n_bp=0, n_ions=80, box=(40.0, 40.0, 60.0),
                                       dcel=0.5, rpore=12.0, half_slab=10.0, dna=False,
                                       table_builder=oracle_table_builder)).copy()
    s.capacity = s.ntot + 150
    if which == "sphere":
        gc = configs.sphere_buffers(s, conc=1.0, shell_len=6.0, mu=(-0.5, -0.4))
    else:
        gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    assert nat[1:].sum() > 0
    o, r = Oracle(s), RefFortran(s)
    o.set_gcmc(gc, ib)
    r.set_gcmc(gc, ib)
    u = np.random.default_rng(4).random(20000)
    pa = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    pb = pa.copy()
    na, nb = nat.copy(), nat.copy()
    ua, ea = o.grand(100, u, na, pa, 1, 0.0)
    ub, eb = r.grand(100, u, nb, pb, 1, 0.0)
    assert ua == ub and np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert ua > 6 * 100 * 0.5        # INSERT redrew: more than the box's draws per attempt
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert o.ntot == r.ntot
    assert np.array_equal(o.positions()[0], r.positions()[0])
    assert np.array_equal(o.ibuffer(), r.ibuffer())


@needs_ref
@pytest.mark.parametrize("which", ["sphere", "ecyl"])
def test_metropolis_vs_reference(which):
    s = shaped(which, with_fixed(pore(seed=36, n_bp=5, n_ions=90), 6))
    o, r = Oracle(s), RefFortran(s)
    u = np.random.default_rng(9).random(40000)
    ikbt = 1.0 / (0.0019872 * 300.0)
    ua, da, ea = o.metropolis(300, 6.0, ikbt, u, 1.5)
    ub, db, eb = r.metropolis(300, 6.0, ikbt, u, 1.5)
    assert da == db == 300 and ua == ub
    assert abs(ea - eb) <= 1e-12 * max(abs(eb), 1.0)
    assert np.array_equal(o.positions()[0], r.positions()[0])


def test_validation():
    from oracle.oracle import oracle_lib
    s = plain(seed=1, n=5)
    t = configs.with_sphere(s, 20.0)
    t.flags |= abi.QECYL
    with pytest.raises(Exception):
        Oracle(t)
    t = configs.with_sphere(s, 20.0)
    t.rsphe = 0.0
    with pytest.raises(Exception):
        Oracle(t)

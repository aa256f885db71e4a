"""Per-species repulsion maps (NMCDEN: vdwgd1trln.f90:61-80, vdwgd1spln.f90:61-80 in
ENERGY, vdwgd0trln.f90:56-75, vdwgd0spln.f90:58-77 in INTERACT): a DNA site reads the map of
its site name, a moving ion the map of its type, a fixed ion map 0.

Pins: (1) with every map equal to the system's one map, energies and forces are those of
the single-map system; (2) each particle's force and INTERACT energy equal those of the
single-map system whose map is that particle's species map (the repulsion term is
per-particle, so only the map choice can differ), which fixes the site order S, P, Ab, Tb,
Cb, Gb and the ion-type numbering; (3) the reference's own routines (oracle/_ref; skipped
where not built) for energies, forces, INTERACT and BD steps.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.system import GridData
from oracle.oracle import Oracle, RefFortran, ref_available

from helpers import energy_errs, force_err, pore_small, with_fixed

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def _nmcden(seed, dna=True, bspline=False, nfix=3):
    s = with_fixed(pore_small(seed=seed, n_ions=120, dna=dna), nfix)
    if bspline:
        s.flags |= abi.QBSPLINE
    return configs.with_nmcden(s, seed=seed, density=0.2)


def _species(s, i):
    if (s.flags & abi.QNUCL) and i < s.nsites:
        return int(s.namsite[i])
    if i < s.nsites + s.nfix:
        return 0
    return (6 if s.flags & abi.QNUCL else 0) + abs(int(s.typei[i])) - s.nold - 1


def _single(s, k):
    t = s.copy()
    t.flags &= ~abi.QNMCDEN
    g = s.phiv
    t.phiv = GridData(np.ascontiguousarray(g.data[k]), g.dcel, g.xbcen, g.ybcen, g.zbcen)
    return t


@pytest.mark.parametrize("dna,bspline", [(True, False), (True, True), (False, False)])
def test_identical_maps_reduce_to_one_map(dna, bspline):
    s = _nmcden(11, dna, bspline)
    base = _single(s, 0)
    t = s.copy()
    t.phiv = GridData(np.stack([base.phiv.data] * s.phiv.data.shape[0]), s.phiv.dcel)
    a, b = Oracle(t), Oracle(base)
    ea, eb = a.energy(), b.energy()
    assert all(ea[k] == eb[k] for k in ea)
    assert np.array_equal(a.forces(), b.forces())


@pytest.mark.parametrize("dna,bspline", [(True, False), (True, True), (False, False)])
def test_each_particle_reads_its_species_map(dna, bspline):
    s = _nmcden(12, dna, bspline)
    nmaps = s.phiv.data.shape[0]
    assert nmaps == (6 if dna else 0) + s.ntype - s.nold
    o = Oracle(s)
    o.energy()
    f, pos = o.forces(), o.positions()[0]
    assert np.max(np.abs(f)) > 0.0
    kinds = np.array([_species(s, i) for i in range(s.ntot)])
    seen = 0
    for k in range(nmaps):
        idx = np.nonzero(kinds == k)[0]
        if not len(idx):
            continue
        r = Oracle(_single(s, k))
        r.energy()
        fk = r.forces()
        move = [i for i in idx if i < s.nsites or i >= s.nsites + s.nfix]
        assert np.array_equal(f[move], fk[move]), k
        for j in idx[:4]:
            jt = 0 if j < s.nsites else abs(int(s.typei[j]))
            assert o.interact(pos[j], jt, j, True) == r.interact(pos[j], jt, j, True), (k, j)
        seen += len(move)
    assert seen == s.ntot - s.nfix
    # a species map that differs changes the particle's force
    last = Oracle(_single(s, nmaps - 1))
    last.energy()
    other = last.forces()
    assert not np.array_equal(f[kinds != nmaps - 1], other[kinds != nmaps - 1])
    # a trial insertion reads its type's map
    for jt in range(s.nold + 1, s.ntype + 1):
        k = (6 if dna else 0) + jt - s.nold - 1
        xyz = (1.5, -2.0, 14.0)
        assert o.interact(xyz, jt, s.ntot, True) == Oracle(_single(s, k)).interact(
            xyz, jt, s.ntot, True)


@needs_ref
@pytest.mark.parametrize("dna,bspline,mode", [(True, False, "scatter"), (True, True, "scatter"),
                                              (False, False, "scatter"), (False, True, "scatter"),
                                              (True, False, "pore"), (False, True, "pore")])
def test_matches_reference(dna, bspline, mode):
    s = _nmcden(13, dna, bspline)
    if mode == "pore":
        s = configs.with_nmcden(with_fixed(pore_small(seed=13, n_ions=120, dna=dna), 3),
                                seed=13, mode="pore")
        if bspline:
            s.flags |= abi.QBSPLINE
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    assert eo["evdwgd"] != 0.0
    assert max(energy_errs(eo, er, floor=1e-12).values()) <= 1e-14
    assert force_err(o.forces(), r.forces()) <= 1e-14
    pos, typ = o.positions()
    js = list(range(0, s.nsites, 3)) + list(range(s.nsites, s.ntot, 7))
    for j in js:
        jt = abs(int(typ[j])) if j >= s.nsites else 0
        a, b = o.interact(pos[j], jt, j, True), r.interact(pos[j], jt, j, True)
        assert abs(a - b) <= 1e-13 * max(abs(b), 1.0), (j, a, b)
    for jt in range(s.nold + 1, s.ntype + 1):
        a, b = o.interact((0.5, 1.0, 12.0), jt, s.ntot, True), r.interact((0.5, 1.0, 12.0), jt,
                                                                         s.ntot, True)
        assert abs(a - b) <= 1e-13 * max(abs(b), 1.0), (jt, a, b)
    noise = np.random.default_rng(4).standard_normal((4, 3 * s.ntot))
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(4):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    assert float(np.max(np.abs(o.positions()[0] - r.positions()[0]))) <= 1e-12

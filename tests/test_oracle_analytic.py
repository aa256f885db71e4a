"""The analytic pair branch -- Coulomb, 12-6 Lennard-Jones and the SRPMF short-range term of
pair types without an EFPOT table (energy.mono.f90:135-178 ion-ion, :623-668 site-ion;
interact.f90:182-204, :365-404) -- in the C oracle, pinned three ways:

* against the reference's own routines (oracle/_ref): ENERGY terms and forces, INTERACT,
  BD trajectories, on all-analytic, mixed tabulated/analytic and site-ion systems;
* against closed forms on two particles: the energy formulas as shell_simul.f90:1737-1742
  states them for its SRPMF energy file (LJ as epp4 r6 (r6 - 1), SRPMF with the uninverted
  c2), the smoothing that takes the SRPMF energy to zero at r = rth, and F = -dE/dr by
  central differences;
* INTERACT's SRPMF smoothing, e0 exp(-k (r - srpx)) - srpy (interact.f90:200), against
  ENERGY's e0 (exp(-k (r - srpx)) - srpy) (energy.mono.f90:154-158): equal below srpx,
  different by srpy (e0 - 1) inside the smoothing band.
"""
import math

import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.system import CELEC, PI
from oracle.oracle import Oracle, RefFortran, oracle_table_builder, ref_available

from helpers import energy_errs, force_err, with_fixed

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def dense_plain(seed, n=40, box=24.0):
    return configs.plain_ions(seed=seed, n_per_type=n, box=box,
                              table_builder=oracle_table_builder)


def dense_pore(seed=5, n_bp=6, n_ions=120):
    return configs.pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=(36.0, 36.0, 48.0),
                            dcel=0.5, rpore=12.0, half_slab=8.0,
                            table_builder=oracle_table_builder)


def variant(name):
    """The analytic systems the tests share: every ion pair analytic, a mixed system (one
    tabulated and two analytic types), LJ alone, Coulomb + SRPMF without LJ, site-ion pairs
    analytic next to tabulated ion-ion pairs, and everything analytic with DNA."""
    if name == "all":
        return with_fixed(configs.with_analytic(dense_plain(1), seed=11), 6)
    if name == "mixed":
        return configs.with_analytic(dense_plain(2), types=[1, 3], seed=12)
    if name == "lj":
        return configs.with_analytic(dense_plain(3), seed=13, srpmf=False, coulomb=False)
    if name == "coulomb_srpmf":
        return configs.with_analytic(dense_plain(4), seed=14, lj=False)
    s = dense_pore()
    if name == "site":
        t = configs.with_analytic(s, types=configs.analytic_pair_types(s, "site"), seed=15)
        return with_fixed(t, 5, nonfree=(0, 4, 9))
    if name == "pore_all":
        return with_fixed(configs.with_analytic(s, seed=16), 3)
    raise KeyError(name)


VARIANTS = ["all", "mixed", "lj", "coulomb_srpmf", "site", "pore_all"]


def test_variants_exercise_every_band():
    """Each system has analytic pairs below srpx, inside the smoothing band srpx..rth and
    beyond rth, so every branch of the SRPMF term runs."""
    for name in VARIANTS:
        s = variant(name)
        pos = np.stack([s.x, s.y, s.z], 1)
        d = np.sqrt(((pos[:, None] - pos[None]) ** 2).sum(-1))[np.triu_indices(s.ntot, 1)]
        rth = 8.0
        assert (d < rth - 0.5).sum() > 50, name
        assert ((d >= rth - 0.5) & (d <= rth)).sum() > 20, name
        assert (d > rth).sum() > 100, name
        assert any(p is not None and not p["efpot"] for p in s.pairs), name


@needs_ref
@pytest.mark.parametrize("name", VARIANTS)
def test_energy_forces_vs_reference(name):
    s = variant(name)
    o, r = Oracle(s), RefFortran(s)
    eo, er = o.energy(), r.energy()
    errs = energy_errs(eo, er, floor=1e-12)
    assert max(errs.values()) <= 1e-14, errs
    assert force_err(o.forces(), r.forces()) <= 1e-14
    cat = ("eqqmx", "evdwmx", "esrpmfmx") if name == "site" else ("eelec", "evdw", "esrpmf")
    if name not in ("lj",):
        assert all(er[k] != 0.0 for k in cat if not (name == "coulomb_srpmf" and k == "evdw"))


@needs_ref
@pytest.mark.parametrize("name", VARIANTS)
def test_interact_vs_reference(name):
    s = variant(name)
    o, r = Oracle(s), RefFortran(s)
    pos, typ = o.positions()
    cases = [(pos[j] + 0.25, abs(int(typ[j])) if j >= s.nsites else 0, j)
             for j in range(0, s.ntot, 7)]
    cases += [((0.5, -1.0, 2.0), jt, s.ntot) for jt in range(s.nold + 1, s.ntype + 1)]
    for xyz, jt, j in cases:
        a, b = o.interact(xyz, jt, j, True), r.interact(xyz, jt, j, True)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1.0), (name, j, a, b)


@needs_ref
@pytest.mark.parametrize("name", ["all", "mixed", "site"])
def test_dynamics_vs_reference(name):
    s = variant(name)
    noise = np.random.default_rng(7).standard_normal((5, 3 * s.ntot))
    o, r = Oracle(s), RefFortran(s)
    o.upload_noise(noise)
    r.upload_noise(noise)
    for k in range(5):
        o.step(1, abi.NOISE_INJECTED, k)
        r.step(1, abi.NOISE_INJECTED, k)
    assert float(np.max(np.abs(o.positions()[0] - r.positions()[0]))) <= 1e-12


# ---------------------------------------------------------------- closed forms
def two_ions(r, seed=21, lj=True, srpmf=True, coulomb=True):
    s = configs.plain_ions(seed=seed, n_per_type=1, box=40.0, table_builder=oracle_table_builder)
    s = configs.with_analytic(s, seed=seed, lj=lj, srpmf=srpmf, coulomb=coulomb)
    s.x[:] = (0.0, r)
    s.y[:] = 0.0
    s.z[:] = 0.0
    return s


def closed_form(s, r):
    """Energy of the two ions' pair at distance r by the formulas of shell_simul.f90:
    eelec = fct / r, evdw = epp4 r6 (r6 - 1) with r6 = (sigma^2 / r^2)^3, and the SRPMF
    c0 exp((c1 - r) / c2) cos(c3 pi (c1 - r)) + c4 (c1 / r)^6 with the input c2 (the
    stored value is its inverse); beyond srpx ENERGY multiplies it by
    exp(-srpk (r - srpx)) - srpy, and it is zero beyond rth."""
    is_ = int(s.pairidx[1 * (s.nttyp + 1) + 2])
    p = s.pairs[is_]
    eel = p["fct"] / r if p["chr"] else 0.0
    r6 = (p["sgp2"] / (r * r)) ** 3
    evd = p["epp4"] * r6 * (r6 - 1.0) if p["lj"] else 0.0
    esr, e0 = 0.0, 0.0
    if p["srpmf"] and r * r <= s.rth:
        c2 = 1.0 / p["c2"]
        e0 = (p["c0"] * math.exp((p["c1"] - r) / c2) * math.cos(p["c3"] * PI * (p["c1"] - r))
              + p["c4"] * (p["c1"] / r) ** 6)
        esr = e0
        if r >= s.srpx:
            esr = e0 * (math.exp(-s.srpk * (r - s.srpx)) - s.srpy)
    return eel, evd, esr, e0


@pytest.mark.parametrize("r", [2.3, 3.1, 5.0, 7.2, 7.6, 7.9, 9.5, 14.0])
def test_two_particle_energy_closed_form(r):
    s = two_ions(r)
    e = Oracle(s).energy()
    eel, evd, esr, _ = closed_form(s, r)
    assert e["eelec"] == pytest.approx(eel, rel=1e-13, abs=1e-15)
    assert e["evdw"] == pytest.approx(evd, rel=1e-12, abs=1e-15)
    assert e["esrpmf"] == pytest.approx(esr, rel=1e-12, abs=1e-15)
    assert e["eefpot"] == 0.0
    # the pair's charges are opposite (K+ / Cl-): Coulomb attracts
    assert eel < 0.0


def test_srpmf_vanishes_at_rth():
    """srpk (rth - srpx) = ln 1000, so exp(-srpk (r - srpx)) - srpy reaches 0 at r = rth: the
    smoothed SRPMF energy goes continuously to zero where the term is cut."""
    s = two_ions(7.5)
    rth = math.sqrt(s.rth)
    assert math.exp(-s.srpk * (rth - s.srpx)) - s.srpy == pytest.approx(0.0, abs=1e-12)
    for r in (rth - 1e-6, rth + 1e-6):
        e = Oracle(two_ions(r)).energy()
        assert abs(e["esrpmf"]) < 1e-7


@pytest.mark.parametrize("r", [2.6, 4.4, 6.9, 7.7, 7.95, 11.0])
@pytest.mark.parametrize("which", ["lj", "coulomb", "srpmf", "all"])
def test_force_is_minus_gradient(r, which):
    kw = dict(lj=which in ("lj", "all"), coulomb=which in ("coulomb", "all"),
              srpmf=which in ("srpmf", "all"))
    h = 1e-6
    o = Oracle(two_ions(r, **kw))
    o.energy()
    f = o.forces()
    ep = Oracle(two_ions(r + h, **kw)).energy()["ener"]
    em = Oracle(two_ions(r - h, **kw)).energy()["ener"]
    dedr = (ep - em) / (2 * h)
    # particle 1 sits at +r on x: its force is -dE/dr along x, particle 0 gets the opposite
    scale = max(abs(dedr), 1e-3)
    assert abs(f[1, 0] + dedr) <= 2e-7 * scale, (f[1], dedr)
    assert abs(f[0, 0] - dedr) <= 2e-7 * scale
    assert np.all(f[:, 1:] == 0.0)


@pytest.mark.parametrize("r", [3.0, 7.0, 7.6, 7.9])
def test_interact_srpmf_quirk(r):
    """INTERACT of ion 1 in the two-ion system is its pair energy with INTERACT's own
    smoothing: below srpx it equals ENERGY's pair energy; in the band srpx..rth it is
    e0 exp(-k (r - srpx)) - srpy, which exceeds ENERGY's e0 (exp(...) - srpy) by
    srpy (e0 - 1)."""
    s = two_ions(r)
    o = Oracle(s)
    e = o.energy()
    eel, evd, _, e0 = closed_form(s, r)
    di = o.interact((r, 0.0, 0.0), 2, 1, False)
    energy_pair = e["eelec"] + e["evdw"] + e["esrpmf"]
    if r < s.srpx:
        assert di == pytest.approx(energy_pair, rel=1e-13)
    else:
        quirk = e0 * math.exp(-s.srpk * (r - s.srpx)) - s.srpy
        assert di == pytest.approx(eel + evd + quirk, rel=1e-12, abs=1e-14)
        assert di - energy_pair == pytest.approx(s.srpy * (e0 - 1.0), rel=1e-9, abs=1e-15)
        assert abs(di - energy_pair) > 1e-5

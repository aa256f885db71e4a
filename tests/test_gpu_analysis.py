"""Per-cycle statistics on the device (bromoc_set_analysis / bromoc_analyze /
bromoc_get_analysis) against the oracle's restatement of simul1.f90:365-368, 477-538.

Bar: histograms, irmax and the prob2/prob3 counts exact (integer counters; the positions of
both paths agree to ~1e-12 A, far inside any bin); the pressure sum within 1e-10 of the
sum of the magnitudes of its terms, steps times (ntot + |virial| / (3 kBT)) times
kB T / V.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import pore_small, with_fixed

pytestmark = pytest.mark.gpu
KB = 1.380648813e-23


def _compare(g, o, what):
    a, b = g.get_analysis(), o.get_analysis()
    for k in ("rho", "rdna", "gr", "prob2", "prob3"):
        if b[k] is None:
            assert a[k] is None
            continue
        assert np.array_equal(a[k].astype(float), b[k].astype(float)), k
    if what & abi.AN_GR:
        assert a["irmax"] == b["irmax"]
    return a, b


def _pres_ok(a, b, s, an, vir_bound):
    scale = b["pres_steps"] * (s.ntot + vir_bound * an.ikbt / 3.0) * KB * 1e25 * an.temp / an.tvol
    assert a["pres_steps"] == b["pres_steps"]
    print(f"pres gpu {a['pres']:.12e} oracle {b['pres']:.12e} scale {scale:.3e}")
    assert abs(a["pres"] - b["pres"]) <= 1e-10 * scale


def _cycles(s, what, ncycle, nbd, graphs=False, seed=3, **kw):
    g, o = Engine(s), Oracle(s)
    g.set_analysis(what=what, **kw)
    o.set_analysis(what=what, **kw)
    noise = np.random.default_rng(seed).standard_normal((ncycle * nbd, 3 * s.ntot))
    g.upload_noise(noise)
    o.upload_noise(noise)
    for c in range(ncycle):
        if graphs:
            g.step(nbd, abi.NOISE_INJECTED, c * nbd)
            o.step(nbd, abi.NOISE_INJECTED, c * nbd)
        else:
            for k in range(nbd):
                g.step(1, abi.NOISE_INJECTED, c * nbd + k)
                o.step(1, abi.NOISE_INJECTED, c * nbd + k)
        g.analyze()
        o.analyze()
    err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
    assert err <= 1e-8, err
    return g, o


def test_histograms_pore_dna():
    s = with_fixed(pore_small(seed=21, n_ions=400), 5)
    what = abi.AN_RHO | abi.AN_RDNA | abi.AN_GR | abi.AN_PROB
    g, o = _cycles(s, what, ncycle=4, nbd=3, igr=2, npoints=300, dr=0.1, czmin=-6.0, czmax=6.0,
                   datom=500)
    a, b = _compare(g, o, what)
    assert b["rho"].sum() > 0 and b["rdna"].sum() == 4 * s.nsites and b["prob2"].sum() > 0


def test_pressure_steps_and_graphs():
    s = with_fixed(pore_small(seed=22, n_ions=300), 4)
    what = abi.AN_PRES | abi.AN_RHO
    g, o = _cycles(s, what, ncycle=2, nbd=40, graphs=True, tvol=2.0e5, temp=300.0)
    a, b = _compare(g, o, what)
    o.energy()
    vir = float(np.abs(o.forces() * o.positions()[0]).sum())
    _pres_ok(a, b, s, g._an, 3.0 * vir)
    assert b["pres_steps"] == 80


def test_pressure_large_ion_system():
    s = configs.plain_ions(seed=23, n_per_type=1500, box=90.0)
    g, o = _cycles(s, abi.AN_PRES | abi.AN_GR, ncycle=2, nbd=3, igr=100, npoints=10000)
    a, b = _compare(g, o, abi.AN_PRES | abi.AN_GR)
    o.energy()
    vir = float(np.abs(o.forces() * o.positions()[0]).sum())
    _pres_ok(a, b, s, g._an, 3.0 * vir)


def test_membrane_channel_prob():
    s = with_fixed(pore_small(seed=24, n_ions=200), 3)
    s.flags &= ~(abi.QPHIV | abi.QPHIX)
    s.phiv = s.phix = None
    s = configs.with_membrane(s, thick=16.0, zmemb=1.0, volts=0.1, debye=9.0,
                              pore=(11.0, 5.0, 2.5))
    rng = np.random.default_rng(6)
    ion = np.arange(s.ntot) >= s.nsites + s.nfix
    r = rng.uniform(0.0, 14.0, ion.sum())
    th = rng.uniform(0, 2 * np.pi, ion.sum())
    s.x[ion], s.y[ion] = r * np.cos(th), r * np.sin(th)
    s.z[ion] = np.where(rng.random(ion.sum()) < 0.5, rng.uniform(-6.0, 8.0, ion.sum()), s.z[ion])
    g, o = _cycles(s, abi.AN_PROB, ncycle=3, nbd=2, datom=300)
    a, b = _compare(g, o, abi.AN_PROB)
    assert b["prob2"][:, 1:].sum() > 0


def test_dynamics_adds_no_pressure_and_reset():
    s = pore_small(seed=25, n_ions=60)
    g = Engine(s)
    g.set_analysis(what=abi.AN_PRES)
    noise = np.random.default_rng(7).standard_normal((4, 3 * s.ntot))
    g.upload_noise(noise)
    g.energy()
    g.dynamics(abi.NOISE_INJECTED, 0)
    assert g.get_analysis()["pres_steps"] == 0
    g.step(3, abi.NOISE_INJECTED, 1)
    assert g.get_analysis()["pres_steps"] == 3
    g.set_analysis(what=abi.AN_PRES)
    a = g.get_analysis()
    assert a["pres_steps"] == 0 and a["pres"] == 0.0

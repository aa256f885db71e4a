"""EFPOT table construction: SPLINEPOT (splinepot.f90:19-96) and DISCRETIZE
(discretize.f90:19-120), both through squarespline (polynom.f90:20-169), by the library
(host side of libbromoc_gpu.so), by the C oracle and by the reference's own routines
(oracle/_ref, skipped where not built), on randomized inputs.

Bar: bit-identical knots and wall/tail coefficients.  Integer powers must follow the
reference compiler's expansion (exponents up to 8 multiplied out in sequence, larger ones
by binary powering) and sin/cos stay separate calls: either difference shows up as
one-ulp mismatches in a fraction of these cases.  DISCRETIZE is also pinned by what its
definition fixes: the wall meets the table at the first knot with the sampled slope, the
tail meets the last sample, and the table passes near the samples between.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import discretize
from oracle.oracle import (oracle_discretize, oracle_table_builder, ref_available,
                           ref_discretize, ref_table_builder)

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")
ENDS = ("sc1", "sc2", "sc3", "dmi", "dm2lo", "dm2hi")


def _samples(rng):
    lo = rng.uniform(1.0, 3.0)
    res = float(rng.choice([0.01, 0.02, 0.05, 0.1]))
    n = int(rng.integers(20, 400))
    r = lo + res * np.arange(n)
    q = rng.uniform(-2, 2)
    u = configs.efpot_curve(r, rng.uniform(3, 4.5), rng.uniform(0.1, 0.5), q, 80.0)
    return r, u, configs.CELEC * q / 80.0


def _pair(rng, srpmf=None, chr_=None):
    p = abi.Pair()
    p.chr = int(rng.integers(0, 2)) if chr_ is None else chr_
    p.lj = 1
    p.srpmf = int(rng.integers(0, 2)) if srpmf is None else srpmf
    p.fct = rng.uniform(-5, 5)
    p.epp4 = rng.uniform(0.1, 2)
    p.sgp2 = rng.uniform(3, 4.5) ** 2
    p.c0, p.c1, p.c2 = rng.uniform(-0.5, 0.5), rng.uniform(2.5, 3.5), rng.uniform(0.2, 1)
    p.c3, p.c4 = rng.uniform(0.5, 2), rng.uniform(0, 0.2)
    res = float(rng.choice([0.01, 0.02, 0.05]))
    nnp = int(rng.integers(100, 200))
    mnp = nnp + int(rng.integers(50, 600))
    return p, (nnp, mnp, res, rng.uniform(20, 150), rng.uniform(3, 8), rng.uniform(0.1, 1),
               rng.uniform(0, 0.2))


@needs_ref
def test_splinepot_random_bitexact():
    rng = np.random.default_rng(21)
    for _ in range(50):
        r, u, fct = _samples(rng)
        chr_ = int(rng.integers(0, 2))
        e_ref, p_ref = ref_table_builder(len(r), r, u, chr_, fct)
        for build in (configs.lib_table_builder, oracle_table_builder):
            e, p = build(len(r), r, u, chr_, fct)
            assert np.array_equal(e, e_ref)
            assert all(getattr(p, f) == getattr(p_ref, f) for f in ENDS)


@needs_ref
def test_discretize_random_bitexact():
    rng = np.random.default_rng(22)
    for _ in range(60):
        p, args = _pair(rng)
        e_ref, s_ref = ref_discretize(p, *args)
        for build in (discretize, oracle_discretize):
            e, sc = build(p, *args)
            assert np.array_equal(e, e_ref)
            assert sc == s_ref


@pytest.mark.parametrize("srpmf,chr_", [(0, 0), (0, 1), (1, 1)])
def test_discretize_ends_and_samples(srpmf, chr_):
    rng = np.random.default_rng(23 + srpmf + 2 * chr_)
    p, (nnp, mnp, res, rth, srpx, srpk, srpy) = _pair(rng, srpmf=srpmf, chr_=chr_)
    ep, (sc1, sc2, sc3) = oracle_discretize(p, nnp, mnp, res, rth, srpx, srpk, srpy)
    x = np.arange(nnp, mnp + 1) * res
    kf = p.fct if chr_ else 0.0
    # the sampled potential: Coulomb + LJ (+ SRPMF inside rth with its smoothing)
    u = kf / x + p.epp4 * ((p.sgp2 / x ** 2) ** 6 - (p.sgp2 / x ** 2) ** 3)
    du = -kf / x ** 2 + p.epp4 * (6.0 / x) * ((p.sgp2 / x ** 2) ** 3 - 2 * (p.sgp2 / x ** 2) ** 6)
    if srpmf:
        m = x ** 2 <= rth
        a = p.c1 - x
        c0 = p.c0 * np.exp(a * p.c2) * np.cos(p.c3 * np.pi * a) + p.c4 * (p.c1 / x) ** 6
        dc = ((-p.c2 * np.cos(p.c3 * np.pi * a) + p.c3 * np.pi * np.sin(p.c3 * np.pi * a)) * p.c0
              * np.exp(a * p.c2) - 6.0 * p.c4 * (p.c1 / x) ** 6 / x)
        fdf = np.exp(-srpk * (x - srpx)) - srpy
        far = x >= srpx
        dc = np.where(far, dc * fdf - c0 * srpk * (fdf + srpy), dc)
        c0 = np.where(far, c0 * fdf, c0)
        u = u + np.where(m, c0, 0.0)
        du = du + np.where(m, dc, 0.0)
    q = ep.reshape(-1, 3)
    table = q[:, 0] + x * (q[:, 1] + x * q[:, 2])
    # wall: value of the table and the sampled slope at the first knot
    assert abs(sc1 / x[0] ** 12 + sc2 - table[0]) <= 1e-12 * abs(table[0])
    assert abs(-12.0 * sc1 / x[0] ** 13 - du[0]) <= 1e-12 * abs(du[0])
    # tail: kf/r + sc3/r^6 through the last sample
    assert abs(kf / x[-1] + sc3 / x[-1] ** 6 - u[-1]) <= 1e-12 * max(abs(u[-1]), 1e-12)
    # the table follows the samples (squarespline's averaging and end correction keep it
    # within a small fraction of the local variation) away from the SRPMF cut at rth and
    # the smoothing onset at srpx, where the samples jump or kink
    keep = np.ones(len(x), bool)
    if srpmf:
        for edge in (np.sqrt(rth), srpx):
            keep &= np.abs(x - edge) > 4 * res
    scale = np.max(np.abs(u[1:] - u[:-1])) + 1e-12
    assert np.max(np.abs(table - u)[keep]) <= 1e-2 * scale


def test_discretize_argument_errors():
    p = abi.Pair()
    with pytest.raises(abi.BromocError):
        discretize(p, 100, 101, 0.01)           # fewer than three knots
    with pytest.raises(abi.BromocError):
        discretize(p, 0, 100, 0.01)             # r = 0 at the first knot

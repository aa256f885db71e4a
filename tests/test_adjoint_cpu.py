"""CPU checks of the adjoint's building blocks: the chain rule through the table build
(exact: the build is affine in the knot values) and the loss of a trajectory computed
from the reference routines and from the C oracle (identical trajectories)."""
import numpy as np
import pytest

from bromoc_amd import configs
from bromoc_amd.adjoint import RdfSpec, knot_gradients
from oracle.oracle import Oracle, RefFortran, oracle_table_builder, ref_available
from oracle.rdf import rdf_loss, trajectory


def small(seed=0, n=30, **kw):
    return configs.adjoint_system(seed=seed, n_per_type=n, box=30.0,
                                  table_builder=oracle_table_builder, **kw)


def test_table_vjp_is_the_transposed_build():
    """grad_u[k] = <table(e_k) - table(0), grad>: the table coefficients are affine in u."""
    s = small()
    r = s.meta["rgrid"]
    n = len(r)
    rng = np.random.default_rng(1)
    gep = rng.standard_normal(3 * len(s.ep) // 3)
    gsc = rng.standard_normal(3 * (s.npair + 1))
    grads = knot_gradients(s, gep, gsc)
    for is_, (u, chr_, fct) in s.meta["params"].items():
        off = s.pairs[is_]["nxi"]
        ge, gs = gep[3 * off:3 * (off + n)], gsc[3 * is_:3 * is_ + 3]
        e0, p0 = oracle_table_builder(n, r, u, chr_, fct)
        for k in (0, 1, 2, 37, 500, n - 2, n - 1):
            u1 = u.copy()
            u1[k] += 1.0
            e1, p1 = oracle_table_builder(n, r, u1, chr_, fct)
            want = (np.dot(e1 - e0, ge) + (p1.sc1 - p0.sc1) * gs[0] +
                    (p1.sc2 - p0.sc2) * gs[1] + (p1.sc3 - p0.sc3) * gs[2])
            got = grads[is_][k]
            assert abs(got - want) <= 1e-7 * max(abs(want), 1.0), (is_, k, got, want)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")
def test_trajectory_loss_reference_equals_oracle():
    s = small(seed=2, n=25)
    spec = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)], nbins=60)
    noise = np.random.default_rng(3).standard_normal((12, 3 * s.ntot))
    xo = trajectory(Oracle(s), noise, 12)
    xr = trajectory(RefFortran(s), noise, 12)
    assert max(float(np.max(np.abs(a - b))) for a, b in zip(xo, xr)) <= 1e-12
    lo, go = rdf_loss(xo, s, spec)
    lr, gr = rdf_loss(xr, s, spec)
    assert abs(lo - lr) <= 1e-12 * lr and np.max(np.abs(go - gr)) <= 1e-12 * np.max(gr)
    assert lo > 0.0


def g_cube(r, L):
    """Pair-distance density of two points uniform in a cube of side L over that of an
    unbounded ideal gas, for r <= L: the cube's distance density r^2 (4 pi - 6 pi r + 8 r^2
    - r^3) in units of L (Philip, "The probability distribution of the distance between two
    random points in a box", 2007) divided by 4 pi r^2."""
    x = r / L
    return 1.0 - 1.5 * x + (2.0 / np.pi) * x ** 2 - x ** 3 / (4.0 * np.pi)


def test_rdf_normalisation_ideal_gas():
    """The loss's RDF normalisation V / (S P 4 pi r_b^2 sigma sqrt(2 pi)) against the closed
    form of an ideal gas in a box: uniform random states of ions (no forces) give
    g(r) = g_cube(r / L), the cube's edge effect, for every group (like, unlike and with a
    fixed ion).  A normalisation off by a factor (a missing 2 for like pairs, sqrt(2 pi),
    the fixed-partner count) misses by far more than the 3 % statistical bar."""
    s = configs.adjoint_system(seed=5, n_per_type=150, box=40.0, table_builder=oracle_table_builder)
    L = 40.0
    spec = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)], r0=2.0, dr=0.25, nbins=24, sigma=0.15,
                   cut=0.75)
    rng = np.random.default_rng(6)
    states = []
    fixed = np.stack([s.x, s.y, s.z], 1)[: s.nfix]
    for _ in range(300):
        x = rng.uniform(-0.5 * L, 0.5 * L, (s.ntot, 3))
        x[: s.nfix] = fixed
        states.append(x)
    _, g = rdf_loss(states, s, spec)
    rb = spec.r0 + (np.arange(spec.nbins) + 0.5) * spec.dr
    for gi in range(3):      # ion-ion groups: pair counts large enough for a 3 % bar
        want = g_cube(rb, L)
        rel = np.abs(g[gi] - want) / want
        assert np.mean(rel) <= 0.03, (gi, np.mean(rel))
    # the group with the fixed site: the fixed particle sits at the centre, where the
    # density around it is that of an unbounded gas for r < L / 2
    assert np.mean(np.abs(g[3] - 1.0)) <= 0.12


def test_flux_term():
    """J = sum q_i dz_i / (S dt) over the moving ions: one ion of charge q moved by dz gives
    J = q dz / (S dt); the loss adds wflux (J - jtarget)^2."""
    from oracle.rdf import charge_current
    s = small(seed=3, n=5)
    x0 = np.stack([s.x, s.y, s.z], 1)
    x1 = x0.copy()
    k = s.nfix + 2
    x1[k, 2] += 0.75
    q = s.cg[abs(int(s.typei[k]))]
    assert charge_current(x0, x1, s, 10) == pytest.approx(q * 0.75 / (10 * s.dt), rel=1e-15)
    x1[0, 2] += 5.0          # a fixed particle does not count
    assert charge_current(x0, x1, s, 10) == pytest.approx(q * 0.75 / (10 * s.dt), rel=1e-15)
    spec = RdfSpec(groups=[(1, 2)], wflux=2.0, jtarget=0.1)
    l1, _ = rdf_loss([x1], s, spec, x0=x0)
    l0, _ = rdf_loss([x1], s, RdfSpec(groups=[(1, 2)]), x0=x0)
    J = charge_current(x0, x1, s, 1)
    assert l1 - l0 == pytest.approx(2.0 * (J - 0.1) ** 2, rel=1e-12)


@pytest.mark.parametrize("n", [2, 3, 4, 37, 1851])
def test_table_vjp_reverse_sweep_equals_columns(n):
    """bromoc_table_vjp's reverse sweep against the Jacobian built column by column: with no
    charge term the table of unit samples e_k is column k exactly, so grad_u[k] =
    <table(e_k), grad> must agree to rounding (1e-12 norm-wise) for every k, including the
    two- and three-sample special cases of squarespline."""
    import ctypes as C
    from bromoc_amd import abi
    from bromoc_amd.system import ptr
    lib = abi.load_library()
    rng = np.random.default_rng(n)
    r = 1.5 + np.cumsum(rng.uniform(0.5, 1.5, n) * (0.01 if n > 100 else 0.2))
    ge, gs = rng.standard_normal(3 * n), rng.standard_normal(3)
    got = np.zeros(n)
    assert lib.bromoc_table_vjp(n, ptr(r), ptr(ge), ptr(gs), ptr(got)) == 0
    want = np.zeros(n)
    ep = np.zeros(3 * n)
    pair = abi.Pair()
    for k in range(n):
        u = np.zeros(n)
        u[k] = 1.0
        assert lib.bromoc_build_table(n, ptr(r), ptr(u), 0, 0.0, ptr(ep), C.byref(pair)) == 0
        want[k] = ep @ ge + pair.sc1 * gs[0] + pair.sc2 * gs[1] + pair.sc3 * gs[2]
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))

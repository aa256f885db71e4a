"""Tile mode of the BD step (tiles.hip: every unordered pair evaluated once, the reaction-
field cross term in the same sweep, partial sums in a fixed order) against the C oracle,
forced through each of its shapes (BROMOC_TPB 17, 18, 20: 1, 2 or 4 targets per lane, tiles
of 64, 128 or 256 slots).

Covered: ion-only systems from a single pair up to several tiles (one tile pair on the
diagonal only, ragged last tiles), fixed ions, CG-DNA with fixed ions and non-free sites,
the static-field and repulsion grids (trilinear and B-spline, per-species maps), analytic
and mixed pair types, MEMBRANE, the reaction-field grids with their cross term (fixed ions
receive its force), MMIJ, diffusion profiles, BD trajectories (direct and graph-replayed),
GRAND changing the particle count, and bit-for-bit repeatability.

Bar as the other parity tests: energies and norm-wise forces 1e-10 relative, positions
1e-9 A after 5 steps and 1e-8 A after 40, integer outputs exact.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from oracle.oracle import Oracle

from helpers import energy_errs, force_err, near_wall, plain, pore_small, with_fixed

pytestmark = pytest.mark.gpu
TOL = 1e-10
SHAPES = [17, 18, 20]


def _check(g, o):
    eg, eo = g.energy(), o.energy()
    bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
    assert not bad, bad
    fo = o.forces()
    if np.max(np.abs(fo)) > 0.0:
        assert force_err(g.forces(), fo) <= TOL
    assert np.array_equal(g.warn(), o.warn())
    return eg, eo


def _rf_plain(seed, n, nfix=0, **kw):
    s = with_fixed(plain(seed=seed, n_per_type=n), nfix)
    return configs.with_rfpar(s, seed=seed, shape=(45, 45, 45), dcel=1.2, **kw)


SYSTEMS = {
    "pair": lambda: plain(seed=1, n_per_type=1),
    "one_tile": lambda: plain(seed=2, n_per_type=20),
    "ragged": lambda: with_fixed(plain(seed=3, n_per_type=157), 9),
    "pore_dna": lambda: with_fixed(pore_small(seed=9, n_bp=6, n_ions=300), 10,
                                   nonfree=(0, 3, 7, 20)),
    "walls": lambda: near_wall(pore_small(seed=5, n_bp=8, n_ions=260), 40, seed=1),
    "bspline": lambda: configs.with_bspline(near_wall(pore_small(seed=6, n_ions=260), 30)),
    "nmcden": lambda: configs.with_nmcden(with_fixed(pore_small(seed=61, n_ions=260), 3),
                                          seed=61, density=0.2),
    "analytic": lambda: configs.with_analytic(pore_small(seed=7, n_bp=6, n_ions=280),
                                              types=[1, 4, 9], seed=3),
    "membrane": lambda: configs.with_membrane(plain(seed=8, n_per_type=150), pore=(8.0, 6.0, 2.0)),
    "rfpar": lambda: _rf_plain(51, 160, nfix=7),
    "rfpar_walls": lambda: _rf_plain(52, 140, nfix=3, wall_radius=14.0, zero_frac=0.1),
    "rfpar_dna": lambda: configs.with_rfpar(pore_small(seed=52, n_ions=260), seed=3,
                                            shape=(37, 37, 55), dcel=1.0, wall_radius=12.0,
                                            zero_frac=0.05),
    # the same without zero cells: a zero cell makes the grid force discontinuous, and the
    # oracle itself turns a 1e-13 A change of the start into 1e-6 A after 30 steps there
    "rfpar_dna_smooth": lambda: configs.with_rfpar(pore_small(seed=52, n_ions=260), seed=3,
                                                   shape=(37, 37, 55), dcel=1.0,
                                                   wall_radius=12.0, zero_frac=0.0),
    "mmij": lambda: configs.with_mmij(with_fixed(pore_small(seed=51, n_ions=260), 3, nonfree=(1,)),
                                      npol=(3, 4, 5), rfscal=0.9, region=(30.0, 30.0, 44.0),
                                      seed=51),
}


@pytest.mark.parametrize("tpb", SHAPES)
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_energy_forces(name, tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = SYSTEMS[name]()
    g, o = Engine(s), Oracle(s)
    assert g.tuning() == tpb
    _check(g, o)


@pytest.mark.parametrize("tpb", SHAPES)
@pytest.mark.parametrize("name", ["ragged", "pore_dna", "walls", "analytic", "rfpar",
                                  "rfpar_dna_smooth"])
def test_steps_and_graphs(name, tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = SYSTEMS[name]()
    noise = np.random.default_rng(3).standard_normal((40, 3 * s.ntot))
    for nsteps, tol in ((5, 1e-9), (40, 1e-8)):
        g, o = Engine(s), Oracle(s)
        g.upload_noise(noise)
        o.upload_noise(noise)
        eg = g.step(nsteps, abi.NOISE_INJECTED)
        eo = o.step(nsteps, abi.NOISE_INJECTED)
        err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
        print(f"{name} tpb={tpb} {nsteps} steps: position err {err:.3e}")
        assert err <= tol
        bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
        assert not bad, bad
        for a, b in zip(g.flux(), o.flux()):
            assert np.array_equal(a, b)
        assert np.array_equal(g.warn(), o.warn())


@pytest.mark.parametrize("tpb", SHAPES)
def test_profile_philox(tpb, monkeypatch):
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    zk, fk = configs.tanh_profile(-15.0, 15.0, 25, 6.0, low=0.3)
    s = configs.with_profile_diffusion(pore_small(seed=44, n_bp=6, n_ions=250), zk, fk)
    g, o = Engine(s), Oracle(s)
    g.step(70, abi.NOISE_PHILOX, seed=99, counter0=5)
    o.step(70, abi.NOISE_PHILOX, seed=99, counter0=5)
    assert float(np.max(np.abs(g.positions()[0] - o.positions()[0]))) <= 1e-8


@pytest.mark.parametrize("tpb", SHAPES)
def test_grand_then_steps(tpb, monkeypatch):
    """GRAND changes the particle count, so the tile count and the captured graphs change."""
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = pore_small(seed=64, n_bp=0, n_ions=300, dna=False).copy()
    s.capacity = s.ntot + 300
    gc = configs.gcmc_buffers(s, conc=1.0, buffer_len=8.0, mu=(-0.5, -0.4))
    ib, nat = configs.assign_buffers(s, gc)
    g, o = Engine(s), Oracle(s)
    g.set_gcmc(gc, ib)
    o.set_gcmc(gc, ib)
    rng = np.random.default_rng(4)
    natg, nato = nat.copy(), nat.copy()
    probg = np.zeros((gc.nbuffer, s.capacity + 1), np.int32)
    probo = probg.copy()
    eng, eno, ctr = 0.0, 0.0, 0
    for cyc in range(3):
        u = rng.random(6 * 60 + 8)
        _, eng = g.grand(60, u, natg, probg, cyc + 1, eng)
        _, eno = o.grand(60, u, nato, probo, cyc + 1, eno)
        assert g.ntot == o.ntot
        g.step(35, abi.NOISE_PHILOX, seed=3, counter0=ctr)
        o.step(35, abi.NOISE_PHILOX, seed=3, counter0=ctr)
        ctr += 35
        err = float(np.max(np.abs(g.positions()[0] - o.positions()[0])))
        assert err <= 1e-8, (cyc, err)
    assert np.array_equal(natg, nato) and np.array_equal(probg, probo)


@pytest.mark.parametrize("tpb", SHAPES)
def test_repeatable_bit_for_bit(tpb, monkeypatch):
    """Every partial sum has one writer and a fixed summation order: two contexts on the
    same input give identical forces, energies and trajectories."""
    monkeypatch.setenv("BROMOC_TPB", str(tpb))
    s = SYSTEMS["rfpar_dna"]()
    out = []
    for _ in range(2):
        g = Engine(s)
        e = g.energy()
        f = g.forces()
        g.step(33, abi.NOISE_PHILOX, seed=5, counter0=0)
        out.append((e, f, g.positions()[0]))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_autotuner_picks_and_full_size_pore_dna():
    """The 2000-ion + 300-site pore workload at full size in every tile shape."""
    s = configs.pore_dna(seed=1)
    o = Oracle(s)
    eo = o.energy()
    fo = o.forces()
    for tpb in SHAPES:
        import os
        os.environ["BROMOC_TPB"] = str(tpb)
        try:
            g = Engine(s)
            eg = g.energy()
        finally:
            del os.environ["BROMOC_TPB"]
        bad = {k: v for k, v in energy_errs(eg, eo).items() if v > TOL}
        assert not bad, (tpb, bad)
        assert force_err(g.forces(), fo) <= TOL
    print("autotuner pick:", Engine(s).tuning())

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

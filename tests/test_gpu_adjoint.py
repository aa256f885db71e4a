"""The trajectory adjoint on the GPU: the forward loss against the CPU oracle's trajectory,
and the gradient with respect to the EFPOT knot values against central differences of the
loss (h = 1e-6 in the knot values) computed with the GPU forward pass and with the CPU
oracle's forward pass.

Bar: loss and RDFs 1e-10 relative to the CPU path; gradient within 1e-6 norm-wise relative
of the central differences over the sampled knots, max |adjoint - cd| / max |cd| (as the
forces' bar): the difference quotients carry an absolute noise of ~1e-7 from the tables'
1e-9 jumps at knots (fixdy, polynom.f90:91-118), which trajectories of the perturbed
potentials cross, so knots with tiny gradients are not held to a per-knot relative bar.
"""
import numpy as np
import pytest

from bromoc_amd import configs
from bromoc_amd.adjoint import RdfLoss, RdfSpec, rdf_loss, table_gradients, tangent
from oracle.oracle import Oracle, oracle_table_builder
from oracle.rdf import rdf_loss as cpu_rdf_loss, trajectory

pytestmark = pytest.mark.gpu

SPEC = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)])
H = 1e-6


def _system(seed=0, n=60, box=36.0, curve_seed=None):
    return configs.adjoint_system(seed=seed, n_per_type=n, box=box, curve_seed=curve_seed)


def _grids(s, field, seed):
    """s with the static field (field True), or with the static field and a soft repulsion
    map with per-type scaling (field "grids")."""
    if field:
        s = configs.with_static_field(s, seed=seed)
    if field == "grids":
        s = configs.with_soft_repulsion(s, seed=seed + 1)
    return s


def _target(s, noise, nsteps, curve_seed):
    """RDFs of the same start and noise under a second seed's potentials."""
    t = configs.with_knots(s, {is_: u for is_, (u, _, _) in configs.adjoint_system(
        seed=s.meta["seed"], n_per_type=(s.ntot - 1) // 2, box=2 * s.box[1],
        curve_seed=curve_seed).meta["params"].items()})
    return rdf_loss(t, noise, nsteps, SPEC, grad=False)[1]


def _pick(grads, rng, n, rmin=2.5, rmax=14.0, r=None):
    """n (pair type, knot) with r in [rmin, rmax]"""
    keys = list(grads)
    out = []
    while len(out) < n:
        is_ = keys[rng.integers(len(keys))]
        k = int(rng.integers(len(r)))
        if rmin <= r[k] <= rmax and (is_, k) not in out:
            out.append((is_, k))
    return out


def _cd(s, noise, nsteps, target, is_, k, forward, h=H):
    vals = []
    for sg in (1.0, -1.0):
        u = s.meta["params"][is_][0].copy()
        u[k] += sg * h
        vals.append(forward(configs.with_knots(s, {is_: u}, oracle_table_builder)))
    return (vals[0] - vals[1]) / (2 * h)


def _normwise(adj, cd):
    adj, cd = np.asarray(adj), np.asarray(cd)
    return float(np.max(np.abs(adj - cd)) / np.max(np.abs(cd)))


@pytest.mark.parametrize("field", [False, True, "grids"], ids=["pairs", "static_field", "grids"])
def test_forward_loss_matches_cpu(field):
    s = _grids(_system(seed=1, n=40), field, 3)
    nsteps = 30
    noise = np.random.default_rng(2).standard_normal((nsteps, 3 * s.ntot))
    target = np.random.default_rng(3).random((len(SPEC.groups), SPEC.nbins))
    lg, gg, _ = rdf_loss(s, noise, nsteps, SPEC, target=target, grad=False)
    lc, gc = cpu_rdf_loss(trajectory(Oracle(s), noise, nsteps), s, SPEC, target)
    assert abs(lg - lc) <= 1e-10 * lc
    assert np.max(np.abs(gg - gc)) <= 1e-10 * np.max(gc)


@pytest.mark.parametrize("nsteps,ckpt,field", [(60, 20, False), (45, 7, False), (45, 7, True),
                                               (45, 7, "grids")])
def test_gradient_vs_gpu_central_differences(nsteps, ckpt, field):
    s = _grids(_system(seed=4, n=50), field, 5)   # the grids' Jacobian enters the reverse pass
    noise = np.random.default_rng(5).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=77)
    loss, _, grads = rdf_loss(s, noise, nsteps, SPEC, target=target, checkpoint_every=ckpt)
    fwd = lambda t: rdf_loss(t, noise, nsteps, SPEC, target=target, grad=False)[0]
    rng = np.random.default_rng(6)
    adj, cds = [], []
    for is_, k in _pick(grads, rng, 16, r=s.meta["rgrid"]):
        cd = _cd(s, noise, nsteps, target, is_, k, fwd)
        print(f"pair {is_} knot {k}: adjoint {grads[is_][k]:.9e} cd {cd:.9e}")
        adj.append(grads[is_][k])
        cds.append(cd)
    assert _normwise(adj, cds) <= 1e-6, _normwise(adj, cds)


def test_gradient_vs_cpu_central_differences():
    """Central differences of the loss through the CPU oracle's trajectory."""
    s = _system(seed=7, n=30, box=30.0)
    nsteps = 25
    noise = np.random.default_rng(8).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=78)
    _, _, grads = rdf_loss(s, noise, nsteps, SPEC, target=target)
    fwd = lambda t: cpu_rdf_loss(trajectory(Oracle(t), noise, nsteps), t, SPEC, target)[0]
    rng = np.random.default_rng(9)
    adj, cds = [], []
    for is_, k in _pick(grads, rng, 6, r=s.meta["rgrid"]):
        cds.append(_cd(s, noise, nsteps, target, is_, k, fwd))
        adj.append(grads[is_][k])
    assert _normwise(adj, cds) <= 1e-6, _normwise(adj, cds)


def test_config_differentiation_64_knots():
    """The differentiation workload: 200 K+ + 200 Cl- + a fixed wall site in a 60 A box,
    4 x 1851 knot parameters, 400 BD steps of fixed noise checkpointed every 20; target
    RDFs from a second seed's potentials; 64 random knots against central differences.
    Over 400 steps the loss is smooth only between jumps of ~1e-8 (the tables jump by
    ~1e-9 at every knot, fixdy, polynom.f90:91-118, and the perturbed trajectories cross
    knots): a difference quotient whose +-h straddles such a jump is off by jump / 2h,
    1e-3 relative at h = 1e-6 for some knots (one knot here: a jump of 7.8e-9 at
    theta - 7e-8 shifts the quotient by exactly that over h = 1e-7..1e-5, while h <= 5e-8
    agrees to 6e-7).  So each knot is compared at h = 1e-7, then 5e-8, 2e-8 and 1e-8
    (rounding noise ~1e-7 absolute, ~1e-5 relative at h = 1e-8), then 2e-7 and 5e-7 (less
    rounding noise, more chance of a jump) until one agrees within 2e-5, twice that noise
    floor; every knot with a gradient above 1% of the largest must, and the median
    agreement at h = 1e-7 must be within 2e-6.  (The gradient's exactness is the
    dot-product test's to 1e-10; this test checks it against the trajectory itself.)"""
    s = _system(seed=10, n=200, box=60.0)
    nsteps = 400
    noise = np.random.default_rng(11).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=12)
    loss, _, grads = rdf_loss(s, noise, nsteps, SPEC, target=target, checkpoint_every=20)
    assert sum(len(g) for g in grads.values()) == 4 * 1851
    fwd = lambda t: rdf_loss(t, noise, nsteps, SPEC, target=target, grad=False)[0]
    rng = np.random.default_rng(13)
    picks = _pick(grads, rng, 64, r=s.meta["rgrid"])
    adj = np.array([grads[is_][k] for is_, k in picks])
    cd = np.array([_cd(s, noise, nsteps, target, is_, k, fwd, h=1e-7) for is_, k in picks])
    big = np.abs(adj) >= 1e-2 * np.max(np.abs(adj))
    first = np.abs(adj - cd) / np.maximum(np.abs(cd), 1e-300)
    best = first.copy()
    for q in np.nonzero(big & (first > 2e-5))[0]:
        is_, k = picks[q]
        for h in (5e-8, 2e-8, 1e-8, 2e-7, 5e-7):
            c2 = _cd(s, noise, nsteps, target, is_, k, fwd, h=h)
            best[q] = min(best[q], abs(adj[q] - c2) / abs(c2))
            if best[q] <= 2e-5:
                break
    print(f"loss {loss:.6e}; {big.sum()} of 64 knots above 1% of the largest gradient: "
          f"h=1e-7 median {np.median(first[big]):.2e}, 90th pct {np.percentile(first[big], 90):.2e}, "
          f"max {first[big].max():.2e}; best over h: max {best[big].max():.2e}")
    assert np.median(first[big]) <= 2e-6
    assert best[big].max() <= 2e-5, best[big].max()


def test_autograd_function():
    """RdfLoss.backward hands back the adjoint's gradient, bit for bit (the gradient sums
    are integer fixed-point sums, independent of the order the pairs arrive in)."""
    import torch
    s = _system(seed=14, n=30, box=30.0)
    nsteps = 20
    noise = np.random.default_rng(15).standard_normal((nsteps, 3 * s.ntot))
    keys = list(s.meta["params"])
    knots = torch.tensor(np.stack([s.meta["params"][k][0] for k in keys]), requires_grad=True)
    loss = RdfLoss.apply(knots, s, noise, nsteps, SPEC, None)
    loss.backward()
    _, _, grads = rdf_loss(s, noise, nsteps, SPEC)
    want = np.stack([grads[k] for k in keys])
    assert np.array_equal(knots.grad.numpy(), want)


def test_gradient_bit_reproducible():
    """Two reverse passes over the same trajectory give identical gradients."""
    s = _system(seed=16, n=80, box=40.0)
    nsteps = 40
    noise = np.random.default_rng(17).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=79)
    a = table_gradients(s, noise, nsteps, SPEC, target=target)
    b = table_gradients(s, noise, nsteps, SPEC, target=target)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.count_nonzero(a[1]) > 100


FLUX = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)], wflux=3.0, jtarget=0.05)


@pytest.mark.parametrize("field", [False, True, "grids"], ids=["pairs", "static_field", "grids"])
@pytest.mark.parametrize("spec", [SPEC, FLUX], ids=["rdf", "rdf+flux"])
@pytest.mark.parametrize("nsteps,ckpt", [(60, 20), (45, 7)])
def test_dot_product_adjoint_vs_tangent(spec, nsteps, ckpt, field):
    """<dL/dth, v> from the reverse pass equals the tangent-linear pass's dL along v, for
    random directions over every table coefficient (ep2, ep3 of every knot; sc1, sc3 of
    every pair type): the two passes are transposes of one linear map, so they agree to
    rounding -- free of the knot-jump noise central differences carry."""
    s = _grids(_system(seed=18, n=60, box=36.0), field, 19)
    noise = np.random.default_rng(19).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=80)
    loss, gep, gsc = table_gradients(s, noise, nsteps, spec, target=target, checkpoint_every=ckpt)
    rng = np.random.default_rng(20)
    for trial in range(3):
        ve = rng.standard_normal(gep.shape)
        vs = rng.standard_normal(gsc.shape) * np.abs(s.ep).max() * 1e-3
        if trial == 2:               # the wall and tail coefficients alone
            ve[:] = 0.0
            vs = rng.standard_normal(gsc.shape)
        ve[0::3] = 0.0               # ep1 and sc2 do not move the forces
        vs[1::3] = 0.0
        lt, dl = tangent(s, noise, nsteps, spec, ve, vs, target=target, checkpoint_every=ckpt)
        dot = float(np.dot(gep, ve) + np.dot(gsc, vs))
        print(f"{spec.wflux} {nsteps}/{ckpt} trial {trial}: adjoint {dot:.15e} tangent {dl:.15e}")
        assert lt == loss
        assert abs(dot - dl) <= 1e-10 * max(abs(dl), 1e-300), (dot, dl)


def test_dot_product_config_workload():
    """The dot-product test on the 400-step, 401-particle differentiation workload."""
    s = _system(seed=10, n=200, box=60.0)
    nsteps = 400
    noise = np.random.default_rng(11).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=12)
    loss, gep, gsc = table_gradients(s, noise, nsteps, FLUX, target=target)
    ve = np.random.default_rng(21).standard_normal(gep.shape)
    ve[0::3] = 0.0
    vs = np.zeros_like(gsc)
    lt, dl = tangent(s, noise, nsteps, FLUX, ve, vs, target=target)
    dot = float(np.dot(gep, ve))
    print(f"400 steps: adjoint {dot:.15e} tangent {dl:.15e}")
    assert abs(dot - dl) <= 1e-10 * abs(dl)


def test_flux_loss_vs_cpu_and_central_differences():
    """The flux term: the GPU forward loss equals the CPU restatement's on the oracle's
    trajectory, and its knot gradient agrees with central differences of the GPU loss."""
    s = _system(seed=22, n=40, box=30.0)
    nsteps = 30
    noise = np.random.default_rng(23).standard_normal((nsteps, 3 * s.ntot))
    spec = RdfSpec(groups=[(1, 1)], wflux=50.0, jtarget=0.0)
    lg, _, grads = rdf_loss(s, noise, nsteps, spec)
    x0 = np.stack([s.x, s.y, s.z], 1)
    lc, _ = cpu_rdf_loss(trajectory(Oracle(s), noise, nsteps), s, spec, None, x0=x0)
    assert abs(lg - lc) <= 1e-10 * lc
    fwd = lambda t: rdf_loss(t, noise, nsteps, spec, grad=False)[0]
    rng = np.random.default_rng(24)
    adj, cds = [], []
    for is_, k in _pick(grads, rng, 8, r=s.meta["rgrid"]):
        cds.append(_cd(s, noise, nsteps, None, is_, k, fwd))
        adj.append(grads[is_][k])
    assert _normwise(adj, cds) <= 1e-6, _normwise(adj, cds)

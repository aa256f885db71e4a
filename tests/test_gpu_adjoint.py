"""The trajectory adjoint on the GPU: the forward loss against the CPU oracle's trajectory,
and the gradient with respect to the EFPOT knot values against central differences of the
loss (h = 1e-6 in the knot values) computed with the GPU forward pass and with the CPU
oracle's forward pass.

Bar: loss and RDFs 1e-10 relative to the CPU path; gradient within 1e-6 relative of the
central differences (a knot's error measured against max(|cd|, 1e-3 max|grad|): knots
whose difference quotient is rounding noise around zero are held to an absolute floor).
"""
import numpy as np
import pytest

from bromoc_amd import configs
from bromoc_amd.adjoint import RdfLoss, RdfSpec, rdf_loss
from oracle.oracle import Oracle, oracle_table_builder
from oracle.rdf import rdf_loss as cpu_rdf_loss, trajectory

pytestmark = pytest.mark.gpu

SPEC = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)])
H = 1e-6


def _system(seed=0, n=60, box=36.0, curve_seed=None):
    return configs.adjoint_system(seed=seed, n_per_type=n, box=box, curve_seed=curve_seed)


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


def _cd(s, noise, nsteps, target, is_, k, forward):
    vals = []
    for sg in (1.0, -1.0):
        u = s.meta["params"][is_][0].copy()
        u[k] += sg * H
        vals.append(forward(configs.with_knots(s, {is_: u}, oracle_table_builder)))
    return (vals[0] - vals[1]) / (2 * H)


def _check(adj, cd, scale):
    err = abs(adj - cd) / max(abs(cd), 1e-3 * scale)
    return err


def test_forward_loss_matches_cpu():
    s = _system(seed=1, n=40)
    nsteps = 30
    noise = np.random.default_rng(2).standard_normal((nsteps, 3 * s.ntot))
    target = np.random.default_rng(3).random((len(SPEC.groups), SPEC.nbins))
    lg, gg, _ = rdf_loss(s, noise, nsteps, SPEC, target=target, grad=False)
    lc, gc = cpu_rdf_loss(trajectory(Oracle(s), noise, nsteps), s, SPEC, target)
    assert abs(lg - lc) <= 1e-10 * lc
    assert np.max(np.abs(gg - gc)) <= 1e-10 * np.max(gc)


@pytest.mark.parametrize("nsteps,ckpt", [(60, 20), (45, 7)])
def test_gradient_vs_gpu_central_differences(nsteps, ckpt):
    s = _system(seed=4, n=50)
    noise = np.random.default_rng(5).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=77)
    loss, _, grads = rdf_loss(s, noise, nsteps, SPEC, target=target, checkpoint_every=ckpt)
    scale = max(float(np.max(np.abs(g))) for g in grads.values())
    fwd = lambda t: rdf_loss(t, noise, nsteps, SPEC, target=target, grad=False)[0]
    rng = np.random.default_rng(6)
    worst = 0.0
    for is_, k in _pick(grads, rng, 16, r=s.meta["rgrid"]):
        cd = _cd(s, noise, nsteps, target, is_, k, fwd)
        e = _check(grads[is_][k], cd, scale)
        print(f"pair {is_} knot {k}: adjoint {grads[is_][k]:.9e} cd {cd:.9e} err {e:.2e}")
        worst = max(worst, e)
    assert worst <= 1e-6, worst


def test_gradient_vs_cpu_central_differences():
    """Central differences of the loss through the CPU oracle's trajectory."""
    s = _system(seed=7, n=30, box=30.0)
    nsteps = 25
    noise = np.random.default_rng(8).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=78)
    _, _, grads = rdf_loss(s, noise, nsteps, SPEC, target=target)
    scale = max(float(np.max(np.abs(g))) for g in grads.values())
    fwd = lambda t: cpu_rdf_loss(trajectory(Oracle(t), noise, nsteps), t, SPEC, target)[0]
    rng = np.random.default_rng(9)
    worst = 0.0
    for is_, k in _pick(grads, rng, 6, r=s.meta["rgrid"]):
        cd = _cd(s, noise, nsteps, target, is_, k, fwd)
        worst = max(worst, _check(grads[is_][k], cd, scale))
    assert worst <= 1e-6, worst


def test_config_differentiation_64_knots():
    """The differentiation workload: 200 K+ + 200 Cl- + a fixed wall site in a 60 A box,
    4 x 1851 knot parameters, 400 BD steps of fixed noise checkpointed every 20; target
    RDFs from a second seed's potentials; 64 random knots against central differences."""
    s = _system(seed=10, n=200, box=60.0)
    nsteps = 400
    noise = np.random.default_rng(11).standard_normal((nsteps, 3 * s.ntot))
    target = _target(s, noise, nsteps, curve_seed=12)
    loss, _, grads = rdf_loss(s, noise, nsteps, SPEC, target=target, checkpoint_every=20)
    assert sum(len(g) for g in grads.values()) == 4 * 1851
    scale = max(float(np.max(np.abs(g))) for g in grads.values())
    fwd = lambda t: rdf_loss(t, noise, nsteps, SPEC, target=target, grad=False)[0]
    rng = np.random.default_rng(13)
    errs = []
    for is_, k in _pick(grads, rng, 64, r=s.meta["rgrid"]):
        errs.append(_check(grads[is_][k], _cd(s, noise, nsteps, target, is_, k, fwd), scale))
    print(f"loss {loss:.6e}; 64 knots: median err {np.median(errs):.2e}, max {max(errs):.2e}")
    assert max(errs) <= 1e-6, max(errs)


def test_autograd_function():
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
    assert np.allclose(knots.grad.numpy(), want, rtol=1e-12, atol=0.0)

"""Gradients of an RDF loss of a BD trajectory with respect to the EFPOT knot values.

    spec = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)])
    loss, g, grads = rdf_loss(system, noise, nsteps=400, spec=spec, target=g_target)
    # grads: {pair type: dL/du over its knots}

The trajectory and its reverse pass run in the HIP library (bromoc_adjoint_rdf); the chain
rule through the table build is bromoc_table_vjp.  RdfLoss wraps both as a
torch.autograd.Function for fitting.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .engine import Engine
from .system import System, ptr


class RdfStruct(C.Structure):
    _fields_ = [("ngroup", C.c_int32), ("nbins", C.c_int32), ("groups", C.POINTER(C.c_int32)),
                ("r0", C.c_double), ("dr", C.c_double), ("sigma", C.c_double),
                ("cut", C.c_double), ("target", C.POINTER(C.c_double)),
                ("wflux", C.c_double), ("jtarget", C.c_double)]


@dataclass
class RdfSpec:
    """Smoothed RDFs of atnam ion-type pairs: bins of dr from r0 (nbins of them), Gaussian
    smoothing sigma, terms beyond cut dropped; with wflux > 0 the loss also holds
    wflux (J - jtarget)^2 for the charge current J through the box (e A / ps)."""
    groups: list = field(default_factory=lambda: [(1, 1), (1, 2), (2, 2)])
    r0: float = 2.0
    dr: float = 0.1
    nbins: int = 130
    sigma: float = 0.15
    cut: float = 0.75
    wflux: float = 0.0
    jtarget: float = 0.0

    def centres(self):
        return self.r0 + (np.arange(self.nbins) + 0.5) * self.dr


def _declare(lib):
    P = C.POINTER
    lib.bromoc_adjoint_rdf.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(RdfStruct),
                                       P(C.c_double), P(C.c_double), P(C.c_double),
                                       P(C.c_double)]
    lib.bromoc_adjoint_rdf.restype = C.c_int
    lib.bromoc_table_vjp.argtypes = [C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double),
                                     P(C.c_double)]
    lib.bromoc_table_vjp.restype = C.c_int
    lib.bromoc_tangent_rdf.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(RdfStruct),
                                       P(C.c_double), P(C.c_double), P(C.c_double),
                                       P(C.c_double)]
    lib.bromoc_tangent_rdf.restype = C.c_int


def _rdf_struct(spec: RdfSpec, target):
    groups = np.ascontiguousarray(np.array(spec.groups, np.int32).reshape(-1))
    tgt = None if target is None else np.ascontiguousarray(np.asarray(target, float).reshape(-1))
    rs = RdfStruct(len(spec.groups), spec.nbins, ptr(groups, C.c_int32), spec.r0, spec.dr,
                   spec.sigma, spec.cut, ptr(tgt) if tgt is not None else None, spec.wflux,
                   spec.jtarget)
    return rs, (groups, tgt)


def rdf_loss(system: System, noise, nsteps, spec: RdfSpec, target=None, grad=True,
             checkpoint_every=20, engine: Engine | None = None):
    """Loss, RDFs [ngroup, nbins] and (if grad) {pair type: dL/du} for every parameter pair
    type in system.meta['params']."""
    eng = engine or Engine(system)
    _declare(eng.lib)
    eng.set_positions(np.stack([system.x, system.y, system.z], 1)[:system.ntot],
                      system.typei[:system.ntot])
    eng.upload_noise(noise)
    ng, nb = len(spec.groups), spec.nbins
    rs, keep = _rdf_struct(spec, target)
    loss = C.c_double()
    g = np.zeros(ng * nb)
    nknots = len(system.ep) // 3
    gep = np.zeros(3 * max(nknots, 1)) if grad else None
    gsc = np.zeros(3 * (system.npair + 1)) if grad else None
    rc = eng.lib.bromoc_adjoint_rdf(eng.ctx, nsteps, checkpoint_every, C.byref(rs),
                                    C.byref(loss), ptr(g), ptr(gep) if grad else None,
                                    ptr(gsc) if grad else None)
    eng._chk("bromoc_adjoint_rdf", rc)
    grads = knot_gradients(system, gep, gsc, eng.lib) if grad else None
    return loss.value, g.reshape(ng, nb), grads


def tangent(system: System, noise, nsteps, spec: RdfSpec, v_ep, v_sc, target=None,
            checkpoint_every=20, engine: Engine | None = None):
    """(loss, <dL/d(ep, sc), v>): the loss's derivative along the table-coefficient
    direction (v_ep [3*nknots], v_sc [3*(npair+1)]) by the tangent-linear pass."""
    eng = engine or Engine(system)
    _declare(eng.lib)
    eng.set_positions(np.stack([system.x, system.y, system.z], 1)[:system.ntot],
                      system.typei[:system.ntot])
    eng.upload_noise(noise)
    rs, keep = _rdf_struct(spec, target)
    ve = np.ascontiguousarray(v_ep, float)
    vs = np.ascontiguousarray(v_sc, float)
    loss, dl = C.c_double(), C.c_double()
    rc = eng.lib.bromoc_tangent_rdf(eng.ctx, nsteps, checkpoint_every, C.byref(rs), ptr(ve),
                                    ptr(vs), C.byref(loss), C.byref(dl))
    eng._chk("bromoc_tangent_rdf", rc)
    return loss.value, dl.value


def table_gradients(system: System, noise, nsteps, spec: RdfSpec, target=None,
                    checkpoint_every=20, engine: Engine | None = None):
    """(loss, gep [3*nknots], gsc [3*(npair+1)]): the reverse pass's gradient with respect to
    the table coefficients themselves, before the chain rule to the knot values."""
    eng = engine or Engine(system)
    _declare(eng.lib)
    eng.set_positions(np.stack([system.x, system.y, system.z], 1)[:system.ntot],
                      system.typei[:system.ntot])
    eng.upload_noise(noise)
    rs, keep = _rdf_struct(spec, target)
    loss = C.c_double()
    nknots = len(system.ep) // 3
    gep = np.zeros(3 * max(nknots, 1))
    gsc = np.zeros(3 * (system.npair + 1))
    rc = eng.lib.bromoc_adjoint_rdf(eng.ctx, nsteps, checkpoint_every, C.byref(rs),
                                    C.byref(loss), None, ptr(gep), ptr(gsc))
    eng._chk("bromoc_adjoint_rdf", rc)
    return loss.value, gep, gsc


def knot_gradients(system: System, gep, gsc, lib=None, ctx=None):
    """{pair type: dL/du} through the table build of each parameter pair type: on the
    device of context ctx when one is given (one thread per knot), else on the host;
    both give the same bits."""
    lib = lib or abi.load_library()
    _declare(lib)
    r = np.ascontiguousarray(system.meta["rgrid"], float)
    n = len(r)
    out = {}
    for is_ in system.meta["params"]:
        off = system.pairs[is_]["nxi"]
        ge = np.ascontiguousarray(gep[3 * off:3 * (off + n)])
        gs = np.ascontiguousarray(gsc[3 * is_:3 * is_ + 3])
        gu = np.zeros(n)
        if ctx is not None:
            rc = lib.bromoc_table_vjp_device(ctx, n, ptr(r), ptr(ge), ptr(gs), ptr(gu))
        else:
            rc = lib.bromoc_table_vjp(n, ptr(r), ptr(ge), ptr(gs), ptr(gu))
        if rc:
            raise abi.BromocError("bromoc_table_vjp_device" if ctx is not None
                                  else "bromoc_table_vjp", rc)
        out[is_] = gu
    return out


try:
    import torch

    class RdfLoss(torch.autograd.Function):
        """loss = RdfLoss.apply(knots, system, noise, nsteps, spec, target): knots is a
        [n_param_pairs, n_knots] float64 tensor in the order of system.meta['params']."""

        @staticmethod
        def forward(ctx, knots, system, noise, nsteps, spec, target):
            from .configs import with_knots
            keys = list(system.meta["params"])
            k = knots.detach().cpu().numpy()
            s = with_knots(system, {is_: k[q] for q, is_ in enumerate(keys)})
            loss, _, grads = rdf_loss(s, noise, nsteps, spec, target)
            ctx.save_for_backward(torch.as_tensor(np.stack([grads[is_] for is_ in keys])))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None, None, None, None
except ImportError:     # torch is only needed for fitting
    RdfLoss = None

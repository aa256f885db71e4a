"""Python front-end of libbromoc_gpu.so, shaped like the reference's routines.

    eng = Engine(system)            # the set-up shell_simul.f90 does before SIMUL
    e   = eng.energy()              # SUBROUTINE ENERGY          (energy.mono.f90:19)
    eng.step(nbd, ...)              # simul1.f90:363-390 BD loop (ENERGY + DYNAMICS0/0NUC)
    d   = eng.interact(xyz, t, j)   # SUBROUTINE INTERACT        (interact.f90:19)
    eng.grand(ngcmc, ...)           # SUBROUTINE GRAND           (grand.f90:19)

Every call goes to the HIP library; a missing library raises ImportError and there is no
CPU fallback.  Errors come back as BromocError (the reference instead stops through
error(...,faterr), error.f90:30-32).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .system import Gcmc, System, ptr


class Engine:
    def __init__(self, system: System, device: int = 0):
        self.lib = abi.load_library()
        self.system = system
        st, keep = system.to_struct()
        ctx = C.c_void_p()
        rc = self.lib.bromoc_create(C.byref(st), device, C.byref(ctx))
        if rc:
            raise abi.BromocError("bromoc_create", rc)
        self.ctx = ctx
        self.ntot = system.ntot
        self._noise_rows = 0

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.bromoc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, name, rc):
        if rc:
            msg = self.lib.bromoc_last_error(self.ctx)
            raise abi.BromocError(name, rc, msg.decode() if msg else "")

    # --------------------------------------------------------------- state
    def positions(self):
        n = self.ntot
        x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
        t = np.zeros(n, np.int32)
        nn = C.c_int32()
        self._chk("bromoc_get_positions",
                  self.lib.bromoc_get_positions(self.ctx, ptr(x), ptr(y), ptr(z),
                                                ptr(t, C.c_int32), C.byref(nn)))
        return np.stack([x, y, z], 1), t

    def set_positions(self, pos, typei=None):
        pos = np.ascontiguousarray(pos, float)
        n = len(pos)
        x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
        tp = None
        if typei is not None:
            ta = np.ascontiguousarray(typei, np.int32)
            tp = ptr(ta, C.c_int32)
        self._chk("bromoc_set_positions",
                  self.lib.bromoc_set_positions(self.ctx, n, ptr(x), ptr(y), ptr(z), tp))
        self.ntot = n

    def forces(self):
        n = self.ntot
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        self._chk("bromoc_get_forces", self.lib.bromoc_get_forces(self.ctx, ptr(fx), ptr(fy), ptr(fz)))
        return np.stack([fx, fy, fz], 1)

    # ------------------------------------------------------------ routines
    def energy(self, forces=True):
        e = abi.Energies()
        self._chk("bromoc_energy", self.lib.bromoc_energy(self.ctx, 1 if forces else 0, C.byref(e)))
        return e.as_dict()

    def upload_noise(self, noise):
        noise = np.ascontiguousarray(noise, float)
        self._chk("bromoc_upload_noise",
                  self.lib.bromoc_upload_noise(self.ctx, noise.shape[0], ptr(noise)))
        self._noise_rows = noise.shape[0]

    def step(self, nsteps, mode=abi.NOISE_INJECTED, first_row=0, seed=0, counter0=0,
             energies=True):
        e = abi.Energies()
        self._chk("bromoc_step",
                  self.lib.bromoc_step(self.ctx, nsteps, mode, first_row, seed, counter0,
                                       C.byref(e) if energies else None))
        return e.as_dict() if energies else None

    def time_steps(self, nsteps, mode=abi.NOISE_PHILOX, first_row=0, seed=0, counter0=0):
        """Run nsteps BD steps with HIP events around every launch: (total_ms, kernel_ms)."""
        tot, ker = C.c_double(), C.c_double()
        self._chk("bromoc_time_steps",
                  self.lib.bromoc_time_steps(self.ctx, nsteps, mode, first_row, seed, counter0,
                                             C.byref(tot), C.byref(ker)))
        return tot.value, ker.value

    def spatial_order(self):
        """(keys, order, rank, cell_start): the spatial order the step kernels use."""
        n = self.ntot
        keys = np.zeros(n, np.uint32)
        order = np.zeros(n, np.int32)
        rank = np.zeros(n, np.int32)
        cs = np.zeros(n + 1, np.int32)
        nc = C.c_int32()
        self._chk("bromoc_spatial_order",
                  self.lib.bromoc_spatial_order(self.ctx, ptr(keys, C.c_uint32), ptr(order, C.c_int32),
                                                ptr(rank, C.c_int32), ptr(cs, C.c_int32),
                                                C.byref(nc)))
        return keys, order, rank, cs[:nc.value + 1]

    def tuning(self):
        """Targets per block the autotuner picked for this system."""
        t = C.c_int32()
        self._chk("bromoc_get_tuning", self.lib.bromoc_get_tuning(self.ctx, C.byref(t)))
        return t.value

    def dynamics(self, mode=abi.NOISE_INJECTED, row=0, seed=0, counter=0):
        self._chk("bromoc_dynamics", self.lib.bromoc_dynamics(self.ctx, mode, row, seed, counter))

    def flux(self):
        s = self.system
        nt, np_ = s.ntype - s.nold, len(s.zcont)
        nf = np.zeros(max(nt * np_, 1), np.int32)
        nb = np.zeros(max(nt * np_, 1), np.int32)
        self._chk("bromoc_get_flux",
                  self.lib.bromoc_get_flux(self.ctx, ptr(nf, C.c_int32), ptr(nb, C.c_int32)))
        return nf[:nt * np_].reshape(nt, np_), nb[:nt * np_].reshape(nt, np_)

    def warn(self):
        w = np.zeros(self.system.nttyp + 1, np.int32)
        self._chk("bromoc_get_warn", self.lib.bromoc_get_warn(self.ctx, ptr(w, C.c_int32)))
        return w

    def interact(self, xyz, jtype, j, alert=False):
        d = C.c_double()
        self._chk("bromoc_interact",
                  self.lib.bromoc_interact(self.ctx, float(xyz[0]), float(xyz[1]), float(xyz[2]),
                                           int(jtype), int(j), 1 if alert else 0, C.byref(d)))
        return d.value

    def set_gcmc(self, g: Gcmc, ibuffer):
        st, self._gkeep = g.to_struct()
        self._chk("bromoc_set_gcmc", self.lib.bromoc_set_gcmc(self.ctx, C.byref(st)))
        ib = np.ascontiguousarray(ibuffer[:self.ntot], np.int32)
        self._chk("bromoc_set_ibuffer", self.lib.bromoc_set_ibuffer(self.ctx, ptr(ib, C.c_int32)))

    def ibuffer(self):
        ib = np.zeros(self.ntot, np.int32)
        self._chk("bromoc_get_ibuffer", self.lib.bromoc_get_ibuffer(self.ctx, ptr(ib, C.c_int32)))
        return ib

    def grand(self, ngcmc, uniforms, nat, prob, icycle, ener):
        u = np.ascontiguousarray(uniforms, float)
        used = C.c_int64()
        e = C.c_double(ener)
        self._chk("bromoc_grand",
                  self.lib.bromoc_grand(self.ctx, ngcmc, ptr(u), len(u), C.byref(used),
                                        ptr(nat, C.c_int32), ptr(prob, C.c_int32), icycle,
                                        C.byref(e)))
        nn = C.c_int32()
        self.lib.bromoc_get_positions(self.ctx, None, None, None, None, C.byref(nn))
        self.ntot = nn.value
        return used.value, e.value

    def metropolis(self, nmcm, mcmax, ikbt, uniforms, ener=0.0):
        """METROPOLIS(nmcm) (metropolis.f90:19-58): returns (uniforms used, trials done,
        ener)."""
        u = np.ascontiguousarray(uniforms, float)
        used, done, e = C.c_int64(), C.c_int32(), C.c_double(ener)
        self._chk("bromoc_metropolis",
                  self.lib.bromoc_metropolis(self.ctx, nmcm, mcmax, ikbt, ptr(u), len(u),
                                             C.byref(used), C.byref(done), C.byref(e)))
        return used.value, done.value, e.value


    # ------------------------------------------------------------- tables
    def update_tables(self, tables):
        """Rebuild tables on the device (bromoc_update_tables): {pair type: (r, u)}."""
        keys = sorted(tables)
        is_ = np.array(keys, np.int32)
        n = np.array([len(tables[k][0]) for k in keys], np.int32)
        r = np.ascontiguousarray(np.concatenate([tables[k][0] for k in keys]), float)
        u = np.ascontiguousarray(np.concatenate([tables[k][1] for k in keys]), float)
        self._chk("bromoc_update_tables",
                  self.lib.bromoc_update_tables(self.ctx, len(keys), ptr(is_, C.c_int32),
                                                ptr(n, C.c_int32), ptr(r), ptr(u)))

    def table_vjp(self, r, grad_ep, grad_sc):
        """bromoc_table_vjp_device on this context's GPU."""
        r = np.ascontiguousarray(r, float)
        out = np.zeros(len(r))
        self._chk("bromoc_table_vjp_device",
                  self.lib.bromoc_table_vjp_device(self.ctx, len(r), ptr(r),
                                                   ptr(np.ascontiguousarray(grad_ep, float)),
                                                   ptr(np.ascontiguousarray(grad_sc, float)),
                                                   ptr(out)))
        return out

    # ------------------------------------------------------- per-cycle statistics
    def set_analysis(self, **kw):
        """The statistics of simul1.f90's cycle loop (analysis_struct builds the settings);
        zeroes the accumulators."""
        self._an = analysis_struct(self.system, **kw)
        self._chk("bromoc_set_analysis", self.lib.bromoc_set_analysis(self.ctx, C.byref(self._an)))

    def analyze(self):
        """One cycle's rho, rDNA, g(r) and prob2/prob3 (simul1.f90:477-538)."""
        self._chk("bromoc_analyze", self.lib.bromoc_analyze(self.ctx))

    def get_analysis(self):
        """dict of the accumulators: rho [ntype, nzmax], rdna [nttyp, nzmax], gr [ntype,
        npoints], irmax, prob2 / prob3 [ntype, datom+1] (row t-1: type t), pres (the
        reference's sum over the steps) and pres_steps; statistics not selected are None."""
        a, s = self._an, self.system
        out = _analysis_arrays(a, s)
        irm, pres, nps = C.c_int32(), C.c_double(), C.c_int64()
        i64 = lambda k: ptr(out[k], C.c_int64) if out[k] is not None else None
        self._chk("bromoc_get_analysis",
                  self.lib.bromoc_get_analysis(self.ctx, i64("rho"), i64("rdna"), i64("gr"),
                                               C.byref(irm), i64("prob2"), i64("prob3"),
                                               C.byref(pres), C.byref(nps)))
        out.update(irmax=irm.value, pres=pres.value, pres_steps=nps.value)
        return out


def discretize(pair, nnp, mnp, res, rth=0.0, srpx=0.0, srpk=0.0, srpy=0.0):
    """DISCRETIZE (discretize.f90:19-120) by the library: the table of an abi.Pair's analytic
    parameters sampled at k*res, k = nnp..mnp.  Returns (ep [3n], (sc1, sc2, sc3))."""
    lib = abi.load_library()
    ep = np.zeros(3 * (mnp - nnp + 1))
    out = abi.Pair()
    rc = lib.bromoc_discretize(C.byref(pair), nnp, mnp, res, rth, srpx, srpk, srpy, ptr(ep),
                               C.byref(out))
    if rc:
        raise abi.BromocError("bromoc_discretize", rc)
    return ep, (out.sc1, out.sc2, out.sc3)


def table_vjp(r, grad_ep, grad_sc):
    """bromoc_table_vjp on the host: d(table)/d(samples) applied to a table gradient."""
    lib = abi.load_library()
    r = np.ascontiguousarray(r, float)
    out = np.zeros(len(r))
    rc = lib.bromoc_table_vjp(len(r), ptr(r), ptr(np.ascontiguousarray(grad_ep, float)),
                              ptr(np.ascontiguousarray(grad_sc, float)), ptr(out))
    if rc:
        raise abi.BromocError("bromoc_table_vjp", rc)
    return out


def analysis_struct(system, what, nzmax=None, zmini=None, idelz=2.0, igr=0, npoints=10000,
                    dr=0.1, datom=23000, czmin=0.0, czmax=0.0, tvol=None, temp=None,
                    ikbt=None):
    """bromoc_analysis with the reference's defaults (simul1.f90:179-230): z bins from lz2m
    at idelz = 2 per Angstrom, nzmax = int(lz * idelz) capped at 10000, dr = 0.1, the box
    volume, the temperature in system.meta (300 K when absent)."""
    lo, hi = system.box[4], system.box[5]
    if zmini is None:
        zmini = lo
    if nzmax is None:
        nzmax = min(int((hi - lo) * idelz), 10000)
    if tvol is None:
        b = system.box
        tvol = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])
    if temp is None:
        temp = system.meta.get("temp", 300.0)
    if ikbt is None:
        from .system import kbt_of
        ikbt = 1.0 / kbt_of(temp)
    return abi.Analysis(what=what, nzmax=nzmax, zmini=zmini, idelz=idelz, igr=igr,
                        npoints=npoints, dr=dr, datom=datom, reserved=0, czmin=czmin,
                        czmax=czmax, tvol=tvol, temp=temp, ikbt=ikbt)


def _analysis_arrays(a, s):
    nt, n2 = s.ntype, s.nttyp
    pick = lambda bit, shape: np.zeros(shape, np.int64) if a.what & bit else None
    return dict(rho=pick(abi.AN_RHO, (nt, a.nzmax)), rdna=pick(abi.AN_RDNA, (n2, a.nzmax)),
                gr=pick(abi.AN_GR, (nt, a.npoints)),
                prob2=pick(abi.AN_PROB, (nt, a.datom + 1)),
                prob3=pick(abi.AN_PROB, (nt, a.datom + 1)))


def philox4x64(ctr, key):
    """Raw Philox4x64-10 blocks computed on the device."""
    lib = abi.load_library()
    ctr = np.ascontiguousarray(ctr, np.uint64).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint64).reshape(-1, 2)
    out = np.zeros_like(ctr)
    rc = lib.bromoc_philox4x64(len(ctr), ptr(ctr, C.c_uint64), ptr(key, C.c_uint64),
                               ptr(out, C.c_uint64))
    if rc:
        raise abi.BromocError("bromoc_philox4x64", rc)
    return out


class Batch:
    """Many ion systems sharing one System's tables (replica batches of the HIP library).

        b = Batch(shared, [Replica(...), ...])
        b.run(nsteps, seed, counter0, gcmc_every, ngcmc)   # GRAND + BD cycles, all replicas
        x, typei, ntot, ibuffer, nat, ener = b.get(r)
    """

    def __init__(self, shared: System, replicas, device: int = 0):
        from .system import Replica  # noqa: F401
        self.lib = abi.load_library()
        self.shared = shared
        self.replicas = list(replicas)
        st, keep = shared.to_struct()
        arr = (abi.ReplicaStruct * len(self.replicas))()
        for k, rep in enumerate(self.replicas):
            arr[k] = rep.to_struct(keep)
        self._keep = keep
        ctx = C.c_void_p()
        rc = self.lib.bromoc_batch_create(C.byref(st), len(self.replicas), arr, device,
                                          C.byref(ctx))
        if rc:
            raise abi.BromocError("bromoc_batch_create", rc)
        self.ctx = ctx
        self.nrep = len(self.replicas)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.bromoc_batch_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, name, rc):
        if rc:
            msg = self.lib.bromoc_batch_last_error(self.ctx)
            raise abi.BromocError(name, rc, msg.decode() if msg else "")

    def run(self, nsteps, seed=0, counter0=0, gcmc_every=0, ngcmc=0, cycle0=0, energies=False):
        e = (abi.Energies * self.nrep)() if energies else None
        self._chk("bromoc_batch_run",
                  self.lib.bromoc_batch_run(self.ctx, nsteps, seed, counter0, gcmc_every, ngcmc,
                                            cycle0, e))
        return [x.as_dict() for x in e] if energies else None

    def energy(self):
        e = (abi.Energies * self.nrep)()
        self._chk("bromoc_batch_energy", self.lib.bromoc_batch_energy(self.ctx, e))
        return [x.as_dict() for x in e]

    def get(self, r):
        rep = self.replicas[r]
        cap = max(rep.capacity, rep.ntot, 1)
        x, y, z = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        t = np.zeros(cap, np.int32)
        ib = np.zeros(cap, np.int32)
        nb = rep.gcmc.nbuffer if rep.gcmc is not None else 0
        nat = np.zeros(nb + 1, np.int32)
        n = C.c_int32()
        en = C.c_double()
        self._chk("bromoc_batch_get",
                  self.lib.bromoc_batch_get(self.ctx, r, ptr(x), ptr(y), ptr(z), ptr(t, C.c_int32),
                                            C.byref(n), ptr(ib, C.c_int32), ptr(nat, C.c_int32),
                                            C.byref(en)))
        k = n.value
        return (np.stack([x[:k], y[:k], z[:k]], axis=1), t[:k].copy(), k, ib[:k].copy(), nat,
                en.value)

    def forces(self, r):
        k = self.get(r)[2]
        fx, fy, fz = np.zeros(max(k, 1)), np.zeros(max(k, 1)), np.zeros(max(k, 1))
        self._chk("bromoc_batch_get_forces",
                  self.lib.bromoc_batch_get_forces(self.ctx, r, ptr(fx), ptr(fy), ptr(fz)))
        return np.stack([fx[:k], fy[:k], fz[:k]], axis=1)

    def flux(self, r):
        s = self.shared
        nt, np_ = s.ntype - s.nold, len(s.zcont)
        nf = np.zeros(max(nt * np_, 1), np.int32)
        nb = np.zeros(max(nt * np_, 1), np.int32)
        self._chk("bromoc_batch_get_flux",
                  self.lib.bromoc_batch_get_flux(self.ctx, r, ptr(nf, C.c_int32), ptr(nb, C.c_int32)))
        return nf[:nt * np_].reshape(nt, np_), nb[:nt * np_].reshape(nt, np_)

    def prob(self, r):
        rep = self.replicas[r]
        nb = rep.gcmc.nbuffer if rep.gcmc is not None else 0
        cap = max(rep.capacity, rep.ntot, 1)
        p = np.zeros((max(nb, 1), cap + 1), np.int32)
        self._chk("bromoc_batch_get_prob",
                  self.lib.bromoc_batch_get_prob(self.ctx, r, ptr(p, C.c_int32)))
        return p[:nb]

    def time_steps(self, nsteps, seed=0, counter0=0):
        tot, ker = C.c_double(), C.c_double()
        self._chk("bromoc_batch_time_steps",
                  self.lib.bromoc_batch_time_steps(self.ctx, nsteps, seed, counter0,
                                                   C.byref(tot), C.byref(ker)))
        return tot.value, ker.value

    def tuning(self):
        """(workgroups of the persistent step kernel or 0, lanes per target)"""
        v, l = C.c_int32(), C.c_int32()
        self._chk("bromoc_batch_get_tuning",
                  self.lib.bromoc_batch_get_tuning(self.ctx, C.byref(v), C.byref(l)))
        return v.value, l.value

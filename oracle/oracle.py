"""ctypes front-ends of the two CPU checkers (TEST INFRASTRUCTURE ONLY).

Oracle     -- oracle/liboracle.so, the C restatement (bromoc_oracle.c)
RefFortran -- oracle/_ref/libbromoc_ref.so, the reference Fortran routines themselves
              (built only where the reference sources exist); with lib=SHIM_SO the same
              harness drives oracle/_ref/libbromoc_shimcheck.so, where the reference's
              ENERGY/DYNAMICS0/DYNAMICS0NUC/INTERACT/GRAND are replaced by the Fortran
              front-end bromoc_amd/fortran/bromoc_gpu.f90 over the C ABI, served by the
              C restatement (shim_adapter.c); with lib=FRONT_GPU_SO the front-end is served
              by the HIP library itself (a GPU is needed)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from bromoc_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbromoc_ref.so")
SHIM_SO = os.path.join(HERE, "_ref", "libbromoc_shimcheck.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libbromoc_ref_omp.so")   # OpenMP ENERGY
FRONT_GPU_SO = os.path.join(HERE, "_ref", "libbromoc_linkcheck.so")   # front-end over HIP

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def ref_available():
    return os.path.exists(REF_SO)


_olib = None


def oracle_lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: make -C oracle")
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        P = C.POINTER
        for name, args, res in [
            ("orc_create", [P(abi.SystemStruct)], vp),
            ("orc_destroy", [vp], None),
            ("orc_spline", [C.c_int32, _dp, _dp, _dp, _dp, _dp], C.c_int),
            ("orc_set_stream", [vp, C.c_uint64], None),
            ("orc_set_positions", [vp, C.c_int32, _dp, _dp, _dp, _ip], C.c_int),
            ("orc_get_positions", [vp, _dp, _dp, _dp, _ip, _ip], C.c_int),
            ("orc_get_forces", [vp, _dp, _dp, _dp], C.c_int),
            ("orc_get_flux", [vp, _ip, _ip], C.c_int),
            ("orc_get_warn", [vp, _ip], C.c_int),
            ("orc_energy", [vp, C.c_int32, P(abi.Energies)], C.c_int),
            ("orc_upload_noise", [vp, C.c_int32, _dp], C.c_int),
            ("orc_dynamics", [vp, C.c_int32, C.c_int32, C.c_uint64, C.c_uint64], C.c_int),
            ("orc_step", [vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_uint64,
                          P(abi.Energies)], C.c_int),
            ("orc_interact", [vp, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                              C.c_int32, _dp], C.c_int),
            ("orc_set_gcmc", [vp, P(abi.GcmcStruct), _ip], C.c_int),
            ("orc_get_ibuffer", [vp, _ip], C.c_int),
            ("orc_grand", [vp, C.c_int32, _dp, C.c_int64, P(C.c_int64), _ip, _ip, C.c_int64,
                           _dp], C.c_int),
            ("orc_build_table", [C.c_int32, _dp, _dp, C.c_int32, C.c_double, _dp,
                                 P(abi.Pair)], C.c_int),
            ("orc_philox4x64", [P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)], None),
            ("orc_get_dids", [vp, _dp], C.c_int),
            ("orc_repwalls", [C.c_int32, C.c_int32, C.c_int32, P(C.c_int8), C.c_int32], C.c_int),
            ("orc_metropolis", [vp, C.c_int32, C.c_double, C.c_double, _dp, C.c_int64,
                                P(C.c_int64), _ip, _dp], C.c_int),
            ("orc_philox_gauss3", [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, _dp], None),
            ("orc_discretize", [P(abi.Pair), C.c_int32, C.c_int32, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, _dp, P(abi.Pair)], C.c_int),
            ("orc_set_analysis", [vp, P(abi.Analysis)], C.c_int),
            ("orc_analyze", [vp], C.c_int),
            ("orc_get_analysis", [vp, _dp, _dp, _dp, _ip, P(C.c_int64), P(C.c_int64), _dp,
                                  P(C.c_int64)], C.c_int),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _olib = L
    return _olib


def oracle_table_builder(n, r, u, chr_, fct):
    ep = np.zeros(3 * n)
    p = abi.Pair()
    rc = oracle_lib().orc_build_table(n, _p(r), _p(u), chr_, fct, _p(ep), C.byref(p))
    if rc:
        raise RuntimeError(f"orc_build_table {rc}")
    return ep, p


class _Base:
    def forces(self):
        n = self.ntot
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        self._get_forces(_p(fx), _p(fy), _p(fz))
        return np.stack([fx, fy, fz], 1)

    def positions(self):
        n = self.ntot
        x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
        t = np.zeros(n, np.int32)
        nn = C.c_int32()
        self._get_positions(_p(x), _p(y), _p(z), _p(t, C.c_int32), C.byref(nn))
        return np.stack([x, y, z], 1), t


class Oracle(_Base):
    """The C restatement."""

    def __init__(self, system, stream=0):
        self.L = oracle_lib()
        st, self._keep = system.to_struct()
        self.ctx = self.L.orc_create(C.byref(st))
        if not self.ctx:
            raise RuntimeError("orc_create failed")
        self.L.orc_set_stream(self.ctx, stream)
        self.system = system
        self.ntot = system.ntot

    def __del__(self):
        if getattr(self, "ctx", None):
            self.L.orc_destroy(self.ctx)
            self.ctx = None

    def _get_forces(self, *a):
        return self.L.orc_get_forces(self.ctx, *a)

    def _get_positions(self, *a):
        return self.L.orc_get_positions(self.ctx, *a)

    def set_positions(self, pos, typei=None):
        pos = np.ascontiguousarray(pos, float)
        n = len(pos)
        t = np.ascontiguousarray(typei if typei is not None else self.positions()[1][:n], np.int32)
        x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
        rc = self.L.orc_set_positions(self.ctx, n, _p(x), _p(y), _p(z), _p(t, C.c_int32))
        assert rc == 0
        self.ntot = n

    def energy(self, forces=True):
        e = abi.Energies()
        rc = self.L.orc_energy(self.ctx, 1 if forces else 0, C.byref(e))
        if rc:
            raise abi.BromocError("orc_energy", rc)
        return e.as_dict()

    def upload_noise(self, noise):
        noise = np.ascontiguousarray(noise, float)
        rc = self.L.orc_upload_noise(self.ctx, noise.shape[0], _p(noise))
        assert rc == 0

    def dynamics(self, mode=abi.NOISE_INJECTED, row=0, seed=0, counter=0):
        rc = self.L.orc_dynamics(self.ctx, mode, row, seed, counter)
        if rc:
            raise abi.BromocError("orc_dynamics", rc)

    def step(self, nsteps, mode=abi.NOISE_INJECTED, first_row=0, seed=0, counter0=0):
        e = abi.Energies()
        rc = self.L.orc_step(self.ctx, nsteps, mode, first_row, seed, counter0, C.byref(e))
        if rc:
            raise abi.BromocError("orc_step", rc)
        return e.as_dict()

    def flux(self):
        s = self.system
        nt, np_ = s.ntype - s.nold, len(s.zcont)
        nf = np.zeros(max(nt * np_, 1), np.int32)
        nb = np.zeros(max(nt * np_, 1), np.int32)
        self.L.orc_get_flux(self.ctx, _p(nf, C.c_int32), _p(nb, C.c_int32))
        return nf[:nt * np_].reshape(nt, np_), nb[:nt * np_].reshape(nt, np_)

    def warn(self):
        w = np.zeros(self.system.nttyp + 1, np.int32)
        self.L.orc_get_warn(self.ctx, _p(w, C.c_int32))
        return w

    def interact(self, xyz, jtype, j, alert=False):
        d = C.c_double()
        rc = self.L.orc_interact(self.ctx, float(xyz[0]), float(xyz[1]), float(xyz[2]),
                                 int(jtype), int(j), 1 if alert else 0, C.byref(d))
        if rc:
            raise abi.BromocError("orc_interact", rc)
        return d.value

    def set_gcmc(self, g, ibuffer):
        st, self._gkeep = g.to_struct()
        ib = np.ascontiguousarray(ibuffer, np.int32)
        self.L.orc_set_gcmc(self.ctx, C.byref(st), _p(ib, C.c_int32))

    def ibuffer(self):
        ib = np.zeros(self.ntot, np.int32)
        self.L.orc_get_ibuffer(self.ctx, _p(ib, C.c_int32))
        return ib

    def grand(self, ngcmc, uniforms, nat, prob, icycle, ener):
        u = np.ascontiguousarray(uniforms, float)
        used = C.c_int64()
        e = C.c_double(ener)
        rc = self.L.orc_grand(self.ctx, ngcmc, _p(u), len(u), C.byref(used),
                              _p(nat, C.c_int32), _p(prob, C.c_int32), icycle, C.byref(e))
        if rc:
            raise abi.BromocError("orc_grand", rc)
        nn = C.c_int32()
        self.L.orc_get_positions(self.ctx, None, None, None, None, C.byref(nn))
        self.ntot = nn.value
        return used.value, e.value

    def dids(self):
        d = np.zeros(5 * self.system.capacity)
        self.L.orc_get_dids(self.ctx, _p(d))
        return d[:5 * self.ntot].reshape(-1, 5)

    def metropolis(self, nmcm, mcmax, ikbt, uniforms, ener=0.0):
        """METROPOLIS(nmcm); returns (uniforms used, trials done, ener)."""
        u = np.ascontiguousarray(uniforms, float)
        used, done, e = C.c_int64(), C.c_int32(), C.c_double(ener)
        rc = self.L.orc_metropolis(self.ctx, nmcm, mcmax, ikbt, _p(u), len(u), C.byref(used),
                                   C.byref(done), C.byref(e))
        if rc:
            raise abi.BromocError("orc_metropolis", rc)
        return used.value, done.value, e.value


    def set_analysis(self, **kw):
        from bromoc_amd.engine import analysis_struct
        self._an = analysis_struct(self.system, **kw)
        assert self.L.orc_set_analysis(self.ctx, C.byref(self._an)) == 0

    def analyze(self):
        assert self.L.orc_analyze(self.ctx) == 0

    def get_analysis(self):
        """As Engine.get_analysis; rho, rdna and gr come back as the reference's reals
        (float64 arrays of whole numbers)."""
        from bromoc_amd.engine import _analysis_arrays
        out = _analysis_arrays(self._an, self.system)
        for k in ("rho", "rdna", "gr"):
            if out[k] is not None:
                out[k] = out[k].astype(float)
        irm, pres, nps = C.c_int32(), C.c_double(), C.c_int64()
        f64 = lambda k: _p(out[k]) if out[k] is not None else None
        i64 = lambda k: _p(out[k], C.c_int64) if out[k] is not None else None
        self.L.orc_get_analysis(self.ctx, f64("rho"), f64("rdna"), f64("gr"), C.byref(irm),
                                i64("prob2"), i64("prob3"), C.byref(pres), C.byref(nps))
        out.update(irmax=irm.value, pres=pres.value, pres_steps=nps.value)
        return out


_rlibs = {}


def ref_lib(path=REF_SO):
    if path not in _rlibs:
        if not os.path.exists(path):
            raise ImportError(f"{path} missing (needs the reference sources; make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER
        for name, args, res in [
            ("ref_set_nindex", [C.c_int32, C.c_int32], None),
            ("ref_init", [P(abi.SystemStruct)], None),
            ("ref_energy", [C.c_int32, _dp], None),
            ("ref_get_forces", [_dp, _dp, _dp], None),
            ("ref_get_positions", [_dp, _dp, _dp, _ip, _ip], None),
            ("ref_set_positions", [C.c_int32, _dp, _dp, _dp, _ip], None),
            ("ref_get_flux", [_ip, _ip], None),
            ("ref_get_warn", [_ip], None),
            ("ref_dynamics", [_dp], None),
            ("ref_interact", [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                              C.c_int32, _dp], None),
            ("ref_set_gcmc", [P(abi.GcmcStruct), _ip], None),
            ("ref_get_ibuffer", [_ip], None),
            ("ref_grand", [C.c_int32, _dp, C.c_int64, P(C.c_int64), _ip, _ip, C.c_int32,
                           C.c_int64, _dp], None),
            ("ref_discretize", [P(abi.Pair), C.c_int32, C.c_int32, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, _dp, P(abi.Pair)], None),
            ("ref_build_table", [C.c_int32, _dp, _dp, C.c_int32, C.c_double, _dp,
                                 P(abi.Pair)], None),
            ("ref_spline", [C.c_int32, _dp, _dp, _dp, _dp, _dp], None),
            ("ref_notrans", [], None),
            ("ref_get_dids", [_dp], None),
            ("ref_write_map", [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                               C.c_double, C.c_double, C.c_double, _dp, P(C.c_float)], None),
            ("ref_readphi", [C.c_char_p, C.c_int32, C.c_int32, _dp, _ip, _dp, P(C.c_float)],
             None),
            ("ref_repwalls", [C.c_int32, C.c_int32, C.c_int32, P(C.c_int8), C.c_int32], None),
            ("ref_readrfpar", [C.c_int32, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, _dp, _ip,
                               _dp, _dp, P(C.c_float), P(C.c_float)], None),
            ("ref_metropolis", [C.c_int32, C.c_double, C.c_double, _dp, C.c_int64,
                                P(C.c_int64), _dp], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        if path in (SHIM_SO, FRONT_GPU_SO):
            L.bromoc_gpu_reset.argtypes = []
            L.bromoc_gpu_reset.restype = None
            L.bromoc_gpu_set_seed.argtypes = [C.c_int64, C.c_int64]
            L.bromoc_gpu_set_seed.restype = None
        _rlibs[path] = L
    return _rlibs[path]


def batch_uniforms(seed, cycle, r, n):
    """The first n GCMC uniforms of replica r in GCMC cycle `cycle` of a replica batch:
    uniform k is word k%4 of Philox4x64-10(counter (cycle, k/4, r, 1), key (seed, 0)),
    (w >> 11) * 2^-53."""
    L = oracle_lib()
    nb = (n + 3) // 4
    out = np.zeros(4 * nb)
    key = np.array([seed, 0], np.uint64)
    w = np.zeros(4, np.uint64)
    for b in range(nb):
        ctr = np.array([cycle, b, r, 1], np.uint64)
        L.orc_philox4x64(_p(ctr, C.c_uint64), _p(key, C.c_uint64), _p(w, C.c_uint64))
        out[4 * b:4 * b + 4] = (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return out[:n]


def _spline_with(fn, x, y):
    x = np.ascontiguousarray(x, float)
    y = np.ascontiguousarray(y, float)
    n = len(x)
    b, c, d = np.zeros(n), np.zeros(n), np.zeros(n)
    fn(n, _p(x), _p(y), _p(b), _p(c), _p(d))
    return b, c, d


def oracle_spline(x, y):
    """(b, c, d) of spline.f90 from the C restatement."""
    return _spline_with(oracle_lib().orc_spline, x, y)


def ref_spline(x, y):
    """(b, c, d) from the reference's own SPLINE."""
    return _spline_with(ref_lib().ref_spline, x, y)


def oracle_discretize(pair, nnp, mnp, res, rth=0.0, srpx=0.0, srpk=0.0, srpy=0.0):
    """DISCRETIZE restated: (ep [3n], sc1, sc2, sc3) for an abi.Pair's parameters."""
    ep = np.zeros(3 * (mnp - nnp + 1))
    out = abi.Pair()
    rc = oracle_lib().orc_discretize(C.byref(pair), nnp, mnp, res, rth, srpx, srpk, srpy,
                                     _p(ep), C.byref(out))
    if rc:
        raise RuntimeError(f"orc_discretize {rc}")
    return ep, (out.sc1, out.sc2, out.sc3)


def ref_discretize(pair, nnp, mnp, res, rth=0.0, srpx=0.0, srpk=0.0, srpy=0.0):
    """The reference's DISCRETIZE (oracle/_ref)."""
    ep = np.zeros(3 * (mnp - nnp + 1))
    out = abi.Pair()
    ref_lib().ref_discretize(C.byref(pair), nnp, mnp, res, rth, srpx, srpk, srpy, _p(ep),
                             C.byref(out))
    return ep, (out.sc1, out.sc2, out.sc3)


def ref_table_builder(n, r, u, chr_, fct):
    ep = np.zeros(3 * n)
    p = abi.Pair()
    ref_lib().ref_build_table(n, _p(r), _p(u), chr_, fct, _p(ep), C.byref(p))
    return ep, p


class RefFortran(_Base):
    """The reference routines (module-global state: one live instance at a time)."""

    def __init__(self, system, lib=REF_SO):
        self.L = ref_lib(lib)
        if lib in (SHIM_SO, FRONT_GPU_SO):
            self.L.bromoc_gpu_reset()
        self.system = system
        self.L.ref_set_nindex(system.ndna, system.nion)
        st, keep = system.to_struct()
        self.L.ref_init(C.byref(st))
        self.ntot = system.ntot
        self._noise = None

    def _get_forces(self, *a):
        return self.L.ref_get_forces(*a)

    def _get_positions(self, *a):
        return self.L.ref_get_positions(*a)

    def set_positions(self, pos, typei=None):
        pos = np.ascontiguousarray(pos, float)
        n = len(pos)
        t = np.ascontiguousarray(typei if typei is not None else self.positions()[1][:n], np.int32)
        x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
        self.L.ref_set_positions(n, _p(x), _p(y), _p(z), _p(t, C.c_int32))
        self.ntot = n

    def energy(self, forces=True):
        e = np.zeros(23)
        self.L.ref_energy(1 if forces else 0, _p(e))
        return dict(zip(abi.ENERGY_FIELDS, e.tolist()))

    def upload_noise(self, noise):
        self._noise = np.ascontiguousarray(noise, float)

    def dynamics(self, mode=abi.NOISE_INJECTED, row=0, seed=0, counter=0):
        if mode == abi.NOISE_INJECTED:
            g = np.ascontiguousarray(self._noise[row][: 3 * self.ntot])
        else:
            g = np.zeros(3 * self.ntot)
            L = oracle_lib()
            for i in range(self.ntot):
                L.orc_philox_gauss3(seed, counter, 0, i, _p(g[3 * i:3 * i + 3]))
        self.L.ref_dynamics(_p(g))

    def step(self, nsteps, mode=abi.NOISE_INJECTED, first_row=0, seed=0, counter0=0):
        e = None
        for s in range(nsteps):
            e = self.energy(True)
            self.dynamics(mode, first_row + s, seed, counter0 + s)
            self.L.ref_notrans()
        return e

    def dids(self):
        d = np.zeros(5 * self.system.capacity)
        self.L.ref_get_dids(_p(d))
        return d[:5 * self.ntot].reshape(-1, 5)

    def metropolis(self, nmcm, mcmax, ikbt, uniforms, ener=0.0):
        u = np.ascontiguousarray(uniforms, float)
        used, e = C.c_int64(), C.c_double(ener)
        self.L.ref_metropolis(nmcm, mcmax, ikbt, _p(u), len(u), C.byref(used), C.byref(e))
        return used.value, nmcm, e.value

    def flux(self):
        s = self.system
        nt, np_ = s.ntype - s.nold, len(s.zcont)
        nf = np.zeros(max(nt * np_, 1), np.int32)
        nb = np.zeros(max(nt * np_, 1), np.int32)
        self.L.ref_get_flux(_p(nf, C.c_int32), _p(nb, C.c_int32))
        return nf[:nt * np_].reshape(nt, np_), nb[:nt * np_].reshape(nt, np_)

    def warn(self):
        w = np.zeros(self.system.nttyp + 1, np.int32)
        self.L.ref_get_warn(_p(w, C.c_int32))
        return w

    def interact(self, xyz, jtype, j, alert=False):
        d = C.c_double()
        self.L.ref_interact(float(xyz[0]), float(xyz[1]), float(xyz[2]), int(jtype), int(j),
                            1 if alert else 0, C.byref(d))
        return d.value

    def set_gcmc(self, g, ibuffer):
        st, self._gkeep = g.to_struct()
        ib = np.ascontiguousarray(ibuffer, np.int32)
        self.L.ref_set_gcmc(C.byref(st), _p(ib, C.c_int32))

    def ibuffer(self):
        ib = np.zeros(self.ntot, np.int32)
        self.L.ref_get_ibuffer(_p(ib, C.c_int32))
        return ib

    def grand(self, ngcmc, uniforms, nat, prob, icycle, ener):
        u = np.ascontiguousarray(uniforms, float)
        used = C.c_int64()
        e = C.c_double(ener)
        pstride = prob.shape[1]
        self.L.ref_grand(ngcmc, _p(u), len(u), C.byref(used), _p(nat, C.c_int32),
                         _p(prob, C.c_int32), pstride, icycle, C.byref(e))
        nn = C.c_int32()
        x = np.zeros(self.system.capacity)
        t = np.zeros(self.system.capacity, np.int32)
        self.L.ref_get_positions(_p(x), _p(x.copy()), _p(x.copy()), _p(t, C.c_int32),
                                 C.byref(nn))
        self.ntot = nn.value
        return used.value, e.value


# ------------------------------------------------------------------ map files
def oracle_repwalls(phiv, walls=126):
    """REPWALLS restated in C (oracle/bromoc_oracle.c)."""
    out = np.ascontiguousarray(phiv, np.int8).copy()
    nx, ny, nz = out.shape
    oracle_lib().orc_repwalls(nx, ny, nz, out.ctypes.data_as(C.POINTER(C.c_int8)), walls)
    return out


def ref_repwalls(phiv, walls=126):
    """REPWALLS of the reference build."""
    out = np.ascontiguousarray(phiv, np.int8).copy()
    nx, ny, nz = out.shape
    ref_lib().ref_repwalls(nx, ny, nz, out.ctypes.data_as(C.POINTER(C.c_int8)), walls)
    return out


def ref_write_map(path, values, dcel, centre=(0.0, 0.0, 0.0), rec2=(80.0, 2.0, 0.0, 0.0, 0.0, 2.0)):
    """A grid file written by Fortran's own unformatted write (the reference build's
    compiler), in the record layout READPHI reads."""
    v = np.ascontiguousarray(values, np.float32)
    nx, ny, nz = v.shape
    r2 = np.ascontiguousarray(rec2, float)
    ref_lib().ref_write_map(str(path).encode(), nx, ny, nz, dcel, *centre, _p(r2),
                            v.ctypes.data_as(C.POINTER(C.c_float)))


def ref_readphi(path, kind, adjust=False, box=(0, 0, 0, 0, 0, 0), maxn=None):
    """READPHI of the reference build: (dims, (dcel, xbcen, ybcen, zbcen), values)."""
    b = np.ascontiguousarray(box, float)
    dims = np.zeros(3, np.int32)
    geom = np.zeros(4)
    out = np.zeros(maxn or (1 << 22), np.float32)
    ref_lib().ref_readphi(str(path).encode(), kind, 1 if adjust else 0, _p(b), _p(dims, C.c_int32),
                          _p(geom), out.ctypes.data_as(C.POINTER(C.c_float)))
    n = int(np.prod(dims))
    return tuple(int(d) for d in dims), tuple(geom), out[:n].reshape(tuple(dims))


def ref_readrfpar(paths, single=False, adjust=False, box=(0, 0, 0, 0, 0, 0), maxn=1 << 22):
    """READRFPAR of the reference build on one or two files."""
    b = np.ascontiguousarray(box, float)
    dims = np.zeros(3, np.int32)
    geom = np.zeros(4)
    rad = np.zeros(2)
    gs = np.zeros(maxn, np.float32)
    gr = np.zeros(maxn, np.float32)
    p2 = paths[1] if len(paths) > 1 else paths[0]
    fp = C.POINTER(C.c_float)
    ref_lib().ref_readrfpar(len(paths), str(paths[0]).encode(), str(p2).encode(),
                            1 if single else 0, 1 if adjust else 0, _p(b), _p(dims, C.c_int32),
                            _p(geom), _p(rad), gs.ctypes.data_as(fp), gr.ctypes.data_as(fp))
    n = int(np.prod(dims)) * len(paths)
    shape = (len(paths),) + tuple(int(d) for d in dims)
    return tuple(int(d) for d in dims), tuple(geom), gs[:n].reshape(shape), gr[:n].reshape(shape)

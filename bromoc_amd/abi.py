"""ctypes mirror of include/bromoc_gpu.h and the loader of the HIP library.

The structs below must match the C header field for field; tests/test_abi.py checks
sizes and offsets against the compiled library.
"""
from __future__ import annotations

import ctypes as C
import os

ABI_VERSION = 5

OK, E_ARG, E_UNSUPPORTED, E_CAPACITY, E_HIP, E_NODEVICE, E_STATE = range(7)
STATUS_NAMES = {0: "OK", 1: "E_ARG", 2: "E_UNSUPPORTED", 3: "E_CAPACITY", 4: "E_HIP",
                5: "E_NODEVICE", 6: "E_STATE"}

QPAR, QNUCL, QNONBOND, QNOBOND = 1 << 0, 1 << 1, 1 << 2, 1 << 3
QPHIX, QPHIV, QSVDW, QSOLV = 1 << 4, 1 << 5, 1 << 6, 1 << 7
QDEBY, QCOUNTION, QDNAFREE = 1 << 8, 1 << 9, 1 << 10
QMEMB, QMMIJ, QBSPLINE, QNMCDEN = 1 << 16, 1 << 17, 1 << 18, 1 << 19
QDEBYHYB, QPROXDIFF, QNOTRANS, QAPFOR = 1 << 20, 1 << 21, 1 << 22, 1 << 23
QCONTRANS, QRFPAR, QSPHERE, QECYL = 1 << 24, 1 << 25, 1 << 26, 1 << 27
QDIFFUSE, QPROFILE, QRFPSIN = 1 << 28, 1 << 29, 1 << 30

SITE_S, SITE_P, SITE_AB, SITE_TB, SITE_CB, SITE_GB = range(6)
NOISE_INJECTED, NOISE_PHILOX = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)


class Pair(C.Structure):
    _fields_ = [("efpot", C.c_int32), ("chr", C.c_int32), ("lj", C.c_int32),
                ("srpmf", C.c_int32), ("nxi", C.c_int32), ("nknots", C.c_int32),
                ("fct", C.c_double), ("epp4", C.c_double), ("sgp2", C.c_double),
                ("c0", C.c_double), ("c1", C.c_double), ("c2", C.c_double),
                ("c3", C.c_double), ("c4", C.c_double),
                ("dmi", C.c_double), ("dm2lo", C.c_double), ("dm2hi", C.c_double),
                ("sc1", C.c_double), ("sc2", C.c_double), ("sc3", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nclx", C.c_int32), ("ncly", C.c_int32), ("nclz", C.c_int32),
                ("nmaps", C.c_int32), ("dcel", C.c_double), ("xbcen", C.c_double),
                ("ybcen", C.c_double), ("zbcen", C.c_double), ("data", C.c_void_p)]


class SystemStruct(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("ntot", C.c_int32), ("nsites", C.c_int32),
        ("nfix", C.c_int32), ("capacity", C.c_int32), ("ntype", C.c_int32),
        ("nold", C.c_int32), ("nttyp", C.c_int32), ("npair", C.c_int32),
        ("flags", C.c_uint32),
        ("x", _dp), ("y", _dp), ("z", _dp), ("typei", _ip), ("typtyp", _ip),
        ("namsite", _ip), ("stfree", _bp),
        ("nwtype", _ip), ("cg", _dp), ("fact1a", _dp), ("fact2a", _dp), ("pairidx", _ip),
        ("pairs", C.POINTER(Pair)),
        ("nknots", C.c_int32), ("reserved0", C.c_int32), ("ep", _dp),
        ("res", C.c_double), ("ires", C.c_double),
        ("rth", C.c_double), ("srpx", C.c_double), ("srpk", C.c_double), ("srpy", C.c_double),
        ("cgnuc", C.c_double), ("epsnuc", C.c_double), ("scalepairing", C.c_double),
        ("fctn", C.c_double), ("kappa", C.c_double), ("ikappa", C.c_double),
        ("epsolv", C.c_double),
        ("nbond", C.c_int32), ("nangle", C.c_int32), ("ndihe", C.c_int32),
        ("nstack", C.c_int32), ("nbp", C.c_int32), ("nex", C.c_int32), ("nqq", C.c_int32),
        ("nsolv", C.c_int32),
        ("sitebond", _ip), ("distbond", _dp), ("siteangle", _ip), ("valangle", _dp),
        ("sitedihe", _ip), ("valdihe", _dp), ("sitestack", _ip), ("sgstack", _dp),
        ("sitebp", _ip), ("sgbp", _dp), ("siteex", _ip), ("sgex", _dp),
        ("siteqq", _ip), ("siteslv", _ip),
        ("phix", Grid), ("phiv", Grid),
        ("svdw", C.c_double), ("thold8", C.c_double), ("vzmin", C.c_double),
        ("vzmax", C.c_double), ("scal", _dp),
        ("bdmax", C.c_double), ("fact0n1", C.c_double), ("fact0n2", C.c_double),
        ("lx2m", C.c_double), ("lx2p", C.c_double), ("ly2m", C.c_double),
        ("ly2p", C.c_double), ("lz2m", C.c_double), ("lz2p", C.c_double),
        ("cntpts", C.c_int32), ("reserved1", C.c_int32), ("zcont", _dp),
        ("dt", C.c_double), ("kbtdt", C.c_double), ("diffusion", _dp),
        ("plength2", C.c_double), ("pcenter", C.c_double), ("ampl3", _dp), ("p3", _dp),
        ("nspline", C.c_int32), ("reserved2", C.c_int32),
        ("xs", _dp), ("ys", _dp), ("bs", _dp), ("cs", _dp), ("ds", _dp),
        ("rf", Grid), ("gsrfen", C.POINTER(C.c_float)), ("greff", C.POINTER(C.c_float)),
        ("sqrfac", C.c_double),
        ("afn", C.c_int32), ("ctn", C.c_int32), ("sn", _ip), ("af", _dp),
        ("csn", _ip), ("kx", _dp), ("ky", _dp), ("kz", _dp),
        ("contrx", _dp), ("contry", _dp), ("contrz", _dp),
        ("unsplit", C.c_int32), ("nsites1st", C.c_int32),
        ("notrans_axes", C.c_int32), ("reserved3", C.c_int32),
        ("notrx", C.c_double), ("notry", C.c_double), ("notrz", C.c_double),
        ("ibeta", C.c_double), ("diff0", C.c_double), ("diffcutoff", C.c_double),
        ("thold27", C.c_double),
        ("qpore", C.c_int32), ("reserved4", C.c_int32),
        ("voltage", C.c_double), ("afact", C.c_double), ("ceps", C.c_double),
        ("zmemb1", C.c_double), ("zmemb2", C.c_double),
        ("ampl1", _dp), ("p1", _dp), ("ampl2", _dp), ("p2", _dp), ("rcylinder", _dp),
        ("mmij_shape", C.c_int32), ("xnpol", C.c_int32), ("ynpol", C.c_int32),
        ("znpol", C.c_int32), ("ntpol", C.c_int32), ("reserved5", C.c_int32),
        ("lstpx", _ip), ("lstpy", _ip), ("lstpz", _ip), ("bnorm", _dp), ("mmij", _dp),
        ("xscal", C.c_double), ("yscal", C.c_double), ("zscal", C.c_double),
        ("rfscal", C.c_double),
        ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double), ("rsphe", C.c_double),
        ("rsphe2", C.c_double), ("iecx", C.c_double), ("iecy", C.c_double),
    ]


ENERGY_FIELDS = ["ener", "eelec", "evdw", "esrpmf", "esrpmfmx", "eefpot", "eefpotmx", "eqqmx",
                 "evdwmx", "ebond", "eang", "edihe", "estack", "ebp", "eex", "eqq", "esolv",
                 "egsbpa", "egsbpb", "evdwgd", "erfpar", "econ", "ememb"]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in ENERGY_FIELDS]

    def as_dict(self):
        return {n: getattr(self, n) for n in ENERGY_FIELDS}


class GcmcStruct(C.Structure):
    _fields_ = [("nbuffer", C.c_int32), ("reserved", C.c_int32),
                ("ibfftyp", _ip), ("avnum", _dp), ("mu", _dp), ("kb", _dp),
                ("lzmin", _dp), ("lzmax", _dp), ("bufferbias", _bp), ("ntotat", _ip),
                ("lx", C.c_double), ("ly", C.c_double), ("cx", C.c_double),
                ("cy", C.c_double), ("cz", C.c_double), ("ikbt", C.c_double),
                ("rmin", _dp), ("rmax", _dp)]


class ReplicaStruct(C.Structure):
    _fields_ = [("ntot", C.c_int32), ("nfix", C.c_int32), ("capacity", C.c_int32),
                ("reserved", C.c_int32), ("x", _dp), ("y", _dp), ("z", _dp), ("typei", _ip),
                ("phix", C.POINTER(C.c_float)), ("gcmc", C.POINTER(GcmcStruct)),
                ("ibuffer", _ip), ("nat", _ip)]


class MapHeader(C.Structure):
    _fields_ = [("nclx", C.c_int32), ("ncly", C.c_int32), ("nclz", C.c_int32),
                ("reserved", C.c_int32), ("dcel", C.c_double), ("xbcen", C.c_double),
                ("ybcen", C.c_double), ("zbcen", C.c_double), ("rec2", C.c_double * 6)]


MAP_PHIX, MAP_PHIV = 0, 1


class Analysis(C.Structure):
    """bromoc_analysis: the per-cycle statistics of simul1.f90 (rho, rDNA, g(r), prob2/prob3,
    virial pressure)."""
    _fields_ = [("what", C.c_int32), ("nzmax", C.c_int32), ("zmini", C.c_double),
                ("idelz", C.c_double), ("igr", C.c_int32), ("npoints", C.c_int32),
                ("dr", C.c_double), ("datom", C.c_int32), ("reserved", C.c_int32),
                ("czmin", C.c_double), ("czmax", C.c_double), ("tvol", C.c_double),
                ("temp", C.c_double), ("ikbt", C.c_double)]


AN_RHO, AN_RDNA, AN_GR, AN_PROB, AN_PRES = 1, 2, 4, 8, 16


class BromocError(RuntimeError):
    def __init__(self, where, code, detail=""):
        self.code = code
        super().__init__(f"{where}: {STATUS_NAMES.get(code, code)} {detail}".strip())


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libbromoc_gpu.so")

_lib = None


def load_library(path: str | None = None):
    """Load libbromoc_gpu.so (the HIP path).  Raises if it is not built: there is no
    fallback implementation behind this package."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = C.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib


def _declare(lib):
    vp = C.c_void_p
    P = C.POINTER
    sig = {
        "bromoc_create": ([P(SystemStruct), C.c_int, P(vp)], C.c_int),
        "bromoc_destroy": ([vp], C.c_int),
        "bromoc_last_error": ([vp], C.c_char_p),
        "bromoc_set_positions": ([vp, C.c_int32, _dp, _dp, _dp, _ip], C.c_int),
        "bromoc_get_positions": ([vp, _dp, _dp, _dp, _ip, _ip], C.c_int),
        "bromoc_get_forces": ([vp, _dp, _dp, _dp], C.c_int),
        "bromoc_energy": ([vp, C.c_int32, P(Energies)], C.c_int),
        "bromoc_upload_noise": ([vp, C.c_int32, _dp], C.c_int),
        "bromoc_step": ([vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_uint64,
                         P(Energies)], C.c_int),
        "bromoc_dynamics": ([vp, C.c_int32, C.c_int32, C.c_uint64, C.c_uint64], C.c_int),
        "bromoc_get_flux": ([vp, _ip, _ip], C.c_int),
        "bromoc_get_warn": ([vp, _ip], C.c_int),
        "bromoc_interact": ([vp, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                             C.c_int32, _dp], C.c_int),
        "bromoc_set_gcmc": ([vp, P(GcmcStruct)], C.c_int),
        "bromoc_grand": ([vp, C.c_int32, _dp, C.c_int64, P(C.c_int64), _ip, _ip, C.c_int64,
                          _dp], C.c_int),
        "bromoc_read_map": ([C.c_char_p, C.c_int32, C.c_int32, _dp, P(MapHeader), C.c_void_p,
                             C.c_int64], C.c_int),
        "bromoc_read_rfpar": ([C.c_int32, P(C.c_char_p), C.c_int32, _dp, P(MapHeader), _dp,
                               P(C.c_float), P(C.c_float), C.c_int64], C.c_int),
        "bromoc_repwalls": ([C.c_int32, C.c_int32, C.c_int32, P(C.c_int8), C.c_int32, C.c_int],
                            C.c_int),
        "bromoc_discretize": ([P(Pair), C.c_int32, C.c_int32, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_double, _dp, P(Pair)], C.c_int),
        "bromoc_update_tables": ([vp, C.c_int32, _ip, _ip, _dp, _dp], C.c_int),
        "bromoc_table_vjp_device": ([vp, C.c_int32, _dp, _dp, _dp, _dp], C.c_int),
        "bromoc_table_vjp_device_noalloc": ([vp, C.c_int32, _dp, _dp, _dp, _dp], C.c_int),
        "bromoc_set_analysis": ([vp, P(Analysis)], C.c_int),
        "bromoc_analyze": ([vp], C.c_int),
        "bromoc_get_analysis": ([vp, P(C.c_int64), P(C.c_int64), P(C.c_int64), _ip,
                                 P(C.c_int64), P(C.c_int64), _dp, P(C.c_int64)], C.c_int),
        "bromoc_metropolis": ([vp, C.c_int32, C.c_double, C.c_double, _dp, C.c_int64,
                               P(C.c_int64), _ip, _dp], C.c_int),
        "bromoc_build_table": ([C.c_int32, _dp, _dp, C.c_int32, C.c_double, _dp, P(Pair)],
                               C.c_int),
        "bromoc_set_ibuffer": ([vp, _ip], C.c_int),
        "bromoc_get_tuning": ([vp, _ip], C.c_int),
        "bromoc_spatial_order": ([vp, P(C.c_uint32), _ip, _ip, _ip, _ip], C.c_int),
        "bromoc_time_steps": ([vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_uint64,
                               _dp, _dp], C.c_int),
        "bromoc_get_ibuffer": ([vp, _ip], C.c_int),
        "bromoc_philox4x64": ([C.c_int32, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)], C.c_int),
        "bromoc_spline": ([C.c_int32, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "bromoc_batch_create": ([P(SystemStruct), C.c_int32, P(ReplicaStruct), C.c_int, P(vp)],
                                C.c_int),
        "bromoc_batch_destroy": ([vp], C.c_int),
        "bromoc_batch_last_error": ([vp], C.c_char_p),
        "bromoc_batch_run": ([vp, C.c_int32, C.c_uint64, C.c_uint64, C.c_int32, C.c_int32,
                              C.c_uint64, P(Energies)], C.c_int),
        "bromoc_batch_energy": ([vp, P(Energies)], C.c_int),
        "bromoc_batch_get": ([vp, C.c_int32, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _dp], C.c_int),
        "bromoc_batch_get_forces": ([vp, C.c_int32, _dp, _dp, _dp], C.c_int),
        "bromoc_batch_get_flux": ([vp, C.c_int32, _ip, _ip], C.c_int),
        "bromoc_batch_get_prob": ([vp, C.c_int32, _ip], C.c_int),
        "bromoc_batch_time_steps": ([vp, C.c_int32, C.c_uint64, C.c_uint64, _dp, _dp], C.c_int),
        "bromoc_batch_get_tuning": ([vp, _ip, _ip], C.c_int),
    }
    for name, (args, res) in sig.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res


def exported_symbols():
    """Names every build of libbromoc_gpu.so must export (the declarations of
    include/bromoc_gpu.h)."""
    hdr = os.path.join(os.path.dirname(PKG_DIR), "include", "bromoc_gpu.h")
    import re
    names = []
    with open(hdr) as f:
        for line in f:
            m = re.match(r"^\s*(?:int|const char \*|void)\s*\*?\s*(bromoc_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names

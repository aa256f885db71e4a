"""Host-side description of one BROMOC system (the state shell_simul.f90 builds before
SIMUL) and its conversion to the C ABI struct.

Arrays use the ABI conventions of include/bromoc_gpu.h: 0-based particle indices, 1-based
type tables (entry 0 unused), component-major site lists.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

# constamod.f90:21-36
PI = 3.14159265358979323846264338327950288419716939937510
KBOLTZ = 1.380648813E-23
AVOGADRO = 6.0221412927E23
KCAL = 4184.0
KCALMOL = KCAL / AVOGADRO
COULOMB = 1.60217656535E-19
EPERMIT = 8.854187817620E-12
ANGSTROM = 1E-10
LITER = 1E-3
CELEC = COULOMB ** 2 / (4.0 * PI * EPERMIT * KCALMOL * ANGSTROM)


def kbt_of(temp):
    """kBT in kcal/mol (shell_simul.f90:865)."""
    return KBOLTZ * temp / KCALMOL


@dataclass
class GridData:
    data: np.ndarray          # (nclx, ncly, nclz), float32 (phix) or int8 (phiv);
    dcel: float               # per-species phiv maps (NMCDEN): (nmaps, nclx, ncly, nclz)
    xbcen: float = 0.0
    ybcen: float = 0.0
    zbcen: float = 0.0

    @property
    def shape(self):
        return self.data.shape[-3:]

    def tran(self):
        n = self.data.shape
        return tuple(0.5 * (k - 1) * self.dcel for k in n)


@dataclass
class System:
    ntot: int
    nsites: int
    nfix: int
    ntype: int
    nold: int
    nttyp: int
    flags: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    typei: np.ndarray
    nwtype: np.ndarray
    cg: np.ndarray
    fact1a: np.ndarray
    fact2a: np.ndarray
    pairidx: np.ndarray
    pairs: list                     # list of dicts, index 0 unused (None)
    ep: np.ndarray
    res: float = 0.01
    capacity: int = 0
    typtyp: np.ndarray = None
    namsite: np.ndarray = None
    stfree: np.ndarray = None
    rth: float = 0.0
    srpx: float = 0.0
    srpk: float = 0.0
    srpy: float = 0.0
    cgnuc: float = -1.0
    epsnuc: float = 0.184
    scalepairing: float = 1.0
    fctn: float = 0.0
    kappa: float = 0.0
    ikappa: float = 0.0
    epsolv: float = 0.0
    lists: dict = field(default_factory=dict)   # name -> (sites int32 [k*n], vals or None)
    phix: GridData | None = None
    phiv: GridData | None = None
    svdw: float = 50.0
    thold8: float = 8.0
    vzmin: float = 0.0
    vzmax: float = 0.0
    scal: np.ndarray = None
    bdmax: float = 1.0
    fact0n1: float = 0.0
    fact0n2: float = 0.0
    box: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)   # lx2m, lx2p, ly2m, ly2p, lz2m, lz2p
    # curved systems (QSPHERE / QECYL): centre (cx, cy, cz), sphere radius rsphe (rsphe2 =
    # rsphe**2 as the reference holds it), inverse semi-axes iecx, iecy of the cylinder
    center: tuple = (0.0, 0.0, 0.0)
    rsphe: float = 0.0
    iecx: float = 0.0
    iecy: float = 0.0
    zcont: np.ndarray = None
    # position-dependent diffusion (DYNAMICS1 / DYNAMICS2)
    dt: float = 0.0
    kbtdt: float = 0.0
    diffusion: np.ndarray = None     # [ntype+1]
    plength2: float = 0.0
    pcenter: float = 0.0
    ampl3: np.ndarray = None         # [ntype+1]
    p3: np.ndarray = None            # [ntype+1]
    spline: tuple = None             # (xs, ys, b, c, d) of the PROFILE command
    # reaction-field parameter grids: rf.data is gsrfen (ngrid, nx, ny, nz) float32
    rf: GridData | None = None
    rf_greff: np.ndarray = None      # (ngrid, nx, ny, nz) float32
    sqrfac: float = 1.0
    # APFOR: (sites [afn] 0-based, forces [afn, 3])
    apfor: tuple = None
    # CONTRANS: dict(csn [ctn] 1-based with 0 = centroid, k [ctn, 3], centre [ctn+1, 3],
    # unsplit, nsites1st)
    contrans: dict = None
    # NOTRANS: (axes bitmask, (notrx, notry, notrz))
    notrans: tuple = None
    # PROXDIFF: 1/beta, diff0, diffcutoff
    ibeta: float = 0.0
    diff0: float = 0.0
    diffcutoff: float = 0.0
    thold27: float = 1.0                 # B-spline wall threshold
    # MEMBRANE: dict(voltage, afact, ceps, zmemb1, zmemb2, qpore, and per atnam2 type
    # [nttyp+1] ampl1, p1, ampl2, p2, rcylinder)
    membrane: dict = None
    # MMIJ reaction field, RECTBOX: dict(npol (xnpol, ynpol, znpol), lstp [ntpol, 3] degrees,
    # bnorm [ntpol], mmij [ntpol, ntpol] symmetric, scal (xscal, yscal, zscal), rfscal)
    mmij: dict = None
    ndna: int = 0
    nion: int = 0
    meta: dict = field(default_factory=dict)

    LIST_NAMES = ("bond", "angle", "dihe", "stack", "bp", "ex", "qq", "slv")
    LIST_ARITY = {"bond": 2, "angle": 3, "dihe": 4, "stack": 2, "bp": 2, "ex": 2, "qq": 2,
                  "slv": 2}

    def __post_init__(self):
        if not self.capacity:
            self.capacity = self.ntot
        ns = self.nsites
        if self.typtyp is None:
            self.typtyp = np.zeros(ns, np.int32)
        if self.namsite is None:
            self.namsite = np.zeros(ns, np.int32)
        if self.stfree is None:
            self.stfree = np.ones(ns, np.int8)
        if self.scal is None:
            self.scal = np.ones(self.nttyp + 1)
        if self.zcont is None:
            self.zcont = np.zeros(0)
        for n in self.LIST_NAMES:
            self.lists.setdefault(n, (np.zeros(0, np.int32), None))

    @property
    def npair(self):
        return len(self.pairs) - 1

    def count(self, name):
        return len(self.lists[name][0]) // self.LIST_ARITY[name]

    def copy(self):
        import copy
        return copy.deepcopy(self)

    def nmcden_maps(self):
        """Repulsion maps an NMCDEN phiv holds (readphi.f90:176-240): the six site maps
        with CG-DNA, then one per ion type nold+1..ntype with ions."""
        return ((6 if self.flags & abi.QNUCL else 0) +
                (self.ntype - self.nold if self.flags & abi.QPAR else 0))

    # ------------------------------------------------------------ C struct
    def to_struct(self):
        """Returns (SystemStruct, keepalive) -- keep both alive while C reads them."""
        keep = []

        def arr(a, dtype, ctype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        s = abi.SystemStruct()
        s.abi_version = abi.ABI_VERSION
        s.ntot, s.nsites, s.nfix = self.ntot, self.nsites, self.nfix
        s.capacity = self.capacity
        s.ntype, s.nold, s.nttyp = self.ntype, self.nold, self.nttyp
        s.npair = self.npair
        s.flags = self.flags
        s.x = arr(self.x[:self.ntot], np.float64, C.c_double)
        s.y = arr(self.y[:self.ntot], np.float64, C.c_double)
        s.z = arr(self.z[:self.ntot], np.float64, C.c_double)
        s.typei = arr(self.typei[:self.ntot], np.int32, C.c_int32)
        s.typtyp = arr(self.typtyp, np.int32, C.c_int32)
        s.namsite = arr(self.namsite, np.int32, C.c_int32)
        s.stfree = arr(self.stfree, np.int8, C.c_int8)
        s.nwtype = arr(self.nwtype, np.int32, C.c_int32)
        s.cg = arr(self.cg, np.float64, C.c_double)
        s.fact1a = arr(self.fact1a, np.float64, C.c_double)
        s.fact2a = arr(self.fact2a, np.float64, C.c_double)
        s.pairidx = arr(self.pairidx, np.int32, C.c_int32)
        parr = (abi.Pair * (self.npair + 1))()
        for k, p in enumerate(self.pairs):
            if p is None:
                continue
            for name, _ in abi.Pair._fields_:
                if name in p:
                    setattr(parr[k], name, p[name])
        keep.append(parr)
        s.pairs = C.cast(parr, C.POINTER(abi.Pair))
        s.nknots = len(self.ep) // 3
        s.ep = arr(self.ep, np.float64, C.c_double)
        s.res, s.ires = self.res, 1.0 / self.res
        s.rth, s.srpx, s.srpk, s.srpy = self.rth, self.srpx, self.srpk, self.srpy
        s.cgnuc, s.epsnuc, s.scalepairing = self.cgnuc, self.epsnuc, self.scalepairing
        s.fctn, s.kappa, s.ikappa, s.epsolv = self.fctn, self.kappa, self.ikappa, self.epsolv
        cnt = {n: self.count(n) for n in self.LIST_NAMES}
        s.nbond, s.nangle, s.ndihe = cnt["bond"], cnt["angle"], cnt["dihe"]
        s.nstack, s.nbp, s.nex, s.nqq, s.nsolv = (cnt["stack"], cnt["bp"], cnt["ex"],
                                                  cnt["qq"], cnt["slv"])
        for name, sfield, vfield in (("bond", "sitebond", "distbond"),
                                     ("angle", "siteangle", "valangle"),
                                     ("dihe", "sitedihe", "valdihe"),
                                     ("stack", "sitestack", "sgstack"),
                                     ("bp", "sitebp", "sgbp"), ("ex", "siteex", "sgex"),
                                     ("qq", "siteqq", None), ("slv", "siteslv", None)):
            sites, vals = self.lists[name]
            setattr(s, sfield, arr(sites if len(sites) else np.zeros(1), np.int32, C.c_int32))
            if vfield:
                v = vals if vals is not None and len(vals) else np.zeros(1)
                setattr(s, vfield, arr(v, np.float64, C.c_double))
        for gname, dtype in (("phix", np.float32), ("phiv", np.int8)):
            g = getattr(self, gname)
            cg = getattr(s, gname)
            if g is None:
                continue
            d = np.ascontiguousarray(g.data, dtype=dtype)
            # the library checks the map count against the flags; a 4-D phiv must hold
            # exactly the maps NMCDEN reads
            assert d.ndim in ((3, 4) if gname == "phiv" else (3,)), (gname, d.shape)
            if d.ndim == 4:
                assert self.flags & abi.QNMCDEN and d.shape[0] == self.nmcden_maps(), (
                    f"phiv of {d.shape[0]} maps; NMCDEN reads {self.nmcden_maps()}")
            keep.append(d)
            cg.nclx, cg.ncly, cg.nclz = d.shape[-3:]
            cg.nmaps = (d.shape[0] if d.ndim == 4 else 1) if gname == "phiv" else 0
            cg.dcel, cg.xbcen, cg.ybcen, cg.zbcen = g.dcel, g.xbcen, g.ybcen, g.zbcen
            cg.data = d.ctypes.data
        s.svdw, s.thold8, s.vzmin, s.vzmax = self.svdw, self.thold8, self.vzmin, self.vzmax
        s.scal = arr(self.scal, np.float64, C.c_double)
        s.bdmax, s.fact0n1, s.fact0n2 = self.bdmax, self.fact0n1, self.fact0n2
        (s.lx2m, s.lx2p, s.ly2m, s.ly2p, s.lz2m, s.lz2p) = self.box
        s.cntpts = len(self.zcont)
        s.zcont = arr(self.zcont if len(self.zcont) else np.zeros(1), np.float64, C.c_double)
        s.dt, s.kbtdt = self.dt, self.kbtdt
        nt1 = self.ntype + 1
        one = np.zeros(1)
        s.diffusion = arr(self.diffusion if self.diffusion is not None else np.zeros(nt1),
                          np.float64, C.c_double)
        s.plength2, s.pcenter = self.plength2, self.pcenter
        s.ampl3 = arr(self.ampl3 if self.ampl3 is not None else np.zeros(nt1), np.float64, C.c_double)
        s.p3 = arr(self.p3 if self.p3 is not None else np.ones(nt1), np.float64, C.c_double)
        sp = self.spline if self.spline is not None else (one,) * 5
        s.nspline = len(sp[0]) if self.spline is not None else 0
        s.xs, s.ys, s.bs, s.cs, s.ds = (arr(a, np.float64, C.c_double) for a in sp)
        if self.rf is not None:
            gs = np.ascontiguousarray(self.rf.data, dtype=np.float32)
            gr = np.ascontiguousarray(self.rf_greff, dtype=np.float32)
            assert gs.shape == gr.shape and gs.ndim == 4
            keep.extend([gs, gr])
            s.rf.nclx, s.rf.ncly, s.rf.nclz = gs.shape[1:]
            s.rf.dcel, s.rf.xbcen, s.rf.ybcen, s.rf.zbcen = (self.rf.dcel, self.rf.xbcen,
                                                             self.rf.ybcen, self.rf.zbcen)
            s.gsrfen = gs.ctypes.data_as(C.POINTER(C.c_float))
            s.greff = gr.ctypes.data_as(C.POINTER(C.c_float))
        else:
            z = np.zeros(1, np.float32)
            keep.append(z)
            s.gsrfen = s.greff = z.ctypes.data_as(C.POINTER(C.c_float))
        s.sqrfac = self.sqrfac
        if self.apfor is not None:
            sn, af = self.apfor
            s.afn = len(sn)
            s.sn = arr(sn if len(sn) else np.zeros(1), np.int32, C.c_int32)
            s.af = arr(np.asarray(af, float).reshape(-1) if len(sn) else np.zeros(3), np.float64,
                       C.c_double)
        ct = self.contrans
        if ct is not None:
            n = len(ct["csn"])
            s.ctn = n
            s.csn = arr(ct["csn"] if n else np.zeros(1), np.int32, C.c_int32)
            k = np.asarray(ct["k"], float).reshape(n, 3) if n else np.zeros((1, 3))
            c0 = np.asarray(ct["centre"], float).reshape(n + 1, 3)
            s.kx, s.ky, s.kz = (arr(k[:, a], np.float64, C.c_double) for a in range(3))
            s.contrx, s.contry, s.contrz = (arr(c0[:, a], np.float64, C.c_double) for a in range(3))
            s.unsplit = 1 if ct.get("unsplit") else 0
            s.nsites1st = int(ct.get("nsites1st", 0))
        if self.notrans is not None:
            s.notrans_axes = int(self.notrans[0])
            s.notrx, s.notry, s.notrz = (float(v) for v in self.notrans[1])
        s.ibeta, s.diff0, s.diffcutoff = self.ibeta, self.diff0, self.diffcutoff
        s.thold27 = self.thold27
        m = self.membrane
        if m is not None:
            s.qpore = 1 if m.get("qpore") else 0
            s.voltage, s.afact, s.ceps = m["voltage"], m["afact"], m["ceps"]
            s.zmemb1, s.zmemb2 = m["zmemb1"], m["zmemb2"]
            s.ampl1 = arr(m["ampl1"], np.float64, C.c_double)
            s.ampl2 = arr(m["ampl2"], np.float64, C.c_double)
            s.rcylinder = arr(m["rcylinder"], np.float64, C.c_double)
            s.p1 = arr(np.stack([m["p1"], 1.0 / np.asarray(m["p1"], float)], 1).reshape(-1),
                       np.float64, C.c_double)
            s.p2 = arr(np.stack([m["p2"], 1.0 / np.asarray(m["p2"], float)], 1).reshape(-1),
                       np.float64, C.c_double)
        g = self.mmij
        if g is not None:
            s.mmij_shape = int(g.get("shape", 1))
            s.xnpol, s.ynpol, s.znpol = (int(v) for v in g["npol"])
            lst = np.asarray(g["lstp"], np.int32).reshape(-1, 3)
            s.ntpol = len(lst)
            s.lstpx, s.lstpy, s.lstpz = (arr(lst[:, a], np.int32, C.c_int32) for a in range(3))
            s.bnorm = arr(g["bnorm"], np.float64, C.c_double)
            s.mmij = arr(np.asarray(g["mmij"], float).reshape(-1), np.float64, C.c_double)
            s.xscal, s.yscal, s.zscal = (float(v) for v in g["scal"])
            s.rfscal = float(g.get("rfscal", 1.0))
        s.cx, s.cy, s.cz = (float(v) for v in self.center)
        s.rsphe, s.rsphe2 = float(self.rsphe), float(self.rsphe) ** 2
        s.iecx, s.iecy = float(self.iecx), float(self.iecy)
        return s, keep


def nindex_table(nttyp, ndna, nion):
    """nindex.f90:19-35 over atnam2 types 1..nttyp, as a dense (nttyp+1)^2 table."""
    t = np.zeros((nttyp + 1, nttyp + 1), np.int32)
    for a in range(1, nttyp + 1):
        for b in range(1, nttyp + 1):
            mx, mn = (a, b) if a > b else (b, a)
            maxi = mx - ndna
            if mn > ndna:
                t[a, b] = maxi * (maxi - 1) // 2 + mn - ndna + ndna * nion
            elif mx > ndna:
                t[a, b] = (mn - 1) * nion + maxi
    return t.reshape(-1)


@dataclass
class Gcmc:
    """GCMC buffers (grandmod) in ABI form; arrays are 1-based (entry 0 unused)."""
    ibfftyp: np.ndarray
    avnum: np.ndarray
    mu: np.ndarray
    kb: np.ndarray
    lzmin: np.ndarray
    lzmax: np.ndarray
    bufferbias: np.ndarray
    ntotat: np.ndarray
    lx: float
    ly: float
    cx: float
    cy: float
    cz: float
    ikbt: float
    rmin: np.ndarray = None          # [nbuffer+1] spherical shells (QSPHERE)
    rmax: np.ndarray = None

    @property
    def nbuffer(self):
        return len(self.ibfftyp) - 1

    def to_struct(self):
        keep = []

        def arr(a, dtype, ctype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        g = abi.GcmcStruct()
        g.nbuffer = self.nbuffer
        g.ibfftyp = arr(self.ibfftyp, np.int32, C.c_int32)
        g.avnum = arr(self.avnum, np.float64, C.c_double)
        g.mu = arr(self.mu, np.float64, C.c_double)
        g.kb = arr(self.kb, np.float64, C.c_double)
        g.lzmin = arr(self.lzmin, np.float64, C.c_double)
        g.lzmax = arr(self.lzmax, np.float64, C.c_double)
        g.bufferbias = arr(self.bufferbias, np.int8, C.c_int8)
        g.ntotat = arr(self.ntotat, np.int32, C.c_int32)
        g.lx, g.ly, g.cx, g.cy, g.cz, g.ikbt = (self.lx, self.ly, self.cx, self.cy, self.cz,
                                                self.ikbt)
        if self.rmin is not None:
            g.rmin = arr(self.rmin, np.float64, C.c_double)
            g.rmax = arr(self.rmax, np.float64, C.c_double)
        return g, keep


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


__all__ = ["System", "GridData", "Gcmc", "nindex_table", "CELEC", "kbt_of", "ptr", "math"]


@dataclass
class Replica:
    """One system of a replica batch: its particles, its static field (on the shared phix
    geometry) and its GCMC buffers."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    typei: np.ndarray
    nfix: int = 0
    capacity: int = 0
    phix: np.ndarray = None          # float32 (nclx, ncly, nclz)
    gcmc: Gcmc | None = None
    ibuffer: np.ndarray = None       # [ntot]
    nat: np.ndarray = None           # [nbuffer+1]
    meta: dict = field(default_factory=dict)

    @property
    def ntot(self):
        return len(self.typei)

    def to_struct(self, keep):
        def arr(a, dtype, ctype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        r = abi.ReplicaStruct()
        r.ntot, r.nfix = self.ntot, self.nfix
        r.capacity = max(self.capacity, self.ntot, 1)
        r.x = arr(self.x, np.float64, C.c_double)
        r.y = arr(self.y, np.float64, C.c_double)
        r.z = arr(self.z, np.float64, C.c_double)
        r.typei = arr(self.typei, np.int32, C.c_int32)
        if self.phix is not None:
            r.phix = arr(self.phix, np.float32, C.c_float)
        if self.gcmc is not None:
            g, gk = self.gcmc.to_struct()
            keep.extend([g, gk])
            r.gcmc = C.pointer(g)
            r.ibuffer = arr(self.ibuffer[:self.ntot], np.int32, C.c_int32)
            r.nat = arr(self.nat, np.int32, C.c_int32)
        return r


def replica_system(shared: System, rep: Replica) -> System:
    """The single system that replica `rep` of a batch over `shared` is."""
    s = shared.copy()
    n = rep.ntot
    s.ntot, s.nsites, s.nfix = n, 0, rep.nfix
    s.capacity = max(rep.capacity, n, 1)
    s.x, s.y, s.z = (np.asarray(a, float).copy() for a in (rep.x, rep.y, rep.z))
    s.typei = np.asarray(rep.typei, np.int32).copy()
    if rep.phix is not None and s.phix is not None:
        s.phix = GridData(np.asarray(rep.phix, np.float32), s.phix.dcel, s.phix.xbcen,
                          s.phix.ybcen, s.phix.zbcen)
    return s

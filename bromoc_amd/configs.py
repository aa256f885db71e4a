"""Synthetic, seeded BROMOC workloads (no reference input decks are shipped).

plain_ions     400-ion BD in a 60 A box, 3 EFPOT pair tables on r = 1.5..20 A at 0.01 A
pore_dna       ions in a cylindrical pore through a slab (phiv/phix grids) with a 50-bp
               CG-DNA duplex on the pore axis (3 sites per nucleotide)
pore_replica   one replica of the batched ion+pore workload (no DNA, GCMC buffers)

Every builder is size-parameterised so tests can build small instances of the same shape.
The EFPOT (r,U) samples are turned into the reference's piecewise-quadratic table by the
host-side SPLINEPOT of the HIP library (bromoc_build_table), or by a caller-given builder.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import abi
from .system import (CELEC, COULOMB, KCALMOL, AVOGADRO, LITER, ANGSTROM, PI, Gcmc, GridData,
                     System, kbt_of, nindex_table)

RMIN, RMAX, RES = 1.5, 20.0, 0.01


def lib_table_builder(n, r, u, chr_, fct):
    lib = abi.load_library()
    ep = np.zeros(3 * n)
    p = abi.Pair()
    rc = lib.bromoc_build_table(n, r.ctypes.data_as(C.POINTER(C.c_double)),
                                u.ctypes.data_as(C.POINTER(C.c_double)), chr_, fct,
                                ep.ctypes.data_as(C.POINTER(C.c_double)), C.byref(p))
    if rc:
        raise abi.BromocError("bromoc_build_table", rc)
    return ep, p


def _pair_dict(p: abi.Pair):
    return {name: getattr(p, name) for name, _ in abi.Pair._fields_}


def knot_grid(rmin=RMIN, rmax=RMAX, res=RES):
    """r samples as an EFPOT file lists them (decimal values, nearest binary64)."""
    n = int(round((rmax - rmin) / res)) + 1
    return np.array([float(f"{rmin + res * k:.6f}") for k in range(n)])


def efpot_curve(r, sig, eps, qq, cdie, osc=0.3):
    """12-6 LJ + screened Coulomb + damped oscillatory tail (kcal/mol)."""
    sr6 = (sig / r) ** 6
    return (4.0 * eps * (sr6 * sr6 - sr6) + qq * CELEC / (cdie * r)
            + osc * np.exp(-r / 3.0) * np.cos(2.0 * PI * r / 2.8))


def build_tables(pairs_spec, table_builder=None, rmin=RMIN, rmax=RMAX, res=RES):
    """pairs_spec: list indexed by pair type (entry 0 None) of dicts with sig, eps, qq,
    cdie.  Returns (pairs list of dicts, ep)."""
    tb = table_builder or lib_table_builder
    r = knot_grid(rmin, rmax, res)
    pairs, eps_all = [None], []
    off = 0
    for spec in pairs_spec[1:]:
        u = efpot_curve(r, spec["sig"], spec["eps"], spec["qq"], spec["cdie"],
                        spec.get("osc", 0.3))
        chr_ = 1 if spec["qq"] != 0.0 else 0
        fct = CELEC * spec["qq"] / spec["cdie"] if chr_ else 0.0
        ep, p = tb(len(r), r, u, chr_, fct)
        d = _pair_dict(p)
        d["nxi"] = off
        d["efpot"] = 1
        off += len(r)
        pairs.append(d)
        eps_all.append(ep)
    return pairs, np.concatenate(eps_all) if eps_all else np.zeros(0)


def _place(rng, n, lo, hi, dmin, accept=None, existing=None, max_tries=2_000_000):
    """Uniform positions in the box [lo,hi], rejection-sampled to min separation dmin."""
    pts = np.zeros((0, 3)) if existing is None else np.asarray(existing, float).reshape(-1, 3)
    out = []
    cell = max(dmin, 1e-9)
    grid = {}

    def key(p):
        return tuple(np.floor(p / cell).astype(int))

    for p in pts:
        grid.setdefault(key(p), []).append(p)
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place particles")
        p = lo + (hi - lo) * rng.random(3)
        if accept is not None and not accept(p):
            continue
        k = key(p)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if np.sum((p - q) ** 2) < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(k, []).append(p)
            out.append(p)
    return np.array(out).reshape(-1, 3)


# ------------------------------------------------------------------ plain ions
def plain_ions(seed=0, n_per_type=200, box=60.0, temp=300.0, cdie=80.0, dt=0.02,
               diff=(0.196, 0.203), charges=(1.0, -1.0), table_builder=None, capacity=None,
               extra_pair_types=0, rmax=RMAX):
    """Plain ion-only BD: equal numbers of K+ and Cl- in a cubic box centred at 0."""
    rng = np.random.default_rng(seed)
    ntypes = len(charges)
    nion = ntypes
    ndna = 0
    nttyp = nion
    sig = rng.uniform(3.0, 4.5, ntypes)
    eps = rng.uniform(0.1, 0.5, ntypes)
    pidx = nindex_table(nttyp, ndna, nion)
    npair = nion * (nion + 1) // 2
    spec = [None] * (npair + 1)
    for a in range(1, nttyp + 1):
        for b in range(a, nttyp + 1):
            ia, ib = a - 1, b - 1
            spec[pidx[a * (nttyp + 1) + b]] = dict(
                sig=0.5 * (sig[ia] + sig[ib]), eps=math.sqrt(eps[ia] * eps[ib]),
                qq=charges[ia] * charges[ib], cdie=cdie)
    pairs, ep = build_tables(spec, table_builder, rmax=rmax)
    half = 0.5 * box
    pos = _place(rng, n_per_type * ntypes, -half, half, 2.0)
    typei = np.repeat(np.arange(1, ntypes + 1, dtype=np.int32), n_per_type)
    order = rng.permutation(len(typei))
    pos, typei = pos[order], typei[order]
    kbt = kbt_of(temp)
    kbtdt = dt * (1.0 / kbt)
    d = np.array((0.0,) + tuple(diff))
    s = System(
        ntot=len(typei), nsites=0, nfix=0, ntype=ntypes, nold=0, nttyp=nttyp,
        flags=abi.QPAR | abi.QNONBOND | abi.QNOBOND,
        x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), typei=typei,
        nwtype=np.arange(0, ntypes + 1, dtype=np.int32),
        cg=np.array((0.0,) + tuple(charges)), fact1a=d * kbtdt, fact2a=np.sqrt(2.0 * dt * d),
        pairidx=pidx, pairs=pairs, ep=ep, res=RES,
        capacity=capacity or len(typei), bdmax=1.0,
        box=(-half, half, -half, half, -half, half), ndna=ndna, nion=nion,
        dt=dt, kbtdt=kbtdt, diffusion=d.copy(),
        meta=dict(workload="plain_ions", seed=seed, temp=temp, dt=dt, kbt=kbt, cdie=cdie),
    )
    s.fact2a[0] = 0.0
    return s


# ------------------------------------------------------------------- CG-DNA
# B-form cylindrical coordinates (r [A], phi [deg], z [A]) of the 3-site model,
# rise 3.38 A and twist 36 deg per base step.
_CYL = {"P": (8.91, 94.9, 2.186), "S": (6.19, 70.5, 1.89), "A": (0.773, 41.905, 0.981),
        "T": (2.349, 86.452, 1.281), "C": (1.837, 85.539, 1.347), "G": (0.817, 40.808, 0.930)}
_COMP = {"A": "T", "T": "A", "C": "G", "G": "C"}
_BASE_CODE = {"A": abi.SITE_AB, "T": abi.SITE_TB, "C": abi.SITE_CB, "G": abi.SITE_GB}


def _dihedral(p1, p2, p3, p4):
    """energy.mono.f90:283-300 convention."""
    v1, v2, v3 = p1 - p2, p3 - p2, p3 - p4
    n1, n2 = np.cross(v1, v2), np.cross(v2, v3)
    return math.atan2(math.sqrt(v2 @ v2) * (v1 @ n2), n1 @ n2)


def _angle(p1, p2, p3):
    v1, v2 = p1 - p2, p3 - p2
    return math.acos((v1 @ v2) / math.sqrt((v1 @ v1) * (v2 @ v2)))


def dna_duplex(seq, z0=0.0, solv=True, layout=None):
    """Native B-form duplex of `seq` (strand 1, 5'->3') plus its complement: site arrays
    and the bonded / nonbonded lists built by the rules of bonds.f90, angles.f90,
    dihedral.f90 and go_qq.f90:45-116 on the native geometry.  layout(pos, step) may move
    the sites (e.g. fold a long duplex) before the native values are taken; step is each
    site's base-pair step along the helix."""
    nb = len(seq)
    rise, twist = 3.38, 36.0
    names, code, strand, pos, nuc = [], [], [], [], []

    def site(kind, step, st, letter, k):
        r, ph, zz = _CYL[kind]
        if st == 2:
            ph, zz = -ph, -zz
        a = math.radians(ph + step * twist)
        pos.append((r * math.cos(a), r * math.sin(a), zz + step * rise))
        names.append(kind)
        code.append(abi.SITE_P if kind == "P" else abi.SITE_S if kind == "S" else _BASE_CODE[kind])
        strand.append(st)
        nuc.append(k)

    k = 0
    for j, b in enumerate(seq):
        for kind in ("P", "S", b):
            site(kind, j, 1, b, k)
        k += 1
    comp = [_COMP[b] for b in reversed(seq)]
    for j, b in enumerate(comp):
        step = nb - 1 - j
        for kind in ("P", "S", b):
            site(kind, step, 2, b, k)
        k += 1
    pos = np.array(pos)
    pos[:, 2] += z0 - 0.5 * (nb - 1) * rise
    if layout is not None:
        steps = np.array([j for j in range(nb) for _ in range(3)] +
                         [nb - 1 - j for j in range(nb) for _ in range(3)])
        pos = layout(pos, steps)
    n = len(names)
    bonds = []
    for st in (1, 2):
        idx = [i for i in range(n) if strand[i] == st]
        for t in range(0, len(idx), 3):
            p, s_, b = idx[t], idx[t + 1], idx[t + 2]
            bonds += [(p, s_), (s_, b)]
            if t + 3 < len(idx):
                bonds.append((s_, idx[t + 3]))
    nbr = {i: set() for i in range(n)}
    for a, b in bonds:
        nbr[a].add(b)
        nbr[b].add(a)
    angles = []
    for c_ in range(n):
        ns_ = sorted(nbr[c_])
        for ii in range(len(ns_)):
            for jj in range(ii + 1, len(ns_)):
                angles.append((ns_[ii], c_, ns_[jj]))
    dihes = []
    for b_, c_ in bonds:
        for a in sorted(nbr[b_] - {c_}):
            for d in sorted(nbr[c_] - {b_}):
                if a != d:
                    dihes.append((a, b_, c_, d))
    # nonbonded lists over every site pair i > j (go_qq.f90:45-116), in (i, j) order
    nm = ["P " if x == "P" else "S " if x == "S" else x + "b" for x in names]
    code_a = np.array([{"P ": 0, "S ": 1, "Ab": 2, "Tb": 3, "Cb": 4, "Gb": 5}[x] for x in nm])
    strand_a = np.array(strand)
    isbase = code_a >= 2
    comp_of = np.array([-1, -1, 3, 2, 5, 4])       # A<->T, C<->G
    bonded = [set() for _ in range(n)]
    for a, b in bonds:
        bonded[max(a, b)].add(min(a, b))
    angend = [set() for _ in range(n)]
    for a, _, d in angles:
        angend[max(a, d)].add(min(a, d))
    parts = {k: [] for k in ("stack", "bp", "ex", "qq", "slv")}
    for i in range(1, n):
        j = np.arange(i)
        keep = np.ones(i, bool)
        if bonded[i]:
            keep[list(bonded[i])] = False
        dij = np.sqrt(np.sum((pos[:i] - pos[i]) ** 2, axis=1))
        same = strand_a[:i] == strand_a[i]
        st_ = keep & same & ((dij - 9.0) < 0.0) & (np.abs(dij - 9.0) >= 1e-6)
        bp_ = keep & ~st_ & (comp_of[code_a[i]] == code_a[:i])
        ex_ = keep & ~st_ & ~bp_
        if st_.any():
            parts["stack"].append((np.full(st_.sum(), i), j[st_], dij[st_]))
        if bp_.any():
            sg = 2.8694 if nm[i] in ("Gb", "Cb") else 2.9002
            parts["bp"].append((np.full(bp_.sum(), i), j[bp_], np.full(bp_.sum(), sg)))
        if ex_.any():
            sg = np.where(isbase[i] & isbase[:i][ex_], 1.0, 6.86)
            parts["ex"].append((np.full(ex_.sum(), i), j[ex_], sg))
        qk = keep & (code_a[i] == 0) & (code_a[:i] == 0)
        if angend[i]:
            qk[list(angend[i])] = False
        if qk.any():
            parts["qq"].append((np.full(qk.sum(), i), j[qk], None))
        if solv and code_a[i] == 1:
            sk = (~same) & (code_a[:i] == 1)
            if sk.any():
                parts["slv"].append((np.full(sk.sum(), i), j[sk], None))

    def cat(name):
        ps = parts[name]
        if not ps:
            return np.zeros(0, np.int32), np.zeros(0)
        a = np.concatenate([q[0] for q in ps]).astype(np.int32)
        b = np.concatenate([q[1] for q in ps]).astype(np.int32)
        v = np.concatenate([q[2] for q in ps]) if ps[0][2] is not None else None
        return np.concatenate([a, b]), v

    def cm(lst, k):
        if not lst:
            return np.zeros(0, np.int32)
        a = np.array([t[:k] for t in lst], np.int32)
        return a.T.reshape(-1).copy()

    lists = {
        "bond": (cm(bonds, 2), np.array([np.linalg.norm(pos[a] - pos[b]) for a, b in bonds])),
        "angle": (cm(angles, 3), np.array([_angle(pos[a], pos[b], pos[c]) for a, b, c in angles])),
        "dihe": (cm(dihes, 4), np.array([_dihedral(pos[a], pos[b], pos[c], pos[d])
                                         for a, b, c, d in dihes])),
        "stack": cat("stack"),
        "bp": cat("bp"),
        "ex": cat("ex"),
        "qq": (cat("qq")[0], None),
        "slv": (cat("slv")[0], None),
    }
    return dict(pos=pos, code=np.array(code, np.int32), strand=np.array(strand), nuc=nuc,
                lists=lists, nnuc=k)


# ----------------------------------------------------------------- pore grids
def pore_repulsion(shape, dcel, rpore, half_slab):
    """int8 map: 1 inside the slab wall (|z| <= half_slab and R > rpore), else 0
    (the binarised phiv of readphi.f90:319-327)."""
    nx, ny, nz = shape
    x = (np.arange(nx) - 0.5 * (nx - 1)) * dcel
    y = (np.arange(ny) - 0.5 * (ny - 1)) * dcel
    z = (np.arange(nz) - 0.5 * (nz - 1)) * dcel
    r2 = x[:, None] ** 2 + y[None, :] ** 2
    wall_xy = r2 > rpore * rpore
    wall_z = np.abs(z) <= half_slab
    return (wall_xy[:, :, None] & wall_z[None, None, :]).astype(np.int8)


def pore_field(shape, dcel, volt, half_slab, rng, noise_volt=0.01, smooth=2.0):
    """float32 static field phix [e/A]: a linear drop of `volt` volts across the slab
    plus Gaussian-smoothed noise, scaled so that E = q*phix*celec (staticf1.f90:110)."""
    from scipy.ndimage import gaussian_filter
    nx, ny, nz = shape
    z = (np.arange(nz) - 0.5 * (nz - 1)) * dcel
    v = volt * np.clip((half_slab - z) / (2.0 * half_slab), 0.0, 1.0)
    noise = gaussian_filter(rng.standard_normal(shape).astype(np.float32), smooth)
    noise *= noise_volt / max(float(noise.std()), 1e-30)
    vol = v[None, None, :] + noise
    return (vol * (COULOMB / KCALMOL) / CELEC).astype(np.float32)


def pore_dna(seed=1, n_bp=50, n_ions=2000, box=(100.0, 100.0, 200.0), dcel=0.5, rpore=12.0,
             half_slab=20.0, volt=0.5, svdw=1000.0, temp=300.0, cdie=80.0, dt=0.02,
             table_builder=None, dna=True, capacity=None, zcont=(-20.0, 0.0, 20.0),
             rmax=RMAX, grid_shape=None, dna_layout=None, workload=None):
    """Ions in a cylindrical pore through a slab, with a CG-DNA duplex on the axis (or laid
    out by dna_layout, see dna_duplex).  The grids cover the box unless grid_shape is
    given (centred on the box)."""
    rng = np.random.default_rng(seed)
    lx, ly, lz = box
    shape = grid_shape or tuple(int(round(L / dcel)) + 1 for L in box)
    phiv = pore_repulsion(shape, dcel, rpore, half_slab)
    phix = pore_field(shape, dcel, volt, half_slab, rng)
    if dna:
        seq = "".join(rng.choice(list("ACGT"), n_bp))
        for b in "ACGT":                      # all four base types present
            if b not in seq:
                seq = b + seq[1:]
        d = dna_duplex(seq, z0=0.0, layout=dna_layout)
        spos = d["pos"]
        ns = len(spos)
        nold = d["nnuc"]
        present = [c for c in (abi.SITE_AB, abi.SITE_TB, abi.SITE_CB, abi.SITE_GB)
                   if c in set(d["code"].tolist())]
        ndna = 2 + len(present)
        t2 = {abi.SITE_S: 1, abi.SITE_P: 2}
        for k, c in enumerate(present):
            t2[c] = 3 + k
        typtyp = np.array([t2[c] for c in d["code"]], np.int32)
    else:
        spos = np.zeros((0, 3))
        ns = 0
        nold = 0
        ndna = 0
        typtyp = np.zeros(0, np.int32)
        d = None
    nion = 2
    ntype = nold + nion
    nttyp = ndna + nion
    charges = (1.0, -1.0)
    diff = (0.196, 0.203)
    isig = rng.uniform(3.0, 4.5, nion)
    ieps = rng.uniform(0.1, 0.5, nion)
    ssig = rng.uniform(3.0, 4.5, max(ndna, 1))
    seps = rng.uniform(0.1, 0.5, max(ndna, 1))
    pidx = nindex_table(nttyp, ndna, nion)
    npair = ndna * nion + nion * (nion + 1) // 2
    spec = [None] * (npair + 1)
    cgnuc = -1.0
    for a in range(1, nttyp + 1):
        for b in range(a, nttyp + 1):
            if b <= ndna:
                continue
            if a > ndna:
                ia, ib = a - ndna - 1, b - ndna - 1
                spec[pidx[a * (nttyp + 1) + b]] = dict(
                    sig=0.5 * (isig[ia] + isig[ib]), eps=math.sqrt(ieps[ia] * ieps[ib]),
                    qq=charges[ia] * charges[ib], cdie=cdie)
            else:
                ib = b - ndna - 1
                qs = cgnuc if a == 2 else 0.0
                spec[pidx[a * (nttyp + 1) + b]] = dict(
                    sig=0.5 * (ssig[a - 1] + isig[ib]), eps=math.sqrt(seps[a - 1] * ieps[ib]),
                    qq=qs * charges[ib], cdie=cdie)
    pairs, ep = build_tables(spec, table_builder, rmax=rmax)

    hx, hy, hz = 0.5 * lx, 0.5 * ly, 0.5 * lz
    sgrid = {}
    for p in spos:
        sgrid.setdefault(tuple(np.floor(p / 4.0).astype(int)), []).append(p)

    def accessible(p):
        if abs(p[2]) <= half_slab + 1.0 and p[0] ** 2 + p[1] ** 2 > (rpore - 1.5) ** 2:
            return False
        k = np.floor(p / 4.0).astype(int)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in sgrid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if np.sum((p - q) ** 2) < 3.5 * 3.5:
                            return False
        return True

    lo = np.array([-hx + 0.5, -hy + 0.5, -hz + 0.5])
    hi = -lo
    ipos = _place(rng, n_ions, lo, hi, 2.0, accept=accessible)
    ityp = np.array([nold + 1 + (k % 2) for k in range(n_ions)], np.int32)
    pos = np.concatenate([spos, ipos]) if ns else ipos
    typei = np.concatenate([np.zeros(ns, np.int32), ityp])
    ntot = ns + n_ions
    kbt = kbt_of(temp)
    kbtdt = dt * (1.0 / kbt)
    dvec = np.zeros(ntype + 1)
    dvec[nold + 1:] = diff
    cg = np.zeros(ntype + 1)
    cg[1:nold + 1] = cgnuc
    cg[nold + 1:] = charges
    nwtype = np.zeros(ntype + 1, np.int32)
    for k in range(nion):
        nwtype[nold + 1 + k] = ndna + 1 + k
    flags = abi.QPAR | abi.QNONBOND | abi.QNOBOND | abi.QPHIX | abi.QPHIV | abi.QCOUNTION
    diffnuc = 0.1
    kw = {}
    if dna:
        flags |= abi.QNUCL | abi.QSOLV
        kw = dict(typtyp=typtyp, namsite=d["code"], stfree=np.ones(ns, np.int8),
                  lists=d["lists"], epsnuc=rng.uniform(0.15, 0.25), scalepairing=1.0,
                  fctn=CELEC * cgnuc ** 2 / cdie, epsolv=0.1)
    tz = 0.5 * (shape[2] - 1) * dcel
    s = System(
        ntot=ntot, nsites=ns, nfix=0, ntype=ntype, nold=nold, nttyp=nttyp, flags=flags,
        x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), typei=typei,
        nwtype=nwtype, cg=cg, fact1a=dvec * kbtdt, fact2a=np.sqrt(2.0 * dt * dvec),
        pairidx=pidx, pairs=pairs, ep=ep, res=RES, capacity=capacity or ntot,
        phix=GridData(phix, dcel), phiv=GridData(phiv, dcel), svdw=svdw, thold8=8.0,
        vzmin=-tz, vzmax=tz, bdmax=1.0, cgnuc=cgnuc,
        fact0n1=diffnuc * kbt * dt, fact0n2=math.sqrt(2.0 * dt * diffnuc),
        box=(-hx, hx, -hy, hy, -hz, hz), zcont=np.array(zcont, float), ndna=ndna, nion=nion,
        dt=dt, kbtdt=kbtdt, diffusion=dvec.copy(),
        meta=dict(workload=workload or ("pore_dna" if dna else "pore_ions"), seed=seed,
                  temp=temp, dt=dt,
                  kbt=kbt, cdie=cdie, n_bp=n_bp if dna else 0, grid=shape, dcel=dcel),
        **kw,
    )
    return s


def serpentine(seg_bp=80, thread_base=-100.0, cis_base=-295.0, columns=None, rise=3.38):
    """dna_duplex layout: a long duplex cut into segments of seg_bp base-pair steps, each
    moved rigidly: the first threads the pore on the z axis from z = thread_base, the others
    stand side by side below the slab (cis side) from z = cis_base, at the given (x, y)."""
    cols = columns or [(40.0, 0.0), (-40.0, 0.0), (0.0, 40.0), (0.0, -40.0), (40.0, 40.0),
                       (-40.0, 40.0), (40.0, -40.0), (-40.0, -40.0), (80.0, 0.0),
                       (-80.0, 0.0), (0.0, 80.0), (0.0, -80.0), (80.0, 80.0),
                       (-80.0, -80.0)]

    def layout(pos, step):
        out = pos.copy()
        zref = float(np.min(pos[:, 2] - step * rise))    # the helix start
        seg = step // seg_bp
        if seg.max() > len(cols):
            raise ValueError("too many segments for the columns")
        for k in range(seg.max() + 1):
            m = seg == k
            base = thread_base if k == 0 else cis_base
            cx, cy = (0.0, 0.0) if k == 0 else cols[k - 1]
            out[m, 0] = pos[m, 0] + cx
            out[m, 1] = pos[m, 1] + cy
            out[m, 2] = pos[m, 2] + (base - zref - k * seg_bp * rise)
        return out
    return layout


def translocation(seed=4, n_bp=1000, n_ions=50000, box=(200.0, 200.0, 600.0), dcel=0.625,
                  grid_shape=(321, 321, 641), rpore=15.0, half_slab=20.0, volt=0.5,
                  table_builder=None, conc=1.0, rfpar=True, profile=True, gcmc=True,
                  capacity=None):
    """The large translocation workload: a 1000-bp CG-DNA (6000 sites) threaded through a
    pore of radius 15 A in a 200 x 200 x 600 A box (the rest of the duplex folded on the cis
    side), 50,000 K+/Cl- ions, repulsion / static-field / reaction-field grids of
    321 x 321 x 641 at 0.625 A, D(z) tanh-reduced to 0.5x inside the pore (PROFILE), and
    10 A GCMC buffers at both box ends."""
    s = pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=box, dcel=dcel, rpore=rpore,
                 half_slab=half_slab, volt=volt, table_builder=table_builder,
                 capacity=capacity or (n_ions + 6 * n_bp + 4096), grid_shape=grid_shape,
                 dna_layout=serpentine(), workload="translocation")
    if rfpar:
        s = with_rfpar(s, seed=seed, shape=grid_shape, dcel=dcel, wall_radius=None)
    if profile:
        zk, fk = tanh_profile(box[2] * -0.5, box[2] * 0.5, 121, half_slab)
        s = with_profile_diffusion(s, zk, fk)
    if gcmc:
        s.meta["gcmc"] = gcmc_buffers(s, conc=conc, buffer_len=10.0)
    return s


def gcmc_buffers(s: System, conc, buffer_len=10.0, mu=(0.0, 0.0), volt=0.0):
    """One buffer per ion type at each box end (2 * nion buffers), as the BUFFER command
    sets them up (shell_simul.f90:1015-1037): avnum = conc * volume, mu shifted by the
    applied voltage."""
    lx2m, lx2p, ly2m, ly2p, lz2m, lz2p = s.box
    lx, ly = lx2p - lx2m, ly2p - ly2m
    ib_types, lzmin, lzmax, avnum, mus = [0], [0.0], [0.0], [0.0], [0.0]
    for end in (0, 1):
        for k in range(s.nion):
            t = s.nold + 1 + k
            lo, hi = (lz2m, lz2m + buffer_len) if end == 0 else (lz2p - buffer_len, lz2p)
            vol = lx * ly * (hi - lo)
            dens = conc * (AVOGADRO / LITER) * ANGSTROM ** 3
            ib_types.append(t)
            lzmin.append(lo)
            lzmax.append(hi)
            avnum.append(dens * vol)
            battery = volt if end == 1 else 0.0
            mus.append(mu[k] + s.cg[t] * battery * COULOMB / KCALMOL)
    nb = len(ib_types)
    return Gcmc(ibfftyp=np.array(ib_types, np.int32), avnum=np.array(avnum), mu=np.array(mus),
                kb=np.zeros(nb), lzmin=np.array(lzmin), lzmax=np.array(lzmax),
                bufferbias=np.zeros(nb, np.int8),
                ntotat=np.array([0] + [int(round(a)) for a in avnum[1:]], np.int32),
                lx=lx, ly=ly, cx=0.0, cy=0.0, cz=0.0, ikbt=1.0 / s.meta["kbt"])


def assign_buffers(s: System, g: Gcmc):
    """ibuffer of each particle: the buffer whose slab holds it and whose type matches
    (0 if none), as the reference's initial count does (count.f90:39-92)."""
    ib = np.zeros(s.capacity, np.int32)
    sphere = bool(s.flags & abi.QSPHERE)
    cx, cy, cz = s.center
    for i in range(s.nsites + s.nfix, s.ntot):
        t = abs(int(s.typei[i]))
        for b in range(1, g.nbuffer + 1):
            if g.ibfftyp[b] == t and g.lzmin[b] <= s.z[i] <= g.lzmax[b]:
                if sphere:      # the shell rmin..rmax too (count.f90:52-55)
                    r2 = (s.x[i] - cx) ** 2 + (s.y[i] - cy) ** 2 + (s.z[i] - cz) ** 2
                    if not (g.rmin[b] ** 2 <= r2 <= g.rmax[b] ** 2):
                        continue
                ib[i] = b
                break
    nat = np.zeros(g.nbuffer + 1, np.int32)
    for b in range(1, g.nbuffer + 1):
        nat[b] = int(np.sum(ib[:s.ntot] == b))
    return ib, nat


def with_static_field(s: System, seed=0, dcel=1.0, volt=0.5, noise_volt=0.05):
    """A static field phix (STATICF1) over the whole box: a linear drop of `volt` volts
    along z plus smoothed noise (pore_field), on a grid of spacing dcel centred on the box."""
    s = s.copy()
    lx2m, lx2p, ly2m, ly2p, lz2m, lz2p = s.box
    shape = tuple(int(math.ceil((hi - lo) / dcel)) + 3
                  for lo, hi in ((lx2m, lx2p), (ly2m, ly2p), (lz2m, lz2p)))
    half = 0.5 * (lz2p - lz2m)
    field = pore_field(shape, dcel, volt, half, np.random.default_rng(seed), noise_volt=noise_volt)
    s.phix = GridData(field, dcel, xbcen=0.5 * (lx2m + lx2p), ybcen=0.5 * (ly2m + ly2p),
                      zbcen=0.5 * (lz2m + lz2p))
    s.flags |= abi.QPHIX
    return s


def with_soft_repulsion(s: System, seed=0, dcel=1.0, svdw=0.5, vmax=3, scal=True):
    """A repulsion map phiv (VDWGD1TRLN) over the whole box with small smooth integer
    values 0..vmax (from smoothed noise), scaled by svdw, with per-type scal factors (QSVDW)
    when scal is set; vzmin..vzmax is the box's z range."""
    from scipy.ndimage import gaussian_filter
    s = s.copy()
    lx2m, lx2p, ly2m, ly2p, lz2m, lz2p = s.box
    shape = tuple(int(math.ceil((hi - lo) / dcel)) + 3
                  for lo, hi in ((lx2m, lx2p), (ly2m, ly2p), (lz2m, lz2p)))
    v = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), 2.0)
    v = (v - v.min()) / max(float(v.max() - v.min()), 1e-30)
    s.phiv = GridData(np.rint(v * vmax).astype(np.int8), dcel, xbcen=0.5 * (lx2m + lx2p),
                      ybcen=0.5 * (ly2m + ly2p), zbcen=0.5 * (lz2m + lz2p))
    s.svdw, s.thold8, s.vzmin, s.vzmax = svdw, 8.0, lz2m, lz2p
    s.flags |= abi.QPHIV
    if scal:
        s.scal = np.linspace(0.5, 1.5, s.nttyp + 1)
        s.flags |= abi.QSVDW
    return s


# ------------------------------------------------------------------ curved systems
def with_sphere(s: System, rsphe, center=(0.0, 0.0, 0.0), seed=0, shell=0.9):
    """SYSTEM ... SPHERE RADI rsphe CX CY CZ (shell_simul.f90:870-912): a spherical system
    whose bounding box is the cube of side 2 rsphe around the centre.  Moving ions outside
    the sphere are moved radially inside it, to between shell*rsphe and rsphe, so that BD
    steps reach the surface."""
    s = s.copy()
    s.flags = (s.flags | abi.QSPHERE) & ~abi.QECYL
    cx, cy, cz = (float(v) for v in center)
    s.center, s.rsphe = (cx, cy, cz), float(rsphe)
    s.box = (cx - 0.5 * 2.0 * rsphe, cx + 0.5 * 2.0 * rsphe, cy - 0.5 * 2.0 * rsphe,
             cy + 0.5 * 2.0 * rsphe, cz - 0.5 * 2.0 * rsphe, cz + 0.5 * 2.0 * rsphe)
    s.x, s.y, s.z = s.x.copy(), s.y.copy(), s.z.copy()
    rng = np.random.default_rng(seed)
    for i in range(s.nsites + s.nfix, s.ntot):
        d = np.array([s.x[i] - cx, s.y[i] - cy, s.z[i] - cz])
        r = float(np.sqrt(d @ d))
        if r > shell * rsphe:
            f = rsphe * rng.uniform(shell, 0.995) / max(r, 1e-9)
            s.x[i], s.y[i], s.z[i] = cx + f * d[0], cy + f * d[1], cz + f * d[2]
    return s


def with_ecyl(s: System, seed=0, shell=0.9):
    """SYSTEM ... ECYL LX LY LZ CX CY CZ (shell_simul.f90:883-895): an elliptic cylinder
    inscribed in the system box (semi-axes lx/2, ly/2 around the box centre; iecx = 2/lx,
    iecy = 2/ly), z bounded by the box.  Moving ions outside the ellipse are moved inside
    it, to between shell and 1 of the way out, so that BD steps reach the wall."""
    s = s.copy()
    s.flags = (s.flags | abi.QECYL) & ~abi.QSPHERE
    lx2m, lx2p, ly2m, ly2p, lz2m, lz2p = s.box
    lx, ly = lx2p - lx2m, ly2p - ly2m
    cx, cy, cz = 0.5 * (lx2m + lx2p), 0.5 * (ly2m + ly2p), 0.5 * (lz2m + lz2p)
    s.center = (cx, cy, cz)
    s.iecx, s.iecy = 2.0 / lx, 2.0 / ly
    s.x, s.y = s.x.copy(), s.y.copy()
    rng = np.random.default_rng(seed)
    for i in range(s.nsites + s.nfix, s.ntot):
        ex, ey = (s.x[i] - cx) * s.iecx, (s.y[i] - cy) * s.iecy
        r = float(np.hypot(ex, ey))
        if r > shell:
            f = rng.uniform(shell, 0.995) / max(r, 1e-9)
            s.x[i], s.y[i] = cx + f * (s.x[i] - cx), cy + f * (s.y[i] - cy)
    return s


def sphere_buffers(s: System, conc, shell_len=6.0, mu=(0.0, 0.0)):
    """GCMC buffers of a spherical system: per ion type, the lower and the upper half of
    the outer shell rsphe - shell_len .. rsphe (BUFFER ... RMIN RMAX LZMIN LZMAX,
    shell_simul.f90:975-1000); avnum = conc * the half shell's volume."""
    cx, cy, cz = s.center
    R, r0 = s.rsphe, s.rsphe - shell_len
    ib_types, lzmin, lzmax, avnum, mus, rmin, rmax = [0], [0.0], [0.0], [0.0], [0.0], [0.0], [0.0]
    for half in (0, 1):
        for k in range(s.nion):
            t = s.nold + 1 + k
            lo, hi = (cz - R, cz) if half == 0 else (cz, cz + R)
            dens = conc * (AVOGADRO / LITER) * ANGSTROM ** 3
            ib_types.append(t)
            lzmin.append(lo)
            lzmax.append(hi)
            rmin.append(r0)
            rmax.append(R)
            avnum.append(dens * 2.0 / 3.0 * PI * (R ** 3 - r0 ** 3))
            mus.append(mu[k])
    nb = len(ib_types)
    return Gcmc(ibfftyp=np.array(ib_types, np.int32), avnum=np.array(avnum), mu=np.array(mus),
                kb=np.zeros(nb), lzmin=np.array(lzmin), lzmax=np.array(lzmax),
                bufferbias=np.zeros(nb, np.int8),
                ntotat=np.array([0] + [int(round(a)) for a in avnum[1:]], np.int32),
                lx=2.0 * R, ly=2.0 * R, cx=cx, cy=cy, cz=cz, ikbt=1.0 / s.meta["kbt"],
                rmin=np.array(rmin), rmax=np.array(rmax))


# ------------------------------------------------------ diffusion and reaction field
def lib_spline(x, y):
    """Spline coefficients (b, c, d) from the HIP library's host routine (spline.f90)."""
    lib = abi.load_library()
    x = np.ascontiguousarray(x, float)
    y = np.ascontiguousarray(y, float)
    n = len(x)
    b, c, d = np.zeros(n), np.zeros(n), np.zeros(n)
    dp = C.POINTER(C.c_double)
    rc = lib.bromoc_spline(n, *(a.ctypes.data_as(dp) for a in (x, y, b, c, d)))
    if rc:
        raise abi.BromocError("bromoc_spline", rc)
    return b, c, d


def with_switch_diffusion(s: System, porelength, pcenter=0.0, lowfraction=0.5,
                          switchlength=4.0):
    """DIFFUSION command (shell_simul.f90:1975-2020): every ion type's D drops to
    lowfraction * D inside a pore of the given length, with a cubic switch of
    switchlength at each end (DYNAMICS1)."""
    s = s.copy()
    s.flags = (s.flags | abi.QDIFFUSE) & ~abi.QPROFILE
    s.plength2 = 0.5 * porelength
    s.pcenter = pcenter
    nt1 = s.ntype + 1
    s.ampl3 = np.zeros(nt1)
    s.p3 = np.ones(nt1)
    lf = np.broadcast_to(np.asarray(lowfraction, float), (s.ntype - s.nold,))
    sl = np.broadcast_to(np.asarray(switchlength, float), (s.ntype - s.nold,))
    s.ampl3[s.nold + 1:] = lf
    s.p3[s.nold + 1:] = sl
    return s


def tanh_profile(z0, z1, n, half_pore, width=3.0, low=0.5):
    """Knots (z, D(z)/D0) of a tanh-shaped reduction to `low` inside |z| < half_pore."""
    z = np.linspace(z0, z1, n)
    f = 1.0 - (1.0 - low) * 0.5 * (np.tanh((z + half_pore) / width) -
                                   np.tanh((z - half_pore) / width))
    return z, f


def with_profile_diffusion(s: System, zk, fk, spline_builder=None):
    """PROFILE command (shell_simul.f90:1955-1972): D(z) = D0 * spline(z) (DYNAMICS2).
    spline_builder(x, y) -> (b, c, d); default: the library's bromoc_spline."""
    s = s.copy()
    s.flags = (s.flags | abi.QPROFILE) & ~abi.QDIFFUSE
    zk = np.asarray(zk, float)
    fk = np.asarray(fk, float)
    b, c, d = (spline_builder or lib_spline)(zk, fk)
    s.spline = (zk.copy(), fk.copy(), b, c, d)
    return s


def with_bspline(s: System, thold27=27.0):
    """Repulsion map by the 27-point quadratic B-spline (Qtrln off: vdwgd1spln.f90,
    vdwgd0spln.f90); thold27 as the SIMUL keyword (default 27, shell_simul.f90:2706)."""
    s = s.copy()
    s.flags |= abi.QBSPLINE
    s.thold27 = float(thold27)
    return s


def with_proxdiff(s: System, beta=2.93, diff0=0.101, diffcutoff=None):
    """PROXDIFF command (shell_simul.f90:1901-1910): ion D relaxes towards diff0 near the
    DNA, D(d) = D - (D - diff0) exp(-d / beta) within diffcutoff of the nearest site's table
    start (DYNAMICS3).  diffcutoff defaults to the box diagonal plus 1 A (maxl)."""
    s = s.copy()
    s.flags |= abi.QPROXDIFF
    s.flags &= ~(abi.QDIFFUSE | abi.QPROFILE)
    s.ibeta = 1.0 / beta
    s.diff0 = float(diff0)
    if diffcutoff is None:
        lx, ly, lz = s.box[1] - s.box[0], s.box[3] - s.box[2], s.box[5] - s.box[4]
        diffcutoff = math.sqrt(max(lx, ly) ** 2 + lz ** 2) + 1.0
    s.diffcutoff = float(diffcutoff)
    if s.diffusion is None:
        raise ValueError("PROXDIFF needs the per-type diffusion constants")
    return s


def with_apfor(s: System, seed=0, n=6, fmax=2.0, repeat=True):
    """APFOR: constant forces on n random sites (one site twice when repeat is set)."""
    rng = np.random.default_rng(seed)
    s = s.copy()
    s.flags |= abi.QAPFOR
    sn = rng.choice(s.nsites, n, replace=False).astype(np.int32)
    if repeat and n > 1:
        sn[-1] = sn[0]
    s.apfor = (sn, rng.uniform(-fmax, fmax, (n, 3)))
    return s


def _centroid(v, lo, hi, insites, scale=1.0):
    t = 0.0
    for a in v[lo:hi]:          # Fortran sum, in order
        t += a
    return t * insites * scale


def with_contrans(s: System, entries, unsplit=False, nsites1st=None):
    """CONTRANS command (shell_simul.f90:724-770): entries of (site, (kx, ky, kz)) with site
    1-based or 0 for the centroid; centres as the command sets them from the current
    positions (the site's position, or the centroid; with unsplit each strand's)."""
    s = s.copy()
    s.flags |= abi.QCONTRANS
    ns = s.nsites
    insites = 1.0 / ns
    n1 = ns // 2 if nsites1st is None else nsites1st
    pos = (s.x, s.y, s.z)
    csn = np.array([e[0] for e in entries], np.int32)
    k = np.array([e[1] for e in entries], float).reshape(len(entries), 3)
    centre = np.zeros((len(entries) + 1, 3))
    for q, (site, _) in enumerate(entries):
        for a in range(3):
            if site == 0:
                if unsplit:
                    centre[q, a] = _centroid(pos[a], 0, n1, insites, 2.0)
                    centre[-1, a] = _centroid(pos[a], n1, ns, insites, 2.0)
                else:
                    centre[q, a] = _centroid(pos[a], 0, ns, insites)
            else:
                centre[q, a] = pos[a][site - 1]
    s.contrans = dict(csn=csn, k=k, centre=centre, unsplit=bool(unsplit), nsites1st=n1)
    return s


def with_notrans(s: System, axes=7):
    """NOTRANS command (shell_simul.f90:692-712): after every step the sites' centroid is
    put back, on the axes of the bitmask (1 x, 2 y, 4 z), where it was at the command."""
    s = s.copy()
    s.flags |= abi.QNOTRANS
    insites = 1.0 / s.nsites
    s.notrans = (axes, tuple(_centroid(v, 0, s.nsites, insites) for v in (s.x, s.y, s.z)))
    return s


def with_membrane(s: System, thick=20.0, zmemb=0.0, epsm=2.0, volts=0.1, debye=9.6,
                  amplmemb=None, pmemb=None, pore=None):
    """MEMBRANE command (shell_simul.f90:1563-1640): a slab of thickness `thick` centred at
    zmemb with dielectric epsm, the Nernst potential of `volts` across it (Debye length
    `debye`, shared with the DNA's Debye screening), per atnam2 type wall heights amplmemb
    and switching lengths pmemb; with pore = (radius, amplpore, ppore) a cylindrical pore
    through it (Qpore)."""
    s = s.copy()
    s.flags |= abi.QMEMB
    nt1 = s.nttyp + 1
    half = 0.5 * thick
    volt = volts * COULOMB / KCALMOL
    s.kappa, s.ikappa = debye, 1.0 / debye
    cdie = s.meta.get("cdie", 80.0)
    m = dict(zmemb1=-half + zmemb, zmemb2=half + zmemb, voltage=volt,
             afact=epsm * volt / (2.0 * epsm + cdie * s.ikappa * thick), ceps=cdie / epsm,
             qpore=pore is not None)
    a1 = np.zeros(nt1)
    p1 = np.ones(nt1)
    a2 = np.zeros(nt1)
    p2 = np.ones(nt1)
    rc = np.ones(nt1)
    a1[1:] = 10.0 if amplmemb is None else amplmemb
    p1[1:] = 2.0 if pmemb is None else pmemb
    if pore is not None:
        rc[1:], a2[1:], p2[1:] = pore
    m.update(ampl1=a1, p1=p1, ampl2=a2, p2=p2, rcylinder=rc)
    s.membrane = m
    return s


def with_debyhyb(s: System, debye=9.6, outside=()):
    """The DNA command's DEBYHYB option (shell_simul.f90:820-846): free DNA (DNAFREE) whose
    P-P Coulomb pairs are Debye-screened (length `debye`) when both sites lie outside the
    box (energy.mono.f90:439-441); `outside` lists sites moved out past the box's +x face."""
    s = s.copy()
    s.flags = (s.flags | abi.QDEBYHYB | abi.QDNAFREE) & ~abi.QDEBY
    s.kappa, s.ikappa = debye, 1.0 / debye
    s.x = s.x.copy()
    for k, i in enumerate(outside):
        s.x[i] = s.box[1] + 1.0 + 0.7 * k
    return s


def with_nmcden(s: System, seed=0, density=0.15, mode="scatter"):
    """The PHIV command's NMCDEN option (shell_simul.f90:2720-2750, readphi.f90:176-240):
    one repulsion map per DNA site name (S, P, Ab, Tb, Cb, Gb, with DNA) and per ion type
    nold+1..ntype, so every species reads a map of its own.  mode "scatter": the system's map
    with a seeded scatter of extra wall nodes (density of the grid) -- every particle sees
    its own map, but the rough maps make BD trajectories chaotic; mode "pore": the wall grown
    by 0, 3 or 6 nodes across x and y (a per-species pore radius) -- smooth trajectories.
    SVDW is off (shell_simul.f90:2875)."""
    s = s.copy()
    s.flags = (s.flags | abi.QNMCDEN) & ~abi.QSVDW
    nmaps = (6 if s.flags & abi.QNUCL else 0) + (s.ntype - s.nold if s.flags & abi.QPAR else 0)
    base = np.asarray(s.phiv.data, np.int8)
    rng = np.random.default_rng(seed)
    if mode == "scatter":
        maps = [base | (rng.random(base.shape) < density).astype(np.int8) for _ in range(nmaps)]
    else:
        maps = []
        for k in range(nmaps):
            m = base.copy()
            for _ in range(3 * ((k + int(seed)) % 3)):
                g = m.copy()                  # one node of wall more along +-x and +-y
                g[1:] |= m[:-1]
                g[:-1] |= m[1:]
                g[:, 1:] |= m[:, :-1]
                g[:, :-1] |= m[:, 1:]
                m = g
            maps.append(m)
    maps = np.stack(maps)
    s.phiv = GridData(maps, s.phiv.dcel, s.phiv.xbcen, s.phiv.ybcen, s.phiv.zbcen)
    return s


def with_mmij(s: System, npol=(4, 4, 5), region=None, rfscal=1.0, seed=0, strength=1e-4):
    """MMIJ reaction field in a rectangular region (shell_simul.f90:2780-2831, READMMIJ,
    RECT_NORM): a region of the given widths centred at 0 (default: 80 % of the box), the
    full tensor basis of xnpol*ynpol*znpol Legendre products (degrees in the basis order a
    GSBP matrix file lists them), normalisation sqrt((2l+1)(2m+1)(2n+1) xscal yscal zscal / 8)
    (rect_norm.f90:28-33) and a seeded symmetric negative-definite reaction-field matrix
    (a dielectric's response lowers the energy) standing in for the file's MMIJ."""
    s = s.copy()
    s.flags |= abi.QMMIJ
    if region is None:
        b = s.box
        region = tuple(0.8 * (b[2 * k + 1] - b[2 * k]) for k in range(3))
    scal = tuple(2.0 / w for w in region)
    nx, ny, nz = npol
    lst = np.array([(a, b_, c) for a in range(nx) for b_ in range(ny) for c in range(nz)],
                   np.int32)
    ilxyz = scal[0] * scal[1] * scal[2] * 0.125
    bnorm = np.sqrt((2.0 * lst[:, 0] + 1.0) * (2.0 * lst[:, 1] + 1.0) * (2.0 * lst[:, 2] + 1.0)
                    * ilxyz)
    n = len(lst)
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    m = -(a @ a.T + 0.1 * np.eye(n)) * strength / ilxyz
    m = 0.5 * (m + m.T)
    s.mmij = dict(shape=1, npol=npol, lstp=lst, bnorm=bnorm, mmij=m, scal=scal, rfscal=rfscal)
    return s


def with_rfpar(s: System, seed=0, shape=None, dcel=None, single=False, reffac=1.5,
               wall_radius=None, zero_frac=0.0, const_reff=None):
    """Synthetic reaction-field parameter grids (RFPAR command, rfpar.f90:373-582): per
    ion type (or one grid with single) a smooth gsrfen field and a positive greff field
    (constant if const_reff is given); gsrfen is zero beyond wall_radius from the z axis
    and at a random zero_frac of the nodes (the reference stops a corner row at a zero
    gsrfen node, rfpar.f90:121).  sqrfac = sqrt(reffac) (shell_simul.f90:2860)."""
    s = s.copy()
    rng = np.random.default_rng(seed)
    lx2m, lx2p, ly2m, ly2p, lz2m, lz2p = s.box
    if dcel is None:
        dcel = 1.0
    if shape is None:
        shape = tuple(int(np.ceil((hi - lo) / dcel)) + 3 for lo, hi in
                      ((lx2m, lx2p), (ly2m, ly2p), (lz2m, lz2p)))
    ng = 1 if single else s.ntype - s.nold
    nx, ny, nz = shape
    x = (np.arange(nx) - 0.5 * (nx - 1)) * dcel
    y = (np.arange(ny) - 0.5 * (ny - 1)) * dcel
    z = (np.arange(nz) - 0.5 * (nz - 1)) * dcel
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    gs = np.zeros((ng,) + tuple(shape), np.float32)
    gr = np.zeros_like(gs)
    for k in range(ng):
        a = rng.uniform(0.2, 0.5, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        f = (-0.6 + 0.15 * np.sin(a[0] * X + ph[0]) * np.cos(a[1] * Y + ph[1]) +
             0.1 * np.sin(a[2] * Z + ph[2]))
        if wall_radius is not None:
            f = np.where(X ** 2 + Y ** 2 > wall_radius ** 2, 0.0, f)
        if zero_frac > 0.0:
            f = np.where(rng.random(f.shape) < zero_frac, 0.0, f)
        gs[k] = f
        if const_reff is not None:
            gr[k] = const_reff
        else:
            gr[k] = 2.0 + 0.5 * np.cos(a[1] * X + ph[2]) * np.sin(a[0] * Z + ph[0]) + 0.3 * k
    s.rf = GridData(gs, dcel)
    s.rf_greff = gr
    s.sqrfac = math.sqrt(reffac)
    s.flags |= abi.QRFPAR
    if single:
        s.flags |= abi.QRFPSIN
    return s


# ------------------------------------------------------------------- replica batches
def replica_batch(seed=0, nrep=256, n_range=(100, 4000), box=(96.0, 96.0, 192.0), dcel=0.75,
                  rpore=12.0, half_slab=20.0, conc_range=(0.1, 2.0), volt_range=(-0.5, 0.5),
                  gcmc=True, buffer_len=10.0, mu=(-0.3, -0.3), cap_factor=1.5,
                  table_builder=None, noise_volt=0.01):
    """Independent ion + pore systems for a replica batch: a shared System (pair tables of
    K+/Cl-, the pore repulsion map, box, flux planes, BD constants) and `nrep` Replicas with
    a log-uniform ion count in n_range, their own static field (a linear drop of a uniform
    voltage in volt_range across the slab plus a smooth perturbation) and, with gcmc, bath
    buffers of buffer_len at both box ends for a uniform concentration in conc_range."""
    from .system import Replica, replica_system
    rng = np.random.default_rng(seed)
    shared = pore_dna(seed=seed, n_ions=0, box=box, dcel=dcel, rpore=rpore,
                      half_slab=half_slab, dna=False, table_builder=table_builder, volt=0.0)
    shared.capacity = 1
    shape = shared.phix.data.shape
    nx, ny, nz = shape
    x = (np.arange(nx) - 0.5 * (nx - 1)) * dcel
    y = (np.arange(ny) - 0.5 * (ny - 1)) * dcel
    z = (np.arange(nz) - 0.5 * (nz - 1)) * dcel
    drop = np.clip((half_slab - z) / (2.0 * half_slab), 0.0, 1.0)      # 1 V profile
    to_field = (COULOMB / KCALMOL) / CELEC
    hx, hy, hz = 0.5 * box[0], 0.5 * box[1], 0.5 * box[2]

    def accessible(p):
        return not (abs(p[2]) <= half_slab + 1.0 and p[0] ** 2 + p[1] ** 2 > (rpore - 1.5) ** 2)

    lo = np.array([-hx + 0.5, -hy + 0.5, -hz + 0.5])
    reps = []
    for r in range(nrep):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        n += n % 2
        volt = rng.uniform(*volt_range)
        conc = rng.uniform(*conc_range)
        a = rng.uniform(0.1, 0.3, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        pert = (np.sin(a[0] * x + ph[0])[:, None, None] * np.sin(a[1] * y + ph[1])[None, :, None] *
                np.sin(a[2] * z + ph[2])[None, None, :])
        field = ((volt * drop)[None, None, :] + noise_volt * pert) * to_field
        pos = _place(rng, n, lo, -lo, 2.0, accept=accessible)
        typei = np.array([shared.nold + 1 + (k % 2) for k in range(n)], np.int32)
        order = rng.permutation(n)
        pos, typei = pos[order], typei[order]
        # GRAND drives the ion count towards the bath concentration: leave room for twice
        # the bulk count of the box at that concentration
        bulk = 2 * conc * (AVOGADRO / LITER) * ANGSTROM ** 3 * box[0] * box[1] * box[2]
        cap = max(int(cap_factor * n) + 64, int(2.0 * bulk) + 64) if gcmc else int(cap_factor * n) + 64
        rep = Replica(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), typei=typei,
                      nfix=0, capacity=cap,
                      phix=field.astype(np.float32), meta=dict(volt=volt, conc=conc))
        if gcmc:
            rs = replica_system(shared, rep)
            g = gcmc_buffers(rs, conc=conc, buffer_len=buffer_len, mu=mu, volt=volt)
            ib, nat = assign_buffers(rs, g)
            rep.gcmc, rep.ibuffer, rep.nat = g, ib[:n].copy(), nat
        reps.append(rep)
    return shared, reps


# ------------------------------------------------------- adjoint (differentiation) system
PARAM_PAIRS = ((1, 1), (1, 2), (2, 2), (1, 3))   # K-K, K-Cl, Cl-Cl, K-W


def adjoint_system(seed=0, n_per_type=200, box=60.0, temp=300.0, cdie=80.0, dt=0.02,
                   curve_seed=None, table_builder=None, w_pos=(0.0, 0.0, 0.0)):
    """Plain K+/Cl- ions in a cubic box plus one fixed uncharged wall site W (nfix = 1, the
    first particle).  Six pair types, all tabulated on r = 1.5..20 A at 0.01 A from 12-6 LJ
    (sigma, eps drawn from curve_seed) + Coulomb + a damped oscillatory tail.  The knot
    values of K-K, K-Cl, Cl-Cl and K-W are the fitting parameters:
    meta['params'][is] = (u, chr, fct) of pair type is, meta['rgrid'] the knots' r."""
    rng = np.random.default_rng(seed)
    crng = np.random.default_rng(seed if curve_seed is None else curve_seed)
    charges = (1.0, -1.0, 0.0)
    diff = (0.196, 0.203, 0.0)
    ntypes = 3
    sig = crng.uniform(3.0, 4.5, ntypes)
    eps = crng.uniform(0.1, 0.5, ntypes)
    pidx = nindex_table(ntypes, 0, ntypes)
    npair = ntypes * (ntypes + 1) // 2
    spec = [None] * (npair + 1)
    for a in range(1, ntypes + 1):
        for b in range(a, ntypes + 1):
            spec[pidx[a * (ntypes + 1) + b]] = dict(
                sig=0.5 * (sig[a - 1] + sig[b - 1]), eps=math.sqrt(eps[a - 1] * eps[b - 1]),
                qq=charges[a - 1] * charges[b - 1], cdie=cdie)
    pairs, ep = build_tables(spec, table_builder)
    r = knot_grid()
    params = {}
    for a, b in PARAM_PAIRS:
        is_ = int(pidx[a * (ntypes + 1) + b])
        sp = spec[is_]
        u = efpot_curve(r, sp["sig"], sp["eps"], sp["qq"], sp["cdie"], sp.get("osc", 0.3))
        chr_ = 1 if sp["qq"] != 0.0 else 0
        params[is_] = (u, chr_, CELEC * sp["qq"] / sp["cdie"] if chr_ else 0.0)
    half = 0.5 * box
    wp = np.asarray(w_pos, float)
    pos = _place(rng, 2 * n_per_type, np.full(3, -half), np.full(3, half), 2.0,
                 accept=lambda p: np.sum((p - wp) ** 2) > 9.0)
    typei = np.repeat(np.arange(1, 3, dtype=np.int32), n_per_type)
    order = rng.permutation(len(typei))
    pos, typei = pos[order], typei[order]
    pos = np.concatenate([wp[None, :], pos])
    typei = np.concatenate([np.array([3], np.int32), typei])
    kbt = kbt_of(temp)
    kbtdt = dt * (1.0 / kbt)
    d = np.array((0.0,) + diff)
    s = System(
        ntot=len(typei), nsites=0, nfix=1, ntype=ntypes, nold=0, nttyp=ntypes,
        flags=abi.QPAR | abi.QNONBOND | abi.QNOBOND,
        x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), typei=typei,
        nwtype=np.arange(0, ntypes + 1, dtype=np.int32),
        cg=np.array((0.0,) + charges), fact1a=d * kbtdt, fact2a=np.sqrt(2.0 * dt * d),
        pairidx=pidx, pairs=pairs, ep=ep, res=RES, capacity=len(typei), bdmax=1.0,
        box=(-half, half, -half, half, -half, half), ndna=0, nion=ntypes,
        dt=dt, kbtdt=kbtdt, diffusion=d.copy(),
        meta=dict(workload="adjoint", seed=seed, temp=temp, dt=dt, kbt=kbt, cdie=cdie,
                  params=params, rgrid=r),
    )
    return s


def type_charges(s: System):
    """Charge of each atnam2 type as shell_simul.f90:602-607 assigns it: the P sites carry
    cgnuc (type 2, when the DNA has more than one site type), ion type i its cg(i)."""
    cg2 = np.zeros(s.nttyp + 1)
    if s.ndna > 1:
        cg2[2] = s.cgnuc
    for i in range(s.nold + 1, s.ntype + 1):
        cg2[s.nwtype[i]] = s.cg[i]
    return cg2


def analytic_pair_types(s: System, kind="all"):
    """Pair types (nindex) of the ion-ion pairs ("ion"), the site-ion pairs ("site") or
    both ("all")."""
    nt1, ndna = s.nttyp + 1, s.ndna
    out = []
    for a in range(1, s.nttyp + 1):
        for b in range(a, s.nttyp + 1):
            if b <= ndna:
                continue
            is_ = int(s.pairidx[a * nt1 + b])
            if (a > ndna and kind in ("all", "ion")) or (a <= ndna and kind in ("all", "site")):
                out.append(is_)
    return sorted(set(out))


def with_analytic(s: System, types=None, seed=0, lj=True, srpmf=True, coulomb=True,
                  rth=8.0, cdie=None):
    """Copy of s whose pair types `types` (default: every ion-ion and site-ion type) take
    ENERGY's and INTERACT's analytic branch (energy.mono.f90:135-178,623-668,
    interact.f90:182-204,365-404) instead of an EFPOT table:

    - Coulomb fct = celec q_a q_b / cdie where both atnam2 types are charged
      (shell_simul.f90:606-617);
    - 12-6 Lennard-Jones, epp4 = 4 eps and sgp2 = sigma^2 (shell_simul.f90:1183-1184),
      sigma in [3.0, 4.5] A and eps in [0.1, 0.5] kcal/mol from the seed;
    - SRPMF, the damped-cosine short-range term c0 exp((c1-r)/c2) cos(c3 pi (c1-r)) +
      c4 (c1/r)^6 (shell_simul.f90:1702-1711, c2 kept as its inverse, :1758) up to
      r = rth, smoothed from srpx = rth - 0.5 with srpk = 6.90775527898 / 0.5 and
      srpy = 0.001 (shell_simul.f90:1682-1686).

    Pair types not listed keep their tables, so a partial list gives a mixed system."""
    t = s.copy()
    rng = np.random.default_rng(seed)
    types = analytic_pair_types(t) if types is None else list(types)
    cdie = cdie if cdie is not None else t.meta.get("cdie", 80.0)
    cg2 = type_charges(t)
    nt1 = t.nttyp + 1
    pt = {}
    for a in range(1, nt1):
        for b in range(a, nt1):
            pt.setdefault(int(t.pairidx[a * nt1 + b]), (a, b))
    t.pairs = [None if p is None else dict(p) for p in t.pairs]
    for is_ in types:
        a, b = pt[is_]
        d = t.pairs[is_]
        qq = cg2[a] * cg2[b]
        sig, eps = rng.uniform(3.0, 4.5), rng.uniform(0.1, 0.5)
        d.update(efpot=0, nknots=0,
                 chr=1 if (coulomb and qq != 0.0) else 0,
                 fct=CELEC * qq / cdie if (coulomb and qq != 0.0) else 0.0,
                 lj=1 if lj else 0, epp4=4.0 * eps if lj else 0.0,
                 sgp2=sig * sig if lj else 0.0,
                 srpmf=1 if srpmf else 0,
                 c0=rng.uniform(-0.6, 0.6) if srpmf else 0.0,
                 c1=rng.uniform(2.8, 4.0) if srpmf else 0.0,
                 c2=1.0 / rng.uniform(0.4, 0.8) if srpmf else 0.0,
                 c3=rng.uniform(0.3, 0.6) if srpmf else 0.0,
                 c4=rng.uniform(0.0, 0.05) if srpmf else 0.0)
    if srpmf:
        t.rth = rth * rth
        t.srpx = rth - 0.5
        t.srpk = 6.90775527898 / 0.5
        t.srpy = 0.001
    t.meta = dict(t.meta, analytic=sorted(types))
    return t


def with_knots(s: System, new_u, table_builder=None):
    """Copy of s whose parameter pair types `is` (keys of new_u) take the knot values
    new_u[is]; their tables are rebuilt in place in ep."""
    tb = table_builder or lib_table_builder
    t = s.copy()
    r = t.meta["rgrid"]
    n = len(r)
    ep = t.ep.copy()
    for is_, u in new_u.items():
        _, chr_, fct = t.meta["params"][is_]
        e, p = tb(n, r, np.ascontiguousarray(u, float), chr_, fct)
        off = t.pairs[is_]["nxi"]
        ep[3 * off:3 * (off + n)] = e
        d = _pair_dict(p)
        d["nxi"] = off
        d["efpot"] = 1
        t.pairs[is_] = d
        t.meta["params"][is_] = (np.asarray(u, float).copy(), chr_, fct)
    t.ep = ep
    return t

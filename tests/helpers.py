"""Shared builders and comparisons for the parity tests."""
from __future__ import annotations

import functools

import numpy as np

from bromoc_amd import abi, configs


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def force_err(f, g):
    """Norm-wise relative force error max|dF| / max|F|."""
    return float(np.max(np.abs(f - g)) / max(np.max(np.abs(g)), 1e-300))


def energy_errs(e, r, floor=1e-6):
    """Per-term relative error, norm-wise for tiny terms: a term smaller than `floor` times
    the largest term (e.g. native-geometry stacking energies, which are pure rounding noise
    around zero) is measured against floor * largest instead of against itself."""
    scale = max(abs(v) for v in r.values())
    out = {}
    for k in r:
        d = abs(e[k] - r[k])
        out[k] = d / max(abs(r[k]), floor * scale, 1e-300)
    return out


@functools.lru_cache(maxsize=None)
def plain(seed=0, n_per_type=200, **kw):
    return configs.plain_ions(seed=seed, n_per_type=n_per_type, **kw)


@functools.lru_cache(maxsize=None)
def pore_small(seed=5, n_bp=8, n_ions=200, dna=True):
    return configs.pore_dna(seed=seed, n_bp=n_bp, n_ions=n_ions, box=(40.0, 40.0, 60.0),
                            dcel=0.5, rpore=12.0, half_slab=10.0, dna=dna)


def with_fixed(s, nfix, nonfree=()):
    """Copy of s whose first nfix ions are fixed and whose listed sites are not free."""
    t = s.copy()
    t.nfix = nfix
    for k in nonfree:
        t.stfree[k] = 0
    return t


def near_wall(s, k, seed=0, rpore=12.0):
    """Copy of s with its last k ions moved to within half an Angstrom of the pore wall
    (the pore systems' slab spans |z| <= 10 A)."""
    s = s.copy()
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, k)
    r = rpore + rng.uniform(-0.5, 0.5, k)
    s.x[s.ntot - k:s.ntot] = r * np.cos(th)
    s.y[s.ntot - k:s.ntot] = r * np.sin(th)
    s.z[s.ntot - k:s.ntot] = rng.uniform(-9.0, 9.0, k)
    return s

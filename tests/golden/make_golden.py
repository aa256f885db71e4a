"""Writes the golden vectors of tests/golden/*.npz from the reference's own routines
(oracle/_ref, compiled from the reference sources): ENERGY terms and forces, 3 BD steps
with injected noise, flux counters and INTERACT values, for small seeded systems.

Run in the build container (the reference is not present on GPU machines):
    python tests/golden/make_golden.py
Each file stores the builder arguments and a digest of the built inputs, so a consumer
rebuilds the identical system and can tell if the builders ever drift.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bromoc_amd import abi, configs  # noqa: E402

CASES = {
    "plain_ions_128": ("plain_ions", dict(seed=17, n_per_type=64)),
    "pore_dna_small": ("pore_dna", dict(seed=23, n_bp=6, n_ions=100, box=[40.0, 40.0, 60.0],
                                        dcel=0.5, rpore=12.0, half_slab=10.0)),
}


def build(kind, kw):
    kw = dict(kw)
    if "box" in kw:
        kw["box"] = tuple(kw["box"])
    return getattr(configs, kind)(**kw)


def digest(s):
    h = hashlib.sha256()
    for a in (s.x, s.y, s.z, s.typei, s.ep):
        h.update(np.ascontiguousarray(a).tobytes())
    for g in (s.phix, s.phiv):
        if g is not None:
            h.update(np.ascontiguousarray(g.data).tobytes())
    for n in s.LIST_NAMES:
        h.update(np.ascontiguousarray(s.lists[n][0]).tobytes())
    return h.hexdigest()


def main():
    from oracle.oracle import RefFortran
    for name, (kind, kw) in CASES.items():
        s = build(kind, kw)
        r = RefFortran(s)
        e = r.energy()
        f = r.forces()
        noise = np.random.default_rng(1000 + len(name)).standard_normal((3, 3 * s.ntot))
        r.upload_noise(noise)
        e3 = r.step(3, abi.NOISE_INJECTED, 0)
        p3 = r.positions()[0]
        f3 = r.forces()
        nf, nb = r.flux()
        # INTERACT at the final positions: a few ions (and sites) plus an insertion trial
        pos, typ = r.positions()
        js = [s.nsites, s.nsites + 1, s.ntot - 1] + ([0, 1] if s.nsites else [])
        inter = [r.interact(pos[j], abs(typ[j]) if j >= s.nsites else 0, j) for j in js]
        trial = (1.0, -2.0, 0.5)
        inter.append(r.interact(trial, s.nold + 1, s.ntot))
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            kind=kind, kwargs=json.dumps(kw), digest=digest(s),
            energy_keys=np.array(abi.ENERGY_FIELDS), energy=np.array([e[k] for k in abi.ENERGY_FIELDS]),
            forces=f, noise=noise, energy3=np.array([e3[k] for k in abi.ENERGY_FIELDS]),
            positions3=p3, forces3=f3, nforward=nf, nbackward=nb,
            interact_j=np.array(js + [s.ntot]), interact=np.array(inter),
            trial=np.array(trial))
        print(name, s.ntot, "ener", e["ener"])


if __name__ == "__main__":
    main()

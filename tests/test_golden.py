"""Golden vectors of the reference routines (tests/golden/make_golden.py) against the C
oracle (CPU) and against the HIP path (GPU): ENERGY terms and forces, 3 BD steps with
injected noise, flux counters and INTERACT.  The systems are rebuilt from the stored
builder arguments and must reproduce the stored input digest."""
import glob
import json
import os

import numpy as np
import pytest

from bromoc_amd import abi
from helpers import energy_errs, force_err

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
IDS = [os.path.basename(f)[:-4] for f in FILES]


def _load(path):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden
    g = np.load(path, allow_pickle=False)
    s = make_golden.build(str(g["kind"]), json.loads(str(g["kwargs"])))
    assert make_golden.digest(s) == str(g["digest"]), "system builders drifted"
    return s, g


def _check(eng, s, g, tol_e, tol_f, tol_x):
    keys = [str(k) for k in g["energy_keys"]]
    e = eng.energy()
    assert max(energy_errs(e, dict(zip(keys, g["energy"]))).values()) <= tol_e
    assert force_err(eng.forces(), g["forces"]) <= tol_f
    eng.upload_noise(g["noise"])
    e3 = eng.step(3, abi.NOISE_INJECTED, 0)
    assert max(energy_errs(e3, dict(zip(keys, g["energy3"]))).values()) <= tol_e
    assert float(np.max(np.abs(eng.positions()[0] - g["positions3"]))) <= tol_x
    assert force_err(eng.forces(), g["forces3"]) <= tol_f
    nf, nb = eng.flux()
    assert np.array_equal(nf, g["nforward"]) and np.array_equal(nb, g["nbackward"])
    pos, typ = eng.positions()
    js = g["interact_j"]
    for j, ref in zip(js[:-1], g["interact"][:-1]):
        j = int(j)
        got = eng.interact(pos[j], abs(typ[j]) if j >= s.nsites else 0, j)
        assert abs(got - ref) <= tol_e * max(abs(ref), 1.0)
    got = eng.interact(g["trial"], s.nold + 1, int(js[-1]))
    assert abs(got - g["interact"][-1]) <= tol_e * max(abs(g["interact"][-1]), 1.0)


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_reference_golden(path):
    from oracle.oracle import Oracle
    s, g = _load(path)
    _check(Oracle(s), s, g, 1e-13, 1e-13, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_reproduces_reference_golden(path):
    from bromoc_amd.engine import Engine
    s, g = _load(path)
    _check(Engine(s), s, g, 1e-10, 1e-10, 1e-9)

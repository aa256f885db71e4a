"""The C ABI without a GPU: the library loads, exports every entry point the header
declares, its structs have the header's layout, and argument validation answers with
status codes before any device is touched."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from bromoc_amd import abi, configs
from oracle.oracle import oracle_table_builder

from helpers import pore_small

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "bromoc_gpu.h")


def test_exports_every_declared_symbol():
    lib = abi.load_library()
    names = abi.exported_symbols()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(\w+)$", out, re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)


def test_struct_layout_matches_header(tmp_path):
    structs = {"bromoc_pair": abi.Pair, "bromoc_grid": abi.Grid, "bromoc_system": abi.SystemStruct,
               "bromoc_energies": abi.Energies, "bromoc_gcmc": abi.GcmcStruct,
               "bromoc_replica": abi.ReplicaStruct, "bromoc_map_header": abi.MapHeader}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = structs[cname]
        got = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(val), (cname, field, got, val)


def _create(s):
    lib = abi.load_library()
    st, keep = s.to_struct()
    ctx = C.c_void_p()
    return lib.bromoc_create(C.byref(st), 0, C.byref(ctx)), ctx


def _plain():
    return configs.plain_ions(seed=1, n_per_type=5, table_builder=oracle_table_builder)


def test_create_rejects_unsupported_options():
    # a curved system needs its shape (rsphe, or iecx and iecy), and only one of the two
    for flag in (abi.QSPHERE, abi.QECYL):
        s = _plain()
        s.flags |= flag
        rc, _ = _create(s)
        assert rc == abi.E_ARG
    s = configs.with_sphere(_plain(), 20.0)
    s.flags |= abi.QECYL
    assert _create(s)[0] == abi.E_ARG
    s = configs.with_mmij(_plain(), npol=(3, 3, 3))
    s.mmij = dict(s.mmij, shape=2)                # the SPHERE reaction field (sphe_rf1.f90)
    assert _create(s)[0] == abi.E_UNSUPPORTED


def test_create_rejects_bad_indices():
    s = _plain()
    s.typei = s.typei.copy()
    s.typei[0] = 7                       # no such atnam type
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.pairidx = np.zeros_like(s.pairidx)  # no pair type for present ions
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.pairs[1] = dict(s.pairs[1], nknots=100)   # table range exceeds its knots
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.capacity = s.ntot - 1
    assert _create(s)[0] == abi.E_ARG


def test_create_rejects_inconsistent_options():
    s = _plain()
    s.flags |= abi.QNMCDEN                        # per-species maps without a PHIV map
    s.flags &= ~abi.QPHIV
    assert _create(s)[0] == abi.E_ARG
    s = configs.with_nmcden(pore_small(seed=3, n_ions=4))
    assert _create(s)[0] in (abi.OK, abi.E_NODEVICE)  # valid: only the device is missing
    s.flags |= abi.QSVDW                          # shell_simul.f90:2875
    assert _create(s)[0] == abi.E_ARG
    # the map count crosses the ABI: a phiv of another map count than NMCDEN reads, or of
    # several maps without NMCDEN, is refused before anything is read
    lib = abi.load_library()
    for nm, flip in ((-1, False), (1, False), (0, True)):
        s = configs.with_nmcden(pore_small(seed=3, n_ions=4))
        st, keep = s.to_struct()
        assert st.phiv.nmaps == 6 + s.ntype - s.nold
        st.phiv.nmaps += nm
        if flip:
            st.flags &= ~abi.QNMCDEN
        ctx = C.c_void_p()
        assert lib.bromoc_create(C.byref(st), 0, C.byref(ctx)) == abi.E_ARG, (nm, flip)
    s = _plain()
    s.flags |= abi.QDIFFUSE | abi.QPROFILE        # simul1.f90:338-340
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.flags |= abi.QPROFILE                       # no spline given
    assert _create(s)[0] == abi.E_ARG
    s = configs.with_profile_diffusion(_plain(), [0.0, 1.0, 2.0], [1.0, 0.5, 1.0])
    s.spline = (np.array([1.0, 0.0, 1.0]),) + s.spline[1:]   # equal end knots
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.flags |= abi.QRFPAR                         # no grids given
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.flags |= abi.QBSPLINE                       # no repulsion map
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.flags |= abi.QMEMB                          # no membrane parameters given
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.flags |= abi.QDEBYHYB                       # needs DNA, DNAFREE and a Debye length
    assert _create(s)[0] == abi.E_ARG
    s = _plain()
    s.flags |= abi.QMMIJ                          # no reaction-field matrix given
    assert _create(s)[0] == abi.E_ARG
    s = configs.with_mmij(_plain(), npol=(2, 3, 3))   # rect_rf1.f90 needs 3 polynomials
    assert _create(s)[0] == abi.E_ARG
    s = configs.with_mmij(_plain(), npol=(3, 3, 3))
    s.mmij = dict(s.mmij, lstp=np.where(s.mmij["lstp"] == 2, 3, s.mmij["lstp"]))
    assert _create(s)[0] == abi.E_ARG             # a degree beyond xnpol - 1
    s = configs.with_rfpar(configs.with_mmij(_plain(), npol=(3, 3, 3)))
    assert _create(s)[0] == abi.E_ARG             # shell_simul.f90:2869
    for flag in (abi.QPROXDIFF, abi.QNOTRANS, abi.QAPFOR, abi.QCONTRANS):
        s = _plain()                              # ions only: these need DNA sites
        s.flags |= flag
        assert _create(s)[0] == abi.E_ARG, flag


def test_create_without_device_reports_nodevice():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc, _ = _create(_plain())
    assert rc == abi.E_NODEVICE


def test_build_table_argument_errors():
    lib = abi.load_library()
    r = np.array([1.0])
    ep = np.zeros(3)
    p = abi.Pair()
    dp = C.POINTER(C.c_double)
    assert lib.bromoc_build_table(1, r.ctypes.data_as(dp), r.ctypes.data_as(dp), 0, 0.0,
                                  ep.ctypes.data_as(dp), C.byref(p)) == abi.E_ARG
    r2 = np.array([2.0, 1.0])          # not increasing
    assert lib.bromoc_build_table(2, r2.ctypes.data_as(dp), r2.ctypes.data_as(dp), 0, 0.0,
                                  np.zeros(6).ctypes.data_as(dp), C.byref(p)) == abi.E_ARG


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        abi.load_library(str(tmp_path / "nope.so"))


def test_batch_create_validation():
    from bromoc_amd.engine import Batch
    sh, reps = configs.replica_batch(seed=1, nrep=2, n_range=(10, 20), box=(30.0, 30.0, 60.0),
                                     dcel=1.5, rpore=8.0, half_slab=8.0,
                                     table_builder=oracle_table_builder)
    for flag in (abi.QNUCL, abi.QBSPLINE, abi.QMEMB, abi.QMMIJ, abi.QNMCDEN):
        bad = sh.copy()
        bad.flags |= flag
        with pytest.raises(abi.BromocError) as ei:
            Batch(bad, reps)
        assert ei.value.code == abi.E_UNSUPPORTED
    r0 = reps[0]
    r0.typei = r0.typei.copy()
    r0.typei[0] = 9                     # not an ion type
    with pytest.raises(abi.BromocError) as ei:
        Batch(sh, reps)
    assert ei.value.code == abi.E_ARG
    r0.typei[0] = 1
    r0.nat = r0.nat.copy()
    r0.nat[1] += 1                      # nat out of step with ibuffer
    with pytest.raises(abi.BromocError) as ei:
        Batch(sh, reps)
    assert ei.value.code == abi.E_ARG


def test_bench_refuses_more_than_one_gpu():
    """One GPU per box: bench.py --gpus N > 1 exits with a message instead of running."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""

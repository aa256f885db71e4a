"""Grid files of the reference through the C ABI: READPHI (bromoc_read_map), READRFPAR
(bromoc_read_rfpar) and REPWALLS on the device (bromoc_repwalls)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def _box(box):
    return None if box is None else np.ascontiguousarray(box, float)


def _hdr(h):
    return dict(nclx=h.nclx, ncly=h.ncly, nclz=h.nclz, dcel=h.dcel, xbcen=h.xbcen,
                ybcen=h.ybcen, zbcen=h.zbcen, rec2=list(h.rec2))


def read_map(path, kind, adjust=False, box=None):
    """(header dict, values (nclx, ncly, nclz)): float32 for MAP_PHIX, int8 for MAP_PHIV."""
    lib = abi.load_library()
    b = _box(box)
    bp = b.ctypes.data_as(C.POINTER(C.c_double)) if b is not None else None
    h = abi.MapHeader()
    rc = lib.bromoc_read_map(str(path).encode(), kind, 1 if adjust else 0, bp, C.byref(h), None, 0)
    if rc:
        raise abi.BromocError("bromoc_read_map", rc)
    out = np.zeros((h.nclx, h.ncly, h.nclz), np.float32 if kind == abi.MAP_PHIX else np.int8)
    rc = lib.bromoc_read_map(str(path).encode(), kind, 1 if adjust else 0, bp, C.byref(h),
                             out.ctypes.data_as(C.c_void_p), out.size)
    if rc:
        raise abi.BromocError("bromoc_read_map", rc)
    return _hdr(h), out


def read_rfpar(paths, adjust=False, box=None):
    """(header dict, radion [n], gsrfen (n, nclx, ncly, nclz), greff (same))."""
    lib = abi.load_library()
    b = _box(box)
    bp = b.ctypes.data_as(C.POINTER(C.c_double)) if b is not None else None
    n = len(paths)
    arr = (C.c_char_p * n)(*[str(p).encode() for p in paths])
    h = abi.MapHeader()
    rad = np.zeros(n)
    dp = C.POINTER(C.c_double)
    rc = lib.bromoc_read_rfpar(n, arr, 1 if adjust else 0, bp, C.byref(h),
                               rad.ctypes.data_as(dp), None, None, 0)
    if rc:
        raise abi.BromocError("bromoc_read_rfpar", rc)
    gs = np.zeros((n, h.nclx, h.ncly, h.nclz), np.float32)
    gr = np.zeros_like(gs)
    fp = C.POINTER(C.c_float)
    rc = lib.bromoc_read_rfpar(n, arr, 1 if adjust else 0, bp, C.byref(h), rad.ctypes.data_as(dp),
                               gs.ctypes.data_as(fp), gr.ctypes.data_as(fp), gs.size)
    if rc:
        raise abi.BromocError("bromoc_read_rfpar", rc)
    return _hdr(h), rad, gs, gr


def repwalls(phiv, walls=126, device=0):
    """REPWALLS layers of a binarised map (int8, (nclx, ncly, nclz)), on the device."""
    lib = abi.load_library()
    out = np.ascontiguousarray(phiv, np.int8).copy()
    nx, ny, nz = out.shape
    rc = lib.bromoc_repwalls(nx, ny, nz, out.ctypes.data_as(C.POINTER(C.c_int8)), walls, device)
    if rc:
        raise abi.BromocError("bromoc_repwalls", rc)
    return out


def write_map(path, values, dcel, centre=(0.0, 0.0, 0.0), rec2=(80.0, 2.0, 0.0, 0.0, 0.0, 2.0)):
    """Write a grid file in the reference's record layout (the layout pb-pnp's writer and
    readphi.f90:43-44 share): 4-byte markers, header records, float32 values."""
    v = np.ascontiguousarray(values, np.float32)
    nx, ny, nz = v.shape
    h1 = np.array([nx, ny, nz], np.int32).tobytes() + np.array([dcel, *centre], float).tobytes()
    h2 = np.array(rec2, float).tobytes()
    with open(path, "wb") as f:
        for rec in (h1, h2, v.tobytes()):
            m = np.array([len(rec)], np.int32).tobytes()
            f.write(m + rec + m)


def write_rfpar(path, gsrfen, greff, dcel, centre=(0.0, 0.0, 0.0), radion=2.0):
    """Write one reaction-field parameter file in the layout READRFPAR reads
    (rfpar.f90:435-470): header, (radion, 5 unused), gsrfen, greff."""
    gs = np.ascontiguousarray(gsrfen, np.float32)
    gr = np.ascontiguousarray(greff, np.float32)
    nx, ny, nz = gs.shape
    h1 = np.array([nx, ny, nz], np.int32).tobytes() + np.array([dcel, *centre], float).tobytes()
    h2 = np.array([radion, 0, 0, 0, 0, 0], float).tobytes()
    with open(path, "wb") as f:
        for rec in (h1, h2, gs.tobytes(), gr.tobytes()):
            m = np.array([len(rec)], np.int32).tobytes()
            f.write(m + rec + m)

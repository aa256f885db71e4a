"""Grid files: the library's READPHI / READRFPAR readers (host code of libbromoc_gpu.so)
against the reference's own routines (oracle/_ref), on files written by Fortran's own
unformatted writer and by bromoc_amd.maps; REPWALLS restated in C against the reference's.
Bar: exact (the readers copy float32 values and integer geometry; the crop's centre is the
same FP64 expression).  CPU only; the reference comparisons are skipped where oracle/_ref
is not built.
"""
import numpy as np
import pytest

from bromoc_amd import abi, configs, maps
from oracle.oracle import (oracle_repwalls, ref_available, ref_readphi, ref_readrfpar,
                           ref_repwalls, ref_write_map)

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built here")


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape).astype(np.float32)
    v[rng.random(shape) < 0.3] = 0.0          # zeros: the repulsion map's wall nodes
    return v


@needs_ref
@pytest.mark.parametrize("writer", ["fortran", "python"])
@pytest.mark.parametrize("kind", [abi.MAP_PHIX, abi.MAP_PHIV])
def test_read_map_matches_readphi(tmp_path, writer, kind):
    v = _field((13, 11, 17), 1)
    path = tmp_path / "map.bin"
    if writer == "fortran":
        ref_write_map(path, v, 0.75, centre=(0.5, -0.25, 1.0))
    else:
        maps.write_map(path, v, 0.75, centre=(0.5, -0.25, 1.0))
    h, got = maps.read_map(path, kind)
    dims, geom, want = ref_readphi(path, kind)
    assert (h["nclx"], h["ncly"], h["nclz"]) == dims
    assert (h["dcel"], h["xbcen"], h["ybcen"], h["zbcen"]) == geom
    assert np.array_equal(got.astype(np.float32), want)
    if kind == abi.MAP_PHIV:
        assert np.array_equal(got, (v == 0.0).astype(np.int8))


@needs_ref
@pytest.mark.parametrize("kind", [abi.MAP_PHIX, abi.MAP_PHIV])
@pytest.mark.parametrize("box", [(-2.0, 2.0, -1.5, 3.0, -4.0, 2.5),
                                 (-100.0, 100.0, -1.0, 1.0, 0.3, 0.9)])
def test_read_map_adjust_matches_readphi(tmp_path, kind, box):
    """The crop to the system box (readphi.f90:98-156), off-centre and past the map."""
    v = _field((21, 19, 23), 2)
    path = tmp_path / "map.bin"
    maps.write_map(path, v, 0.5, centre=(0.25, 0.5, -0.5))
    h, got = maps.read_map(path, kind, adjust=True, box=box)
    dims, geom, want = ref_readphi(path, kind, adjust=True, box=box)
    assert (h["nclx"], h["ncly"], h["nclz"]) == dims
    assert (h["dcel"], h["xbcen"], h["ybcen"], h["zbcen"]) == geom
    assert np.array_equal(got.astype(np.float32), want)


@needs_ref
@pytest.mark.parametrize("adjust", [False, True])
def test_read_rfpar_matches_readrfpar(tmp_path, adjust):
    rng = np.random.default_rng(3)
    shape = (9, 10, 12)
    paths = []
    gs_w, gr_w = [], []
    for k in range(2):
        gs = rng.standard_normal(shape).astype(np.float32)
        gr = (2.0 + rng.random(shape)).astype(np.float32)
        p = tmp_path / f"rf{k}.bin"
        maps.write_rfpar(p, gs, gr, 1.0, centre=(0.0, 0.5, 0.0), radion=1.5 + k)
        paths.append(p)
        gs_w.append(gs)
        gr_w.append(gr)
    box = (-2.0, 2.5, -3.0, 3.0, -1.0, 4.0)
    h, rad, gs, gr = maps.read_rfpar(paths, adjust=adjust, box=box)
    dims, geom, gs_r, gr_r = ref_readrfpar(paths, adjust=adjust, box=box)
    assert (h["nclx"], h["ncly"], h["nclz"]) == dims
    assert (h["dcel"], h["xbcen"], h["ybcen"], h["zbcen"]) == geom
    assert np.array_equal(gs, gs_r) and np.array_equal(gr, gr_r)
    assert list(rad) == [1.5, 2.5]
    if not adjust:
        assert np.array_equal(gs, np.stack(gs_w)) and np.array_equal(gr, np.stack(gr_w))


def test_read_rfpar_rejects_mixed_lattices(tmp_path):
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    z = np.zeros((4, 4, 4), np.float32)
    maps.write_rfpar(a, z, z + 1, 1.0)
    maps.write_rfpar(b, z, z + 1, 0.5)
    with pytest.raises(abi.BromocError):
        maps.read_rfpar([a, b])


def test_read_map_rejects_a_truncated_file(tmp_path):
    p = tmp_path / "m.bin"
    maps.write_map(p, np.ones((5, 5, 5), np.float32), 1.0)
    raw = p.read_bytes()
    p.write_bytes(raw[:-10])
    with pytest.raises(abi.BromocError):
        maps.read_map(p, abi.MAP_PHIX)


def _body_map(shape=(30, 28, 32), seed=4):
    """A pore map with closed wall bodies (boxes and a ball) as REPWALLS layers them."""
    phiv = configs.pore_repulsion(shape, 1.0, 6.0, 5.0).astype(np.int8)
    rng = np.random.default_rng(seed)
    for _ in range(3):
        lo = rng.integers(1, 12, 3)
        hi = lo + rng.integers(5, 14, 3)
        phiv[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1
    x, y, z = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    phiv[(x - 20) ** 2 + (y - 8) ** 2 + (z - 22) ** 2 < 40] = 1
    return phiv


@needs_ref
@pytest.mark.parametrize("walls", [1, 3, 126])
def test_repwalls_oracle_matches_reference(walls):
    phiv = _body_map()
    o, r = oracle_repwalls(phiv, walls), ref_repwalls(phiv, walls)
    assert np.array_equal(o, r)
    assert o.max() == min(walls + 1, o.max()) and o.max() > 1


def test_repwalls_layers_by_distance():
    """On a solid block every layer sits one node deeper: a node's final value is one plus
    its Chebyshev distance to the nearest non-wall node, capped at walls + 1."""
    phiv = np.zeros((20, 20, 20), np.int8)
    phiv[3:17, 2:18, 4:15] = 1
    out = oracle_repwalls(phiv, 126)
    idx = np.argwhere(phiv == 1)
    free = np.argwhere(phiv == 0)
    d = np.array([np.min(np.max(np.abs(free - p), axis=1)) for p in idx])
    assert np.array_equal(out[tuple(idx.T)], d)
    out2 = oracle_repwalls(phiv, 2)
    assert np.array_equal(out2[tuple(idx.T)], np.minimum(d, 3))

"""REPWALLS on the device (bromoc_repwalls) against the C oracle's restatement of
repwalls.f90, on maps with closed wall bodies and on the full-size pore map of the
pore + DNA workload (201 x 201 x 401).  Bar: exact (integer layers)."""
import numpy as np
import pytest

from bromoc_amd import configs, maps
from oracle.oracle import oracle_repwalls

from test_maps import _body_map

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("walls", [1, 3, 126])
def test_repwalls_matches_oracle(walls):
    phiv = _body_map()
    g, o = maps.repwalls(phiv, walls), oracle_repwalls(phiv, walls)
    assert np.array_equal(g, o)


def test_repwalls_full_size_pore_map():
    phiv = configs.pore_repulsion((201, 201, 401), 0.5, 12.0, 20.0).astype(np.int8)
    g, o = maps.repwalls(phiv, 126), oracle_repwalls(phiv, 126)
    print("layers", int(g.max()), flush=True)
    assert np.array_equal(g, o) and g.max() > 10

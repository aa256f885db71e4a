"""bromoc_table_vjp_device_noalloc: the table-build Jacobian-vector product with no
allocation (inputs, outputs and workspace staged in a fixed device array per GPU), called
through the C ABI and checked bit for bit against bromoc_table_vjp_device, the entry point it
varies, on two calls in a row -- the second with fewer samples, so stale workspace from the
first call would show -- and its size limit."""
import ctypes as C

import numpy as np
import pytest

from bromoc_amd import abi, configs
from bromoc_amd.engine import Engine
from bromoc_amd.system import ptr

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    r = configs.knot_grid(1.5, 1.5 + 0.01 * (n - 1), 0.01)
    return r, rng.standard_normal(3 * n), rng.standard_normal(3)


def _call(lib, e, name, r, gep, gsc):
    out = np.full(len(r), np.nan)
    rc = getattr(lib, name)(e.ctx, len(r), ptr(np.ascontiguousarray(r)),
                            ptr(np.ascontiguousarray(gep)), ptr(np.ascontiguousarray(gsc)),
                            ptr(out))
    return rc, out


def test_noalloc_matches_device_vjp_twice():
    lib = abi.load_library()
    e = Engine(configs.plain_ions(seed=1, n_per_type=4))
    for n, seed in ((1851, 1), (377, 2)):
        r, gep, gsc = _case(n, seed)
        rc0, want = _call(lib, e, "bromoc_table_vjp_device", r, gep, gsc)
        rc1, got = _call(lib, e, "bromoc_table_vjp_device_noalloc", r, gep, gsc)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(got, want), (n, float(np.max(np.abs(got - want))))
        assert np.count_nonzero(got) > n // 2


def test_noalloc_rejects_beyond_its_workspace():
    lib = abi.load_library()
    e = Engine(configs.plain_ions(seed=1, n_per_type=4))
    n = 190650
    r = 1.0 + 0.001 * np.arange(n)
    rc, _ = _call(lib, e, "bromoc_table_vjp_device_noalloc", r, np.zeros(3 * n), np.zeros(3))
    assert rc == abi.E_CAPACITY
    assert b"workspace" in lib.bromoc_last_error(e.ctx)

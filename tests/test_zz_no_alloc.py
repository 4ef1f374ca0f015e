"""GPU: max_reduce_noalloc (the max_reduce variant that allocates nothing, its partial maxima
in a static device array) against max_reduce (max_reduce.cu:14-41 semantics, which folds
into the default context's residual slot), through the C-ABI: the same value bit for bit on
two calls in a row, the second with a smaller input whose partials must not mix with the
first call's, plus the limits and argument errors."""
import ctypes
import sys

import numpy as np
import pytest

from cfd_amd import lib as L
from cfd_amd.hipmem import DeviceArray
from cfd_amd.params import make_params
from helpers import small_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_max_reduce_noalloc_matches_max_reduce():
    say("case: max_reduce_noalloc vs max_reduce, 2600x1900 then 37x11, NaN, empty")
    lib = L.load()
    prm, x, y, eps, lp = small_case()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    L.check(lib.cfd_legacy_init(ctypes.byref(prm)))
    L.check(lib.cfd_legacy_set_mesh(dp(x), dp(y), dp(eps), dp(lp)))
    try:
        rng = np.random.default_rng(71)
        out = DeviceArray(np.full(1, -1.0))
        # first call: a large input whose maximum sits in the last workgroup's range;
        # second call: a smaller input whose true maximum is far below the first's
        big = rng.uniform(-1.0, 1.0, 2600 * 1900 + 50)
        big[2600 * 1900 - 3] = -1.0e6
        small = rng.uniform(-1.0, 1.0, 37 * 11 + 9)
        small[37 * 11 + 4] = 5.0e3               # past n_i*n_j: not seen
        for a, n_i, n_j in ((big, 2600, 1900), (small, 37, 11)):
            da = DeviceArray(a)
            ref = lib.max_reduce(da.ptr, n_i, n_j)
            assert lib.max_reduce_noalloc(da.ptr, n_i, n_j, out.ptr, None) == 0, L.last_error()
            got = float(out.get()[0])
            say(n_i, n_j, "max_reduce", ref, "noalloc", got)
            assert got == ref == float(np.max(np.abs(a[:n_i * n_j])))
        # NaN wins in both; no entries give 0 in both
        a = rng.uniform(-1.0, 1.0, 5000)
        a[4321] = np.nan
        da = DeviceArray(a)
        assert lib.max_reduce_noalloc(da.ptr, 50, 100, out.ptr, None) == 0
        assert np.isnan(out.get()[0]) and np.isnan(lib.max_reduce(da.ptr, 50, 100))
        assert lib.max_reduce_noalloc(da.ptr, 0, 100, out.ptr, None) == 0
        assert out.get()[0] == 0.0 == lib.max_reduce(da.ptr, 0, 100)
        # limits and arguments
        assert lib.max_reduce_noalloc(da.ptr, 1 << 20, 1 << 17, out.ptr, None) == -1
        assert "maximum" in L.last_error()
        assert lib.max_reduce_noalloc(da.ptr, -1, 5, out.ptr, None) == -1
        assert lib.max_reduce_noalloc(da.ptr, 5, 5, None, None) == -1
    finally:
        lib.desalocar_globais()

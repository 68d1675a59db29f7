"""lb_stats_noalloc (no counterpart in the reference; for callers that cannot
allocate): through the C ABI, its eight outputs equal lb_stats' bit for bit,
on two calls in a row -- a level of a 3-level case, then a smaller level of
a second, smaller case -- so rows left in the shared scratch array by the
larger call do not leak into the smaller one."""
import ctypes as C

import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.solver import Solver, build_case


def _stats(lib, s, level_id, fn):
    out = np.full(8, np.nan)
    rc = getattr(lib, fn)(s._ctx, level_id, N.ptr(out, C.c_double))
    assert rc == 0, lib.lb_last_error(s._ctx).decode()
    return out


@pytest.mark.gpu
def test_stats_noalloc_equals_stats(hip_device):
    lib = N.hip_lib()
    built = []
    for res, lv in ((16, 3), (8, 2)):
        cfg, mesh = cases.tiny_sphere(res, lv)
        grids, params, mesh = build_case(cfg, mesh)
        s = Solver(cfg, grids, params, mesh)
        s.init_equilibrium()
        s.execute_timestep_batch(1, 4, 0.01)
        s.synchronize()
        built.append((s, grids))
    big, small = built
    # the largest level of the first case, then the smallest level of the second
    lb = max(big[1], key=lambda g: g.n_blocks).level_id
    ls = min(small[1], key=lambda g: g.n_blocks).level_id
    assert big[1][lb - 1].n_blocks > small[1][ls - 1].n_blocks
    bytes0 = (big[0].device_bytes, small[0].device_bytes)
    a1 = _stats(lib, big[0], lb, "lb_stats_noalloc")
    a2 = _stats(lib, small[0], ls, "lb_stats_noalloc")
    # nothing was allocated by the two calls
    assert (big[0].device_bytes, small[0].device_bytes) == bytes0
    b1 = _stats(lib, big[0], lb, "lb_stats")
    b2 = _stats(lib, small[0], ls, "lb_stats")
    assert big[0].device_bytes > bytes0[0]          # lb_stats allocated its partial rows
    for a, b in ((a1, b1), (a2, b2)):
        assert np.array_equal(a, b), (a, b)
        assert a[5] > 0 and np.isfinite(a).all()
    # and again in the other order, after lb_stats has run
    assert np.array_equal(_stats(lib, small[0], ls, "lb_stats_noalloc"), b2)
    assert np.array_equal(_stats(lib, big[0], lb, "lb_stats_noalloc"), b1)

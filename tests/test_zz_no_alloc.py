"""cpmp_reduce_noalloc (partials in a static device array, nothing allocated by
the call) against cpmp_reduce, through the C ABI: two calls in a row, the second
on a smaller lattice, each bit for bit, then the NaN report."""
import ctypes as C

import numpy as np
import pytest

import gpu_cases as gc

pytestmark = pytest.mark.gpu


def values(s):
    return [s.mass, s.max_speed, s.bad_count, s.first_bad, s.rho_min, s.rho_max, *s.momentum, s.kinetic]


def both(cpmp, g):
    """(rc, stats) of cpmp_reduce_noalloc, then of cpmp_reduce, on the same state"""
    a, b = cpmp.Stats(), cpmp.Stats()
    ra = cpmp.lib.cpmp_reduce_noalloc(g.h, C.byref(a))
    rb = cpmp.lib.cpmp_reduce(g.h, C.byref(b))
    return ra, values(a), rb, values(b)


def test_reduce_noalloc_matches_reduce(cpmp):
    # first call: a 64x16x64 FP32 lattice (fused density path) after 30 steps
    big = dict(gc.CASES["rke_coalesce"], dims=(64, 16, 64), spheres=[(20, 8, 32, 10), (43, 8, 32, 10)])
    _, g1 = gc.build(big, cpmp, "fp32")
    g1.step(30)
    ra, a, rb, b = both(cpmp, g1)
    print(f"[no-alloc] 64x16x64 fp32: mass {a[0]:.12e} vs {b[0]:.12e}", flush=True)
    assert ra == rb == 0
    assert a == b
    # second call, smaller input: a 16x12x20 FP64 walled lattice
    small = dict(gc.CASES["ref_prw_walls"], dims=(16, 12, 20), radius=5)
    _, g2 = gc.build(small, cpmp, "fp64")
    g2.step(10)
    ra, a, rb, b = both(cpmp, g2)
    print(f"[no-alloc] 16x12x20 fp64: mass {a[0]:.12e} vs {b[0]:.12e}", flush=True)
    assert ra == rb == 0
    assert a == b
    # the first lattice again (the array keeps the last call's contents; nothing stale leaks)
    ra, a, rb, b = both(cpmp, g1)
    assert ra == rb == 0 and a == b
    # a bad density is reported the same way (CPMP_ENAN, first bad node)
    rho = g2.get("rho")
    rho[(3 * 12 + 4) * 20 + 7] = np.nan
    g2.set("rho", rho)
    ra, a, rb, b = both(cpmp, g2)
    assert ra == rb == -5
    assert a[2] == b[2] == 1 and a[3] == b[3] == (3 * 12 + 4) * 20 + 7
    g1.close()
    g2.close()

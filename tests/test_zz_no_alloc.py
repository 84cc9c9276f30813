"""wos_closest_points_noalloc (the entry point that allocates no device
memory: its staging lives in the library's fixed per-GPU array) against
wos_closest_points, bit for bit, on two calls in a row (the second smaller,
so stale staging from the first would show), FP64 and FP32, 2D and 3D; its
size limit and its refusal to build the FP32 grid."""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

MAX_3D, MAX_2D = 246704, 322614


def _same(a, b, label):
    for k in ("dist", "lower", "nearest", "prim"):
        assert np.array_equal(a[k], b[k], equal_nan=True), (label, k)


@pytest.mark.parametrize("dim", [3, 2])
def test_noalloc_equals_closest_points(dim):
    ctx = Context(0)
    if dim == 3:
        ctx.set_mesh(*M.icosphere(4))
        lo, hi = [-1.2] * 3, [1.2] * 3
    else:
        ctx.set_mesh(*M.annulus(128))
        lo, hi = [-1.2] * 2, [1.2] * 2
    rng = np.random.default_rng(dim)
    big = rng.uniform(lo, hi, (200000 if dim == 3 else 300000, dim))
    small = rng.uniform(lo, hi, (777, dim))
    # FP32 with relax needs the grid: refused before any FP32 call built it
    with pytest.raises(RuntimeError, match="distance grid"):
        ctx.closest_points(small, WOS_FP32_PERSISTENT, 0.2, noalloc=True)
    for prec, relax in ((WOS_FP64_PLAIN, 0.0), (WOS_FP32_PERSISTENT, 0.0),
                        (WOS_FP32_PERSISTENT, 0.2)):
        for pts in (big, small):                   # two calls, the second smaller
            want = ctx.closest_points(pts, prec, relax)
            got = ctx.closest_points(pts, prec, relax, noalloc=True)
            _same(got, want, (dim, prec, relax, len(pts)))
    limit = MAX_3D if dim == 3 else MAX_2D
    ok = ctx.closest_points(rng.uniform(lo, hi, (limit, dim)), WOS_FP64_PLAIN, noalloc=True)
    assert np.all(np.isfinite(ok["dist"]))
    with pytest.raises(RuntimeError, match="stages at most"):
        ctx.closest_points(rng.uniform(lo, hi, (limit + 1, dim)), WOS_FP64_PLAIN, noalloc=True)
    ctx.close()


def test_noalloc_on_second_context_of_the_device():
    """Two contexts on one GPU share its array: calls in turn stay correct."""
    a, b = Context(0), Context(0)
    a.set_mesh(*M.icosphere(3))
    b.set_mesh(*M.torus(1.0, 0.35, 60, 30))
    rng = np.random.default_rng(9)
    pa = rng.uniform(-1.1, 1.1, (5000, 3))
    pb = rng.uniform([-1.4, -1.4, -0.4], [1.4, 1.4, 0.4], (3000, 3))
    for _ in range(2):
        _same(a.closest_points(pa, noalloc=True), a.closest_points(pa), "a")
        _same(b.closest_points(pb, noalloc=True), b.closest_points(pb), "b")
    a.close()
    b.close()

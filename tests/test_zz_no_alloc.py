"""appfis_reduce_device_noalloc: the reduction entry point that allocates nothing (its
scratch is a fixed device array per GPU).  Through the C ABI it must give exactly what
appfis_reduce_device gives, on consecutive calls (the second on a smaller input, so a
stale partial or ticket left by the first would show), and it must replay correctly
from a captured graph."""
import ctypes as C
import math

import numpy as np
import pytest

from appfis_amd import hip

pytestmark = pytest.mark.gpu

# (numpy shape (z, y, x) padded, box {x0,x1,y0,y1,z0,z1}); the second is much smaller
SHAPES = [
    ((41, 70, 130), [1, 129, 1, 69, 1, 40]),
    ((19, 35), [1, 34, 1, 18]),
]
CASES = [(hip.U8, hip.SUM), (hip.U8, hip.MAX), (hip.I32, hip.MIN), (hip.I32, hip.SUM),
         (hip.F32, hip.SUM), (hip.F32, hip.MAX), (hip.F64, hip.SUM), (hip.F64, hip.MIN)]


def _input(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    npd = hip.NP_DTYPE[dtype]
    if dtype == hip.U8:
        return rng.integers(0, 256, size=shape, dtype=np.uint8)
    if dtype == hip.I32:
        return rng.integers(-2**31, 2**31 - 1, size=shape, dtype=np.int32)
    return (rng.standard_normal(size=shape) * 1e3).astype(npd)


def _reduce(fn, ctx, arr, box, op, out):
    L = hip.lib()
    dims = (C.c_int * arr.ndim)(*arr.dims)
    bx = (C.c_int * len(box))(*box)
    hip.check(getattr(L, fn)(ctx.handle, arr.ptr, arr.dtype, arr.ndim, dims, bx, op, out.ptr), fn)
    return out.download().tobytes()


def _expected(a, box, op):
    sl = tuple(slice(box[2 * k], box[2 * k + 1]) for k in reversed(range(a.ndim)))
    v = a[sl]
    if op == hip.SUM:
        return int(v.astype(np.int64).sum()) if v.dtype.kind in "ui" else math.fsum(v.astype(np.float64).ravel())
    return v.max() if op == hip.MAX else v.min()


@pytest.mark.parametrize("dtype,op", CASES, ids=lambda x: str(x))
def test_noalloc_matches_reduce_device_on_two_calls(ctx, dtype, op):
    outs = {fn: ctx.array((1,), hip.I64) for fn in ("appfis_reduce_device", "appfis_reduce_device_noalloc")}
    for k, (shape, box) in enumerate(SHAPES):
        a = _input(shape, dtype, 100 * dtype + 10 * op + k)
        d = ctx.from_numpy(a)
        print(f"[noalloc] dtype {dtype} op {op} shape {shape}", flush=True)
        got = {fn: _reduce(fn, ctx, d, box, op, out) for fn, out in outs.items()}
        assert got["appfis_reduce_device_noalloc"] == got["appfis_reduce_device"]
        # and both are the reduction itself
        raw = got["appfis_reduce_device_noalloc"]
        if op == hip.SUM:
            val = np.frombuffer(raw, np.int64 if dtype in (hip.U8, hip.I32) else np.float64)[0]
            exp = _expected(a, box, op)
            if dtype in (hip.U8, hip.I32):
                assert int(val) == exp
            else:  # compensated FP64 sum: within a few ulps of the exactly rounded sum
                sl = tuple(slice(box[2 * q], box[2 * q + 1]) for q in reversed(range(a.ndim)))
                assert abs(float(val) - exp) <= 1e-14 * float(np.abs(a[sl].astype(np.float64)).sum())
        else:
            val = np.frombuffer(raw[:a.itemsize], a.dtype)[0]
            assert val == _expected(a, box, op)


def test_noalloc_replays_from_a_graph(ctx):
    """Captured once, replayed three times with eager work and synchronisations in
    between; every replay equals the eager result on the same data."""
    shape, box = (66, 258), [1, 257, 1, 65]
    a = _input(shape, hip.F64, 7)
    d = ctx.from_numpy(a)
    out_g, out_e = ctx.array((1,), hip.F64), ctx.array((1,), hip.F64)
    L = hip.lib()
    dims = (C.c_int * 2)(*d.dims)
    bx = (C.c_int * 4)(*box)
    ctx.graph_begin()
    st = L.appfis_reduce_device_noalloc(ctx.handle, d.ptr, hip.F64, 2, dims, bx, hip.SUM, out_g.ptr)
    g = ctx.graph_end()
    hip.check(st, "noalloc under capture")
    try:
        for rep in range(3):
            d.upload(a * (rep + 1))
            ctx.graph_launch(g, 1)
            ctx.sync()
            hip.check(L.appfis_reduce_device_noalloc(ctx.handle, d.ptr, hip.F64, 2, dims, bx, hip.SUM,
                                                    out_e.ptr), "noalloc eager")
            ge, ee = out_g.download()[0], out_e.download()[0]
            print(f"[noalloc] graph replay {rep}: {ge!r} eager {ee!r}", flush=True)
            assert ge == ee
            v = (a[1:65, 1:257] * (rep + 1)).ravel()
            assert abs(ge - math.fsum(v)) <= 1e-14 * float(np.abs(v).sum())
    finally:
        ctx.graph_destroy(g)

"""GPU parity of the runtime entry points (C ABI) against the oracle.

seeded inputs, Grid::SetBoundary / SetBoundaryCopy, the single-rank Exchange and the
Execute*Reduce box reductions: bit-exact wherever the result is integer/byte work or a
pure data movement; compensated sums within a few ulps of math.fsum.
"""
import math

import numpy as np
import pytest

import oracle
from appfis_amd import hip

pytestmark = pytest.mark.gpu

NPDT = {hip.U8: np.uint8, hip.I32: np.int32, hip.F32: np.float32, hip.F64: np.float64, hip.F16: np.float16}


def test_abi_version_and_roundtrip(ctx):
    assert hip.lib().appfis_abi_version() == 2
    a = np.arange(1000, dtype=np.float64)
    d = ctx.from_numpy(a)
    assert np.array_equal(d.download(), a)


@pytest.mark.parametrize("nbytes,src_off,dst_off", [(4096, 0, 0), (1000003, 0, 0), (77, 16, 32), (5000, 3, 0),
                                                     (16, 0, 16), (100000, 16, 5)])
def test_memcpy_d2d(ctx, nbytes, src_off, dst_off):
    # 16-byte aligned copies run the library's copy kernel (+ a tail copy), others the runtime's
    rs = np.random.RandomState(nbytes)
    a = rs.randint(0, 256, size=nbytes + 64).astype(np.uint8)
    src = ctx.from_numpy(a)
    dst = ctx.array((nbytes + 64,), hip.U8)
    hip.check(hip.lib().appfis_memcpy_d2d(ctx.handle, dst.ptr + dst_off, src.ptr + src_off, nbytes), "d2d")
    got = dst.download()
    assert np.array_equal(got[dst_off:dst_off + nbytes], a[src_off:src_off + nbytes])
    assert not got[:dst_off].any() and not got[dst_off + nbytes:].any()


@pytest.mark.parametrize("dt", [hip.F64, hip.F32, hip.F16])
@pytest.mark.parametrize("shape,ghost", [((37, 29), 1), ((9, 13, 17), 4), ((100,), 2)])
def test_fill_uniform_matches_oracle(ctx, dt, shape, ghost):
    d = ctx.array(shape, dt)
    ctx.fill_uniform(d, ghost, 1234, -0.5, 2.0)
    ref = np.zeros(shape, dtype=NPDT[dt])
    oracle.fill_uniform(ref, ghost, 1234, -0.5, 2.0)
    assert np.array_equal(d.download().view(np.uint8), ref.view(np.uint8))


def test_fill_uniform_bf16_and_bernoulli(ctx):
    d = ctx.array((12, 20, 18), hip.BF16)
    ctx.fill_uniform(d, 4, 77)
    ref = np.zeros((12, 20, 18), dtype=np.uint16)
    oracle.fill_uniform(ref, 4, 77, bf16=True)
    assert np.array_equal(d.download(), ref)
    g = ctx.array((130, 258), hip.U8)
    ctx.fill_bernoulli(g, 1, 1002, 0.25)
    r = np.zeros((130, 258), dtype=np.uint8)
    oracle.fill_bernoulli(r, 1, 1002, 0.25)
    assert np.array_equal(g.download(), r)


@pytest.mark.parametrize("dt", [hip.U8, hip.I32, hip.F64, hip.F16])
@pytest.mark.parametrize("shape,ghost", [((11, 14), 1), ((10, 12, 9), 2), ((16,), 3), ((13, 9, 11), 1)])
def test_boundary_and_exchange_match_oracle(ctx, dt, shape, ghost):
    rs = np.random.RandomState(sum(shape) + ghost)
    base = rs.randint(0, 100, size=shape).astype(NPDT[dt])
    nd = len(shape)
    for face in [-1] + list(range(2 * nd)):
        d = ctx.from_numpy(base, dt)
        ctx.set_boundary(d, ghost, face, 7)
        r = base.copy()
        oracle.set_boundary(r, ghost, face, 7)
        assert np.array_equal(d.download(), r), ("set_boundary", face)
        d = ctx.from_numpy(base, dt)
        ctx.set_boundary_copy(d, ghost, face)
        r = base.copy()
        oracle.set_boundary_copy(r, ghost, face)
        assert np.array_equal(d.download(), r), ("set_boundary_copy", face)
    if all(s - 2 * ghost >= 2 * ghost for s in shape):
        d = ctx.from_numpy(base, dt)
        ctx.exchange(d, ghost, True)
        r = base.copy()
        oracle.exchange(r, ghost, True)
        assert np.array_equal(d.download(), r)


def test_exchange_rejects_undersized_subgrid(ctx):
    d = ctx.array((5, 5), hip.F64)
    with pytest.raises(hip.AppfisError):
        ctx.exchange(d, 2, True)


@pytest.mark.parametrize("dt", [hip.U8, hip.I32, hip.F32, hip.F64])
def test_reduce_matches_oracle(ctx, dt):
    rs = np.random.RandomState(dt)
    shape = (21, 37, 45)
    if dt in (hip.U8, hip.I32):
        a = rs.randint(0, 3 if dt == hip.U8 else 1000, size=shape).astype(NPDT[dt])
        if dt == hip.I32:
            a -= 500
    else:
        a = (rs.rand(*shape) - 0.3).astype(NPDT[dt])
    d = ctx.from_numpy(a, dt)
    for box in ([1, 44, 1, 36, 1, 20], [0, 45, 0, 37, 0, 21], [5, 6, 3, 30, 2, 3]):
        for op in (hip.MAX, hip.MIN):
            assert ctx.reduce(d, box, op) == oracle.reduce(a, box, op), (box, op)
        sub = a[box[4]:box[5], box[2]:box[3], box[0]:box[1]]
        w = ctx.reduce_sum_wide(d, box)
        if dt in (hip.U8, hip.I32):
            assert w == int(sub.astype(np.int64).sum())
            # the T-typed SUM wraps like the reference's accumulation in T
            bits = 8 if dt == hip.U8 else 32
            wrapped = int(sub.astype(np.int64).sum()) % (1 << bits)
            assert int(ctx.reduce(d, box, hip.SUM)) % (1 << bits) == wrapped
        else:
            exact = math.fsum(sub.astype(np.float64).ravel())
            assert abs(w - exact) <= 4 * np.spacing(abs(exact))


def test_reduce_max_quirk_and_large(ctx):
    a = -np.ones((64, 64), dtype=np.float32)
    d = ctx.from_numpy(a)
    assert ctx.reduce(d, [1, 63, 1, 63], hip.MAX) == np.finfo(np.float32).tiny
    # a box larger than one grid-stride pass, repeated calls reuse the ticket
    big = np.zeros((257, 257, 130), dtype=np.float64)
    oracle.fill_uniform(big, 1, 99)
    db = ctx.from_numpy(big)
    box = [1, 129, 1, 256, 1, 256]
    exact = math.fsum(big[1:256, 1:256, 1:129].ravel())
    for _ in range(3):
        assert abs(ctx.reduce_sum_wide(db, box) - exact) <= 4 * np.spacing(exact)


def test_graph_capture_of_plain_heat2d_and_device_reduction(ctx):
    """The capture-safe entry points captured into one graph (PLAIN Heat2D sweeps, then a
    device-side SUM of the result) and replayed three times, with eager uploads and
    synchronisations between the replays: every replay equals the oracle.  Entry points
    that would synchronise or allocate refuse under the capture and enqueue nothing."""
    import ctypes as C

    nx, ny, it = 130, 67, 10
    a = np.zeros((ny + 2, nx + 2))
    oracle.fill_uniform(a, 1, 1001)
    d0, d1 = ctx.from_numpy(a), ctx.array(a.shape, hip.F64)
    out = ctx.array((1,), hip.F64)
    L = hip.lib()
    dims = (C.c_int * 2)(nx + 2, ny + 2)
    box = (C.c_int * 4)(1, nx + 1, 1, ny + 1)
    # an uncaptured call sizes the context scratch of appfis_reduce_device
    hip.check(L.appfis_reduce_device(ctx.handle, d1.ptr, hip.F64, 2, dims, box, hip.SUM, out.ptr), "warm")
    ctx.sync()
    ctx.graph_begin()
    try:
        ctx.heat2d(d0, d1, nx, ny, 0.1, it, hip.PLAIN)
        st_dev = L.appfis_reduce_device(ctx.handle, d0.ptr, hip.F64, 2, dims, box, hip.SUM, out.ptr)
        host = np.zeros(1)
        st_host = L.appfis_reduce(ctx.handle, d0.ptr, hip.F64, 2, dims, box, hip.SUM, host.ctypes.data)
        msg = L.appfis_last_error().decode()
    finally:
        g = ctx.graph_end()
    try:
        assert st_dev == 0
        assert st_host == 3 and "captured" in msg, (st_host, msg)  # APPFIS_ERR_UNSUPPORTED
        for rep in range(3):
            scale = 1.0 + rep
            src = a * scale
            d0.upload(src)
            ctx.graph_launch(g, 1)
            ctx.sync()
            ref = oracle.heat2d(src.copy(), np.zeros_like(src), nx, ny, 0.1, it)
            got = d0.download()
            print(f"[graph] replay {rep}: field bit-exact {np.array_equal(got, ref)}", flush=True)
            assert np.array_equal(got, ref)
            exp = math.fsum(ref[1:-1, 1:-1].ravel())
            assert abs(out.download()[0] - exp) <= 1e-14 * float(np.abs(ref).sum())
    finally:
        ctx.graph_destroy(g)


def test_kernel_timing_of_gol_calls(ctx):
    # appfis_kernel_timing: events around the generation kernel's launches of a gol call
    # (what bench.py's headline roofline divides by); the call's results are unchanged
    nx, ny = 2048, 64
    a = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
    oracle.fill_bernoulli(a, 1, 77, 0.3)
    d0, d1 = ctx.from_numpy(a), ctx.array(a.shape, hip.U8)
    with pytest.raises(hip.AppfisError):
        ctx.kernel_timing(True)
        ctx.kernel_time()  # nothing timed yet
    ctx.gol(d0, d1, nx, ny, 5, hip.PLAIN)
    ms, launches, gens = ctx.kernel_time()
    assert (launches, gens) == (5, 5) and ms > 0.0
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, 20)
    names = hip.gol_candidates()
    # a full-row variant runs all 20 generations in ceil(20 / 8) launches, the first
    # packing the uint8 field as it loads, the last writing both uint8 buffers
    d0.upload(a)
    ctx.gol(d0, d1, nx, ny, 20, hip.CANDIDATE0 + names.index("tw1 k3 g8 w16"))
    ms, launches, gens = ctx.kernel_time()
    assert (launches, gens) == (3, 20) and ms > 0.0
    assert np.array_equal(d0.download(), ref)
    # the other forms: k_pack, generations 1 .. 19 blocked (ceil(19 / 4) launches), the
    # last one fused with the unpacking
    d0.upload(a)
    ctx.gol(d0, d1, nx, ny, 20, hip.CANDIDATE0 + names.index("tb k10 g4 w16"))
    ms, launches, gens = ctx.kernel_time()
    assert (launches, gens) == (5, 19) and ms > 0.0
    assert np.array_equal(d0.download(), ref)
    ctx.kernel_timing(False)
    ctx.gol(d0, d1, nx, ny, 2, hip.PLAIN)
    with pytest.raises(hip.AppfisError):
        ctx.kernel_time()  # timing is off: no call was timed

"""Edge cases of the C ABI on the GPU: zero step counts, the smallest grids each entry
point accepts, and arguments it must refuse without touching device memory.

Zero steps must leave both buffers byte-for-byte as they were (the reference's time
loops simply do not run, e.g. heat_appfis.hpp:86-91).  The smallest grids run through
the same kernels as the large ones (their bounds checks and periodic wrap) and are
compared with the oracle bit for bit.  Refused arguments raise AppfisError, and the
context keeps working afterwards.
"""
import numpy as np
import pytest

import oracle
from appfis_amd import hip

pytestmark = pytest.mark.gpu

FDTD_COEFF = [0.5, 0.125, 0.0625, 0.125 / 3, 0.03125]


def _pair(ctx, a, dt):
    b = np.full_like(a, 7) if a.dtype == np.uint8 else np.full_like(a, 0.25)
    return ctx.from_numpy(a, dt), ctx.from_numpy(b, dt), b


def test_zero_steps_leave_buffers_untouched(ctx):
    h = np.zeros((19, 23))
    oracle.fill_uniform(h, 1, 11)
    g = np.zeros((12, 34), np.uint8)
    oracle.fill_bernoulli(g, 1, 12, 0.25)
    f = np.zeros((13, 14, 15), np.float32)
    oracle.fill_uniform(f, 4, 13)
    for a, dt, run in [
        (h, hip.F64, lambda d0, d1: ctx.heat2d(d0, d1, 21, 17, 0.1, 0, hip.PLAIN)),
        (h, hip.F64, lambda d0, d1: ctx.heat2d(d0, d1, 21, 17, 0.1, 0, hip.CHOSEN)),
        (g, hip.U8, lambda d0, d1: ctx.gol(d0, d1, 32, 10, 0, hip.PLAIN)),
        (g, hip.U8, lambda d0, d1: ctx.gol(d0, d1, 32, 10, 0, hip.CHOSEN)),
        (f, hip.F32, lambda d0, d1: ctx.fdtd3d(d0, d1, 7, 6, 5, FDTD_COEFF, 0, hip.PLAIN)),
        (f, hip.F32, lambda d0, d1: ctx.fdtd3d(d0, d1, 7, 6, 5, FDTD_COEFF, 0, hip.CHOSEN)),
    ]:
        d0, d1, b = _pair(ctx, a, dt)
        run(d0, d1)
        assert np.array_equal(d0.download(), a) and np.array_equal(d1.download(), b)
    ens = [_pair(ctx, h, hip.F64)[:2]]
    ctx.heat2d_ensemble([(d0, d1, 21, 17) for d0, d1 in ens], 0.1, 0)
    assert np.array_equal(ens[0][0].download(), h)


def test_himeno_zero_iterations(ctx):
    shape = (5, 4, 6)
    arrs = [np.full(shape, 0.5 + i, np.float64) for i in range(14)]
    dev = [ctx.from_numpy(a, hip.F64) for a in arrs]
    gosa = ctx.himeno(dev, 6, 4, 5, 0.8, 0)
    assert len(gosa) == 0
    assert all(np.array_equal(d.download(), a) for d, a in zip(dev, arrs))


@pytest.mark.parametrize("variant", [hip.PLAIN, hip.CHOSEN])
def test_smallest_grids_bit_exact(ctx, variant):
    # heat 1 x 1 and 1 x 3 interiors
    for nx, ny in [(1, 1), (1, 3), (3, 1)]:
        a = np.zeros((ny + 2, nx + 2))
        oracle.fill_uniform(a, 1, 21 + nx)
        d0, d1 = ctx.from_numpy(a, hip.F64), ctx.array(a.shape, hip.F64)
        ctx.heat2d(d0, d1, nx, ny, 0.1, 5, variant)
        assert np.array_equal(d1.download(), oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, 5))
    # periodic GoL at its 2 x 2 minimum, 3 x 3 (buffer not a multiple of 16 bytes), and
    # one packed word per row (32 x 2) for the bit-packed form
    for nx, ny, gens in [(2, 2, 3), (3, 3, 4), (32, 2, 5), (2, 40, 6)]:
        a = np.zeros((ny + 2, nx + 2), np.uint8)
        oracle.fill_bernoulli(a, 1, 31 + nx + ny, 0.5)
        d0, d1 = ctx.from_numpy(a), ctx.array(a.shape, hip.U8)
        ctx.gol(d0, d1, nx, ny, gens, variant)
        ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens)
        assert np.array_equal((d0, d1)[gens % 2].download(), ref), (nx, ny, gens)
    # FDTD3d with a one-point interior and with one-point-thin slabs
    for n in [(1, 1, 1), (2, 1, 9), (1, 7, 2)]:
        nx, ny, nz = n
        a = np.zeros((nz + 8, ny + 8, nx + 8), np.float32)
        oracle.fill_uniform(a, 4, 41 + nz)
        d0, d1 = ctx.from_numpy(a, hip.F32), ctx.array(a.shape, hip.F32)
        ctx.fdtd3d(d0, d1, nx, ny, nz, FDTD_COEFF, 3, variant)
        ref = oracle.fdtd3d(a.copy(), np.zeros_like(a), nx, ny, nz, FDTD_COEFF, 3)
        assert np.array_equal(d1.download().view(np.uint8), ref.view(np.uint8)), n


def test_himeno_smallest_grid(ctx):
    # 3 x 3 x 3 including the boundary shell: one interior point
    shape = (3, 3, 3)
    arrs = [np.zeros(shape) for _ in range(14)]
    oracle.fill_uniform(arrs[0], 0, 51)
    arrs[1][:] = arrs[0]
    for idx, v in enumerate([1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]):
        arrs[2 + idx][:] = v
    dev = [ctx.from_numpy(a, hip.F64) for a in arrs]
    gosa = ctx.himeno(dev, 3, 3, 3, 0.8, 3)
    ref = [a.copy() for a in arrs]
    ref_gosa = oracle.himeno(ref, 3, 3, 3, 0.8, 3)
    assert np.array_equal(dev[1].download(), ref[1])
    assert np.allclose(gosa, ref_gosa, rtol=1e-12, atol=0)


def test_invalid_arguments_are_refused(ctx):
    h = ctx.array((6, 6), hip.F64)
    h2 = ctx.array((6, 6), hip.F64)
    g = ctx.array((6, 6), hip.U8)
    g2 = ctx.array((6, 6), hip.U8)
    f = ctx.array((9, 9, 9), hip.F32)
    f2 = ctx.array((9, 9, 9), hip.F32)
    refused = [
        ("heat empty interior", lambda: ctx.heat2d(h, h2, 0, 4, 0.1, 1)),
        ("heat negative iters", lambda: ctx.heat2d(h, h2, 4, 4, 0.1, -1)),
        ("heat one buffer twice", lambda: ctx.heat2d(h, h, 4, 4, 0.1, 1)),
        ("heat beyond int index", lambda: ctx.heat2d(h, h2, 50000, 50000, 0.1, 1)),
        ("heat unknown candidate", lambda: ctx.heat2d(h, h2, 4, 4, 0.1, 1, hip.CANDIDATE0 + 999)),
        ("gol below 2 x 2", lambda: ctx.gol(g, g2, 1, 4, 1)),
        ("gol negative gens", lambda: ctx.gol(g, g2, 4, 4, -2)),
        ("gol beyond int index", lambda: ctx.gol(g, g2, 60000, 60000, 1)),
        ("gol unknown candidate", lambda: ctx.gol(g, g2, 4, 4, 1, hip.CANDIDATE0 + 999)),
        ("fdtd empty interior", lambda: ctx.fdtd3d(f, f2, 1, 1, 0, FDTD_COEFF, 1)),
        ("fdtd beyond int index", lambda: ctx.fdtd3d(f, f2, 2000, 2000, 2000, FDTD_COEFF, 1)),
        ("fdtd unknown candidate", lambda: ctx.fdtd3d(f, f2, 1, 1, 1, FDTD_COEFF, 1, hip.CANDIDATE0 + 999)),
        ("heat ensemble of no grids", lambda: ctx.heat2d_ensemble([], 0.1, 1)),
    ]
    for what, call in refused:
        with pytest.raises(hip.AppfisError):
            call()
            pytest.fail(what)
    # the context still runs work correctly after the refusals
    a = np.zeros((6, 6))
    oracle.fill_uniform(a, 1, 61)
    h.upload(a)
    ctx.heat2d(h, h2, 4, 4, 0.1, 3)
    assert np.array_equal(h2.download(), oracle.heat2d(a.copy(), np.zeros_like(a), 4, 4, 0.1, 3))

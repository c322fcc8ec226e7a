"""Pins of the CPU oracle (oracle/liborc.so) -- CPU only.

Every oracle function is checked against something that is not a retyping of its own
formula: the reference's own serial programs (golden fixtures written by
tests/golden/make_ref_golden.py), published known-answer vectors, independent
library routines (numpy's IEEE binary16 cast, torch's bfloat16 cast, numpy padding
modes, scipy convolution), closed forms and invariants chosen so that a dropped term,
a wrong sign/index or a transposed operand fails at least one check.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_serial.json")


# ---------------------------------------------------------------------------------
# seeded inputs
# ---------------------------------------------------------------------------------
def test_splitmix64_known_answers():
    # SplitMix64 reference outputs (Vigna's splitmix64.c) for seed 1234567 and seed 0
    assert [oracle.rng_u64(1234567, n) for n in range(3)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423]
    assert oracle.rng_u64(0, 0) == 0xE220A8397B1DCDAF


def test_fill_uniform_and_bernoulli_definitions():
    a = np.zeros((6, 7), dtype=np.float64)
    oracle.fill_uniform(a, 1, 1001, 0.0, 1.0)
    assert np.all(a[0] == 0) and np.all(a[:, 0] == 0)  # ghost untouched
    # interior cell (i, j) = (2, 3) -> interior linear index i + nx*j with nx = 5
    r = oracle.rng_u64(1001, 2 + 5 * 3)
    assert a[3 + 1, 2 + 1] == (r >> 11) * 2.0 ** -53
    f = np.zeros((6, 7), dtype=np.float32)
    oracle.fill_uniform(f, 1, 1001, 0.0, 1.0)
    assert f[4, 3] == np.float32((r >> 40) * 2.0 ** -24)
    g = np.zeros((34, 34), dtype=np.uint8)
    oracle.fill_bernoulli(g, 1, 7, 0.25)
    exp = np.array([[(oracle.rng_u64(7, i + 32 * j) >> 62) == 0 for i in range(32)] for j in range(32)])
    assert np.array_equal(g[1:-1, 1:-1].astype(bool), exp)


# ---------------------------------------------------------------------------------
# storage conversions: against numpy's float16 and torch's bfloat16 casts
# ---------------------------------------------------------------------------------
def _sample_f32_bits(n=200000, seed=3):
    rs = np.random.RandomState(seed)
    bits = rs.randint(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    special = np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x477FEFFF, 0x477FF000, 0x477FF001,
                        0x38800000, 0x387FFFFF, 0x33800000, 0x33000000, 0x33000001, 0x3F800000,
                        0x3F801000, 0x3F802000, 0x3F803000], dtype=np.uint32)
    # also dense coverage of the binary16 range
    dense = (rs.randint(0x33000000, 0x47800000, size=n, dtype=np.int64)).astype(np.uint32)
    return np.concatenate([bits, special, dense])


def test_f16_conversion_matches_numpy():
    bits = _sample_f32_bits()
    vals = bits.view(np.float32)
    vals = vals[~np.isnan(vals)]
    ours = np.array([oracle.f32_to_f16_bits(float(v)) for v in vals[:60000]], dtype=np.uint16)
    ref = vals[:60000].astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)


def test_bf16_conversion_matches_torch():
    torch = pytest.importorskip("torch")
    bits = _sample_f32_bits()
    vals = bits.view(np.float32)
    vals = vals[~np.isnan(vals)][:60000]
    ours = np.array([oracle.f32_to_bf16_bits(float(v)) for v in vals], dtype=np.uint16)
    ref = torch.from_numpy(vals.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, ref)


# ---------------------------------------------------------------------------------
# Grid::SetBoundary / SetBoundaryCopy / periodic Exchange
# ---------------------------------------------------------------------------------
def test_set_boundary_matches_slices():
    a = np.arange(7 * 8 * 9, dtype=np.float64).reshape(7, 8, 9)
    b = a.copy()
    oracle.set_boundary(b, 2, -1, -1.5)
    exp = a.copy()
    exp[:2] = exp[-2:] = -1.5
    exp[:, :2] = exp[:, -2:] = -1.5
    exp[:, :, :2] = exp[:, :, -2:] = -1.5
    assert np.array_equal(b, exp)
    # single 3D faces follow APPFIS::DIRECTION: DOWN = z low, FRONT = y low, RIGHT = x high
    c = a.copy()
    oracle.set_boundary(c, 1, 3, 0.0)
    assert np.all(c[0] == 0) and np.array_equal(c[1:], a[1:])
    c = a.copy()
    oracle.set_boundary(c, 1, 4, 0.0)
    assert np.all(c[:, 0] == 0) and np.array_equal(c[:, 1:], a[:, 1:])
    c = a.copy()
    oracle.set_boundary(c, 1, 1, 0.0)
    assert np.all(c[:, :, -1] == 0) and np.array_equal(c[:, :, :-1], a[:, :, :-1])


def test_set_boundary_copy_ghost1_is_edge_padding():
    rs = np.random.RandomState(0)
    inner = rs.rand(5, 6)
    a = np.zeros((7, 8))
    a[1:-1, 1:-1] = inner
    oracle.set_boundary_copy(a, 1, -1)
    assert np.array_equal(a, np.pad(inner, 1, mode="edge"))


def test_set_boundary_copy_ghost2_reference_order():
    # Grid.hpp:682-830 walks g from ghost-1 down to 0; for the high face layer
    # dim-ghost+g is copied from dim-ghost+g-1, so with ghost 2 the outermost cell gets
    # the OLD inner ghost value, then the inner ghost gets the last interior value.
    a = np.array([10.0, 11.0, 1.0, 2.0, 3.0, 4.0, 20.0, 21.0])
    oracle.set_boundary_copy(a, 2, 1)  # RIGHT
    assert list(a) == [10.0, 11.0, 1.0, 2.0, 3.0, 4.0, 4.0, 20.0]
    b = np.array([10.0, 11.0, 1.0, 2.0, 3.0, 4.0, 20.0, 21.0])
    oracle.set_boundary_copy(b, 2, 0)  # LEFT: g=1 <- 2, then g=0 <- 1 (already copied)
    assert list(b) == [1.0, 1.0, 1.0, 2.0, 3.0, 4.0, 20.0, 21.0]


@pytest.mark.parametrize("shape,g", [((9, 10), 1), ((12, 11), 2), ((8, 9, 10), 1), ((10, 12, 9), 2), ((20,), 3)])
def test_periodic_exchange_is_wrap_padding(shape, g):
    rs = np.random.RandomState(1)
    inner = rs.randint(0, 255, size=tuple(s - 2 * g for s in shape)).astype(np.int32)
    a = np.zeros(shape, dtype=np.int32)
    a[tuple(slice(g, s - g) for s in shape)] = inner
    oracle.exchange(a, g, True)
    assert np.array_equal(a, np.pad(inner, g, mode="wrap"))
    b = np.zeros(shape, dtype=np.int32)
    b[tuple(slice(g, s - g) for s in shape)] = inner
    before = b.copy()
    oracle.exchange(b, g, False)  # MPI_PROC_NULL neighbours: nothing moves
    assert np.array_equal(b, before)


# ---------------------------------------------------------------------------------
# Execute*Reduce
# ---------------------------------------------------------------------------------
def test_reduce_against_numpy():
    rs = np.random.RandomState(2)
    a = (rs.rand(9, 10, 11) - 0.5).astype(np.float64)
    box = [1, 10, 2, 9, 1, 8]
    sub = a[1:8, 2:9, 1:10]
    assert oracle.reduce(a, box, oracle.MAX) == sub.max()
    assert oracle.reduce(a, box, oracle.MIN) == sub.min()
    assert abs(oracle.reduce(a, box, oracle.SUM) - sub.sum()) < 1e-12
    assert abs(oracle.reduce_sum_wide(a, box) - math.fsum(sub.ravel())) < 1e-15
    u = rs.randint(0, 2, size=(40, 50)).astype(np.uint8)
    assert oracle.reduce_sum_wide(u, [1, 49, 1, 39]) == int(u[1:39, 1:49].sum())


def test_reduce_max_starts_at_numeric_limits_min():
    # Runtime.hpp:696 initialises MAX with numeric_limits<T>::min(): for floating types
    # that is the smallest positive normal, so an all-negative field reports it.
    a = -np.ones((5, 5), dtype=np.float32)
    assert oracle.reduce(a, [1, 4, 1, 4], oracle.MAX) == np.finfo(np.float32).tiny
    d = -np.ones((5, 5), dtype=np.float64)
    assert oracle.reduce(d, [1, 4, 1, 4], oracle.MAX) == np.finfo(np.float64).tiny


# ---------------------------------------------------------------------------------
# Heat2D config form: closed-form eigenmode and maximum principle
# ---------------------------------------------------------------------------------
def test_heat2d_eigenmode_closed_form():
    n, c, iters, p, q = 40, 0.1, 60, 2, 3
    i = np.arange(n + 2)
    mode = np.outer(np.sin(q * np.pi * i / (n + 1)), np.sin(p * np.pi * i / (n + 1)))
    b0 = mode.copy()
    b0[0, :] = b0[-1, :] = b0[:, 0] = b0[:, -1] = 0.0
    b1 = np.zeros_like(b0)
    out = oracle.heat2d(b0, b1, n, n, c, iters)
    lam = 1 - 4 * c * (np.sin(p * np.pi / (2 * (n + 1))) ** 2 + np.sin(q * np.pi / (2 * (n + 1))) ** 2)
    assert np.max(np.abs(out - lam ** iters * np.pad(mode[1:-1, 1:-1], 1))) < 1e-13


def test_heat2d_maximum_principle_and_f32():
    a = np.zeros((34, 34))
    oracle.fill_uniform(a, 1, 1001)
    b = np.zeros_like(a)
    out = oracle.heat2d(a.copy(), b, 32, 32, 0.1, 30)
    assert out.min() >= 0.0 and out.max() <= a.max()
    af = a.astype(np.float32)
    outf = oracle.heat2d(af.copy(), np.zeros_like(af), 32, 32, 0.1, 30)
    assert np.max(np.abs(outf - out)) < 1e-5


# ---------------------------------------------------------------------------------
# GameOfLife: reference serial program (golden), patterns, independent convolution
# ---------------------------------------------------------------------------------
def _libc_rand_init(nx, ny):
    libc = ctypes.CDLL(None)
    libc.srand(0)
    vals = np.array([libc.rand() % 2 for _ in range(nx * ny)], dtype=np.int32).reshape(ny, nx)
    return vals


def test_gol_reference_rule_matches_reference_serial_program():
    golden = json.load(open(GOLDEN))
    for case in golden["gol"]:
        nx, ny, it = case["nx"], case["ny"], case["iterations"]
        a = np.zeros((ny + 2, nx + 2), dtype=np.int32)
        a[1:-1, 1:-1] = _libc_rand_init(nx, ny)
        b = np.zeros_like(a)
        out = oracle.gol_ref_i32(a, b, nx, ny, it, True)
        assert int(out[1:-1, 1:-1].sum()) == case["live_count"], case


def _life_conv(g):
    from scipy.signal import convolve2d

    k = np.ones((3, 3), dtype=np.int32)
    k[1, 1] = 0
    n = convolve2d(g.astype(np.int32), k, mode="same", boundary="wrap")
    return ((n == 3) | ((n == 2) & (g == 1))).astype(np.uint8)


def _run_gol_u8(inner, gens):
    ny, nx = inner.shape
    a = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
    a[1:-1, 1:-1] = inner
    b = np.zeros_like(a)
    return oracle.gol_u8(a, b, nx, ny, gens)


def test_gol_b3s23_against_scipy_convolution():
    rs = np.random.RandomState(5)
    g = (rs.rand(37, 64) < 0.3).astype(np.uint8)
    ref = g.copy()
    for _ in range(25):
        ref = _life_conv(ref)
    out = _run_gol_u8(g, 25)
    assert np.array_equal(out[1:-1, 1:-1], ref)
    assert np.array_equal(out, np.pad(out[1:-1, 1:-1], 1, mode="wrap"))  # ring exchanged


def test_gol_patterns():
    blinker = np.zeros((8, 8), dtype=np.uint8)
    blinker[3, 2:5] = 1
    assert np.array_equal(_run_gol_u8(blinker, 2)[1:-1, 1:-1], blinker)
    assert not np.array_equal(_run_gol_u8(blinker, 1)[1:-1, 1:-1], blinker)
    block = np.zeros((6, 6), dtype=np.uint8)
    block[0:2, 5] = 1
    block[0:2, 0] = 1  # block split across the periodic seam
    assert np.array_equal(_run_gol_u8(block, 7)[1:-1, 1:-1], block)
    n = 12
    glider = np.zeros((n, n), dtype=np.uint8)
    glider[0, 1] = glider[1, 2] = glider[2, 0] = glider[2, 1] = glider[2, 2] = 1
    assert np.array_equal(_run_gol_u8(glider, 4 * n)[1:-1, 1:-1], glider)


# ---------------------------------------------------------------------------------
# Heat3D reference functor vs the reference serial program
# ---------------------------------------------------------------------------------
def _heat3d_gauss(nx, ny, nz):
    def var(n):
        m = (n - 1) / 2.0
        return sum((i - m) ** 2 for i in range(n)) / (n - 1), m

    vx, mx = var(nx)
    vy, my = var(ny)
    vz, mz = var(nz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.exp(-((i - mx) ** 2 / (2 * vx) + (j - my) ** 2 / (2 * vy) + (k - mz) ** 2 / (2 * vz)))


def test_heat3d_reference_functor_matches_reference_serial_program():
    golden = json.load(open(GOLDEN))
    cxyz = (0.1 * 0.01) / 0.1 ** 2
    for case in golden["heat3d"]:
        nx, ny, nz, it = case["nx"], case["ny"], case["nz"], case["iterations"]
        a = np.zeros((nz + 2, ny + 2, nx + 2))
        a[1:-1, 1:-1, 1:-1] = _heat3d_gauss(nx, ny, nz)
        out = oracle.heat3d_ref(a, np.zeros_like(a), nx, ny, nz, cxyz, cxyz, cxyz, it)
        inner = out[1:-1, 1:-1, 1:-1]
        assert float(f"{inner.max():.6g}") == case["max"]
        assert float(f"{inner.min():.6g}") == case["min"]


# ---------------------------------------------------------------------------------
# FDTD3d config form: scipy convolution, growth factor, linearity, axis symmetry
# ---------------------------------------------------------------------------------
COEFF = [0.5, 0.125, 0.0625, 0.125 / 3, 0.03125]


def _star_kernel(coeff):
    k = np.zeros((9, 9, 9))
    k[4, 4, 4] = coeff[0]
    for r in range(1, 5):
        for s in (-r, r):
            k[4 + s, 4, 4] = k[4, 4 + s, 4] = k[4, 4, 4 + s] = coeff[r]
    return k


def _fdtd_field(n, seed, dtype=np.float64):
    a = np.zeros((n + 8,) * 3, dtype=dtype)
    oracle.fill_uniform(a, 4, seed)
    return a


def test_fdtd3d_against_scipy_convolution():
    from scipy.ndimage import convolve

    n, steps = 14, 3
    a = _fdtd_field(n, 1004)
    ref = a.copy()
    for _ in range(steps):
        full = convolve(ref, _star_kernel(COEFF), mode="constant", cval=0.0)
        ref = np.zeros_like(ref)
        ref[4:-4, 4:-4, 4:-4] = full[4:-4, 4:-4, 4:-4]
    out = oracle.fdtd3d(a, np.zeros_like(a), n, n, n, COEFF, steps)
    assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < 1e-13
    # an index slip (c2 and c3 swapped) must be visible
    bad = list(COEFF)
    bad[2], bad[3] = bad[3], bad[2]
    out_bad = oracle.fdtd3d(_fdtd_field(n, 1004), np.zeros_like(a), n, n, n, bad, steps)
    assert np.max(np.abs(out_bad - ref)) / np.max(np.abs(ref)) > 1e-3


def test_fdtd3d_growth_linearity_symmetry():
    n = 24
    a = np.zeros((n + 8,) * 3, dtype=np.float64)
    a[4:-4, 4:-4, 4:-4] = 1.0
    out = oracle.fdtd3d(a, np.zeros_like(a), n, n, n, COEFF, 2)
    centre = out[4 + 8:-4 - 8, 4 + 8:-4 - 8, 4 + 8:-4 - 8]
    assert np.allclose(centre, 2.0625 ** 2, rtol=1e-14, atol=0)
    x = _fdtd_field(12, 11)
    y = _fdtd_field(12, 12)
    ox = oracle.fdtd3d(x.copy(), np.zeros_like(x), 12, 12, 12, COEFF, 2)
    oy = oracle.fdtd3d(y.copy(), np.zeros_like(x), 12, 12, 12, COEFF, 2)
    oxy = oracle.fdtd3d(2 * x + 3 * y, np.zeros_like(x), 12, 12, 12, COEFF, 2)
    assert np.max(np.abs(oxy - (2 * ox + 3 * oy))) < 1e-12
    t = np.ascontiguousarray(x.transpose(2, 0, 1))
    ot = oracle.fdtd3d(t, np.zeros_like(t), 12, 12, 12, COEFF, 2)
    assert np.max(np.abs(ot - ox.transpose(2, 0, 1))) < 1e-14


def test_fdtd3d_reference_functor_growth():
    # fdtd3d_appfis.hpp:129-143: coeff * (u + 24 neighbours): a constant core grows by 2.5
    n = 20
    a = np.zeros((n + 8,) * 3, dtype=np.float32)
    a[4:-4, 4:-4, 4:-4] = 1.0
    out = oracle.fdtd3d_ref(a, np.zeros_like(a), n, n, n, 0.1, 4, 1)
    assert np.allclose(out[8:-8, 8:-8, 8:-8], 2.5, rtol=1e-6)
    assert out[4, 4, 4] < 2.5  # corner of the interior sees the zero ghost


def test_fdtd3d_half_storage_is_rounded_f32_arithmetic():
    n = 10
    a = _fdtd_field(n, 21, np.float32)
    h = a.astype(np.float16)
    out16 = oracle.fdtd3d(h.copy(), np.zeros_like(h), n, n, n, COEFF, 1)
    out32 = oracle.fdtd3d(h.astype(np.float32), np.zeros((n + 8,) * 3, np.float32), n, n, n, COEFF, 1)
    assert np.array_equal(out16, out32.astype(np.float16))


# ---------------------------------------------------------------------------------
# Himeno: closed forms for the axis terms and each cross term
# ---------------------------------------------------------------------------------
def _himeno_arrays(nx, ny, nz, p, dtype=np.float64, a=(1, 1, 1, 1 / 6), b=(0, 0, 0), c=(1, 1, 1), wrk1=0.0):
    shape = (nz, ny, nx)
    arrs = [p.astype(dtype).copy(), p.astype(dtype).copy()]
    for v in list(a) + list(b) + list(c) + [1.0, wrk1]:
        arrs.append(np.full(shape, v, dtype=dtype))
    return arrs


def test_himeno_quadratic_closed_form():
    nx, ny, nz = 9, 10, 17
    k = np.arange(nz).reshape(nz, 1, 1) * np.ones((1, ny, nx))
    p = (k / (nz - 1)) ** 2
    arrs = _himeno_arrays(nx, ny, nz, p)
    gosa = oracle.himeno(arrs, nx, ny, nz, 0.8, 1)
    ss = 1.0 / (3 * (nz - 1) ** 2)
    exp = (nx - 2) * (ny - 2) * (nz - 2) * ss * ss
    assert abs(gosa[0] - exp) / exp < 1e-10
    assert np.allclose(arrs[1][1:-1, 1:-1, 1:-1], p[1:-1, 1:-1, 1:-1] + 0.8 * ss, rtol=0, atol=1e-15)


def test_himeno_linear_field_is_fixed_point():
    nx, ny, nz = 8, 8, 12
    k = np.arange(nz).reshape(nz, 1, 1) * np.ones((1, ny, nx))
    arrs = _himeno_arrays(nx, ny, nz, k / (nz - 1))
    gosa = oracle.himeno(arrs, nx, ny, nz, 0.8, 3)
    assert np.all(gosa < 1e-28)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_himeno_axis_orientation(axis):
    # a_axis multiplies the forward neighbour p(+1) and c_axis the backward one p(-1)
    # (himeno_appfis.hpp:184-193).  With p = the coordinate along `axis`, a_axis = 2,
    # c_axis = 0 and the other a, c = 1, a3 = 1/6: s0 = 2(x+1) + 4x = 6x + 2, so
    # ss = 1/3 and the new p = x + omega/3 everywhere inside; the transposed operands
    # (a with p(-1), c with p(+1)) would give ss = -1/3, and a wrong axis ss = 0.
    nx, ny, nz = 7, 8, 9
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = [i, j, k][axis].astype(np.float64)
    a = [1.0, 1.0, 1.0, 1.0 / 6.0]
    c = [1.0, 1.0, 1.0]
    a[axis], c[axis] = 2.0, 0.0
    arrs = _himeno_arrays(nx, ny, nz, p, a=a, c=c)
    omega = 0.8
    gosa = oracle.himeno(arrs, nx, ny, nz, omega, 1)
    inner = (slice(1, -1),) * 3
    assert np.allclose(arrs[1][inner], p[inner] + omega / 3.0, rtol=0, atol=1e-13)
    assert gosa[0] == pytest.approx((nx - 2) * (ny - 2) * (nz - 2) / 9.0, rel=1e-12)


@pytest.mark.parametrize("term", [0, 1, 2])
def test_himeno_cross_terms(term):
    # p = product of two coordinates makes exactly one of the three corner terms equal
    # to 4 (b0: x*y, b1: y*z, b2: x*z) while the axis terms and the other corners see a
    # linear function; with a = 0, c = 0, b_term = 1, a3 = 1: ss = 4 - p.
    nx, ny, nz = 7, 8, 9
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = [i * j, j * k, i * k][term].astype(np.float64)
    b = [0.0, 0.0, 0.0]
    b[term] = 1.0
    arrs = _himeno_arrays(nx, ny, nz, p, a=(0, 0, 0, 1.0), b=b, c=(0, 0, 0))
    gosa = oracle.himeno(arrs, nx, ny, nz, 1.0, 1)
    inner = (slice(1, -1),) * 3
    ss = 4.0 - p[inner]
    assert abs(gosa[0] - np.sum(ss * ss)) < 1e-9 * np.sum(ss * ss)
    assert np.allclose(arrs[1][inner], 4.0)

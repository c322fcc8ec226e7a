"""GPU parity of the hand-written config kernels against the oracle.

Heat2D, GameOfLife and FDTD3d are integer work or fixed-order floating point with no
FMA contraction on either side, so the bar is bit-exact: the PLAIN and CHOSEN forms
must reproduce the oracle's bytes.  Full config sizes are covered through the oracle
for a few sweeps plus size-independent properties for the full sweep count.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from appfis_amd import hip

pytestmark = pytest.mark.gpu

FDTD_COEFF = [0.5, 0.125, 0.0625, 0.125 / 3, 0.03125]


def _heat_pair(ctx, nx, ny, dt, seed=1001):
    a = np.zeros((ny + 2, nx + 2), dtype=np.float64 if dt == hip.F64 else np.float32)
    oracle.fill_uniform(a, 1, seed)
    return a, ctx.from_numpy(a, dt), ctx.array(a.shape, dt)


@pytest.mark.parametrize("dt", [hip.F64, hip.F32])
@pytest.mark.parametrize("nx,ny,iters", [(37, 29, 11), (64, 64, 16), (130, 67, 9), (1, 5, 3), (300, 250, 37)])
@pytest.mark.parametrize("variant", [hip.PLAIN, hip.CHOSEN])
def test_heat2d_bit_exact(ctx, dt, nx, ny, iters, variant):
    a, d0, d1 = _heat_pair(ctx, nx, ny, dt)
    ctx.heat2d(d0, d1, nx, ny, 0.1, iters, variant)
    ref = oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, iters)
    got = (d0, d1)[iters % 2].download()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("tsteps", [1, 3, 4, 8, 11, 16])
def test_heat2d_chosen_depths(ctx, tsteps):
    nx, ny, iters = 200, 150, 23
    a, d0, d1 = _heat_pair(ctx, nx, ny, hip.F64)
    ctx.heat2d(d0, d1, nx, ny, 0.1, iters, hip.CHOSEN, tsteps)
    ref = oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, iters)
    assert np.array_equal((d0, d1)[iters % 2].download(), ref)


@pytest.mark.parametrize("dt", [hip.F64, hip.F32])
@pytest.mark.parametrize("nx,ny,iters", [(200, 150, 37), (5, 300, 19), (1030, 9, 21)])
def test_heat2d_every_candidate_bit_exact(ctx, dt, nx, ny, iters):
    # every temporal-blocking candidate the autotuner may pick, forced one at a time
    # (partial tiles in both axes; a grid narrower and one shorter than any tile)
    a, d0, d1 = _heat_pair(ctx, nx, ny, dt)
    ref = oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, iters)
    names = hip.heat2d_candidates(dt)
    assert len(names) >= 2
    for i, name in enumerate(names):
        d0.upload(a)
        ctx.heat2d(d0, d1, nx, ny, 0.1, iters, hip.CANDIDATE0 + i)
        assert np.array_equal((d0, d1)[iters % 2].download(), ref), name


@pytest.mark.parametrize("dt", [hip.F64, hip.F32])
@pytest.mark.parametrize("iters", [1, 8, 23, 64])
def test_heat2d_ensemble_matches_oracle(ctx, dt, iters):
    # uneven grids advanced together (one launch per chunk over all grids, graph-replayed)
    sizes = [(1, 5), (37, 29), (130, 67), (300, 250), (64, 64), (2, 1), (513, 31)]
    grids, refs = [], []
    for k, (nx, ny) in enumerate(sizes):
        a, d0, d1 = _heat_pair(ctx, nx, ny, dt, seed=2000 + k)
        grids.append((d0, d1, nx, ny))
        refs.append(oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, iters))
    ctx.heat2d_ensemble(grids, 0.1, iters)
    for (d0, d1, nx, ny), ref in zip(grids, refs):
        assert np.array_equal((d0, d1)[iters % 2].download(), ref), (nx, ny)


def test_heat2d_ensemble_repeated_calls(ctx):
    # each call captures and replays its launch chain as a graph; three calls in a row on
    # the same grids (eager reductions between them) continue the same trajectory
    sizes = [(40, 33), (129, 70), (7, 200)]
    grids, arrs = [], []
    for k, (nx, ny) in enumerate(sizes):
        a, d0, d1 = _heat_pair(ctx, nx, ny, hip.F64, seed=2100 + k)
        grids.append((d0, d1, nx, ny))
        arrs.append(a)
    done, prev = 0, 0
    for call, iters in enumerate((9, 16, 7)):
        if prev % 2:  # the last call left the field in buffer 1: bring it to buffer 0
            for d0, d1, _, _ in grids:
                d0.upload(d1.download())
        ctx.heat2d_ensemble(grids, 0.1, iters)
        done += iters
        prev = iters
        for (d0, d1, nx, ny), a in zip(grids, arrs):
            ref = oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, done)
            got = (d0, d1)[iters % 2].download()
            assert np.array_equal(got, ref), (call, nx, ny)
            ctx.reduce_sum_wide(d0, [1, nx + 1, 1, ny + 1])


def test_heat2d_ensemble_graph_reuse():
    # the launch chain is captured on the first call and replayed by later calls with
    # the same shapes and sweep count: a second set of grids (new buffers, new values)
    # must be advanced, not the first; after the context workspace is reallocated (a
    # flood run in between, on a fresh context whose workspace only held the heat
    # table) the chain is captured again
    sizes = [(40, 33), (129, 70), (7, 200)]
    iters = 16
    keep = []

    def one_set(c, seed):
        grids, arrs = [], []
        for k, (nx, ny) in enumerate(sizes):
            a, d0, d1 = _heat_pair(c, nx, ny, hip.F64, seed=seed + k)
            grids.append((d0, d1, nx, ny))
            arrs.append(a)
            keep.extend((d0, d1))
        return grids, arrs

    def check(c, grids, arrs, tag):
        c.heat2d_ensemble(grids, 0.1, iters)
        for (d0, d1, nx, ny), a in zip(grids, arrs):
            ref = oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, iters)
            assert np.array_equal((d0, d1)[iters % 2].download(), ref), (tag, nx, ny)

    with hip.Context(0) as c:
        check(c, *one_set(c, 2200), "first")
        check(c, *one_set(c, 2300), "same shapes, new buffers")
        # grow the workspace: a flood run keeps three field copies of its grid there
        n = 300
        z = oracle.flood_dem(n, n, 10.0, 2400)
        zero = np.zeros_like(z)
        d = [c.from_numpy(x, hip.F64) for x in (zero, zero, zero, z, np.full_like(z, 0.035 ** 2))]
        keep.extend(d)
        c.flood_run(d[0], d[1], d[2], d[3], d[4], n, n, hip.FloodParams(10.0, 0.5, 1e-3, 1.0, 0.0, 1e-5, 0.0), 2,
                    hip.PLAIN)
        check(c, *one_set(c, 2500), "after the workspace grew")
        for x in keep:
            x.free()


def test_heat2d_full_config(ctx):
    # 1024^2 FP64, 500 sweeps: the whole config against the oracle
    nx = ny = 1024
    a, d0, d1 = _heat_pair(ctx, nx, ny, hip.F64)
    ref = oracle.heat2d(a.copy(), np.zeros_like(a), nx, ny, 0.1, 500)
    for variant in (hip.PLAIN, hip.CHOSEN):
        d0.upload(a)
        ctx.heat2d(d0, d1, nx, ny, 0.1, 500, variant)
        got = d0.download()
        assert np.array_equal(got, ref), variant
    print(f"[heat2d] checksum {ref.sum():.17g}")


def _gol_pair(ctx, nx, ny, seed=1002, p=0.25):
    a = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
    oracle.fill_bernoulli(a, 1, seed, p)
    return a, ctx.from_numpy(a), ctx.array(a.shape, hip.U8)


@pytest.mark.parametrize("nx,ny,gens", [(14, 9, 5), (62, 40, 13), (64, 64, 10), (94, 33, 7), (126, 77, 21),
                                        (253, 101, 6), (1000, 19, 4), (510, 257, 31)])
def test_gol_plain_bit_exact(ctx, nx, ny, gens):
    # widths cover W = nx + 2 = 0 .. 15 mod 16 across the cases
    a, d0, d1 = _gol_pair(ctx, nx, ny)
    ctx.gol(d0, d1, nx, ny, gens, hip.PLAIN)
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens)
    assert np.array_equal((d0, d1)[gens % 2].download(), ref)


@pytest.mark.parametrize("nx", list(range(14, 30)))
def test_gol_plain_all_residues(ctx, nx):
    ny, gens = 23, 9
    a, d0, d1 = _gol_pair(ctx, nx, ny, seed=nx)
    ctx.gol(d0, d1, nx, ny, gens, hip.PLAIN)
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens)
    assert np.array_equal((d0, d1)[gens % 2].download(), ref)
    # the other buffer holds generation gens-1 with its ring exchanged
    prev = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens - 1)
    assert np.array_equal((d0, d1)[(gens + 1) % 2].download(), prev)


def test_gol_full_config_prefix(ctx):
    # 8192^2 config: the first 12 generations against the oracle bit for bit
    nx = ny = 8192
    a, d0, d1 = _gol_pair(ctx, nx, ny)
    ctx.gol(d0, d1, nx, ny, 12, hip.PLAIN)
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, 12)
    assert np.array_equal(d0.download(), ref)


@pytest.mark.parametrize("nx,ny,gens", [(32, 2, 3), (32, 9, 1), (64, 64, 10), (96, 7, 13), (160, 33, 21),
                                        (256, 100, 2), (2048, 17, 5), (94, 33, 7), (50, 40, 6),
                                        (2048, 3, 40), (2048, 130, 37), (4096, 50, 33), (16384, 20, 19),
                                        (8192, 9, 2), (2048, 64, 1)])
def test_gol_chosen_both_buffers(ctx, nx, ny, gens):
    # bit-packed form (nx % 32 == 0; other widths take the byte kernel; nx % 2048 == 0
    # also runs the temporally blocked kernel): buf[gens % 2] holds generation gens and
    # the other buffer generation gens-1, rings exchanged
    a, d0, d1 = _gol_pair(ctx, nx, ny, seed=nx * 7 + ny)
    ctx.gol(d0, d1, nx, ny, gens, hip.CHOSEN)
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens)
    assert np.array_equal((d0, d1)[gens % 2].download(), ref)
    prev = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens - 1)
    assert np.array_equal((d0, d1)[(gens + 1) % 2].download(), prev)


@pytest.mark.parametrize("nx,ny,gens", [(2048, 45, 17), (4096, 30, 9), (8192, 21, 6),
                                        (8192, 32, 9), (8192, 64, 20), (8192, 96, 27)])
def test_gol_every_blocked_candidate(ctx, nx, ny, gens):
    # every blocked kernel of the bit-packed form, forced one at a time (candidates whose
    # block does not span nx refuse with an error); the 8192-cell rows with ny of one,
    # two and three 32-row bands put a band's periodic neighbours on itself or on the
    # same other band
    a, d0, d1 = _gol_pair(ctx, nx, ny, seed=nx + gens)
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens)
    prev = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens - 1)
    ran = []
    for i, name in enumerate(hip.gol_candidates()):
        d0.upload(a)
        try:
            ctx.gol(d0, d1, nx, ny, gens, hip.CANDIDATE0 + i)
        except hip.AppfisError as e:
            assert "spanned" in str(e) or "needs nx" in str(e), e
            continue
        ran.append(name)
        assert np.array_equal((d0, d1)[gens % 2].download(), ref), name
        assert np.array_equal((d0, d1)[(gens + 1) % 2].download(), prev), name
    assert len(ran) >= 3


@pytest.mark.parametrize("gens", [2, 7, 8, 9, 16, 17, 24, 25])
def test_gol_full_row_launch_split(ctx, gens):
    # the full-row form runs a call's generations in at least two launches, the first
    # packing and the last unpacking the uint8 field (up to G = 8 generations: n-1 and 1;
    # beyond: G per launch and the rest in the last): generation counts around G and 2G,
    # both buffers against the oracle
    nx, ny = 2048, 40
    names = hip.gol_candidates()
    a, d0, d1 = _gol_pair(ctx, nx, ny, seed=gens + 500)
    ctx.gol(d0, d1, nx, ny, gens, hip.CANDIDATE0 + names.index("tw1 k3 g8 w16"))
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens)
    prev = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, gens - 1)
    assert np.array_equal((d0, d1)[gens % 2].download(), ref)
    assert np.array_equal((d0, d1)[(gens + 1) % 2].download(), prev)


def test_gol_chosen_glider_period(ctx):
    # a glider on an N x N torus returns to its start after 4N generations
    n = 64
    a = np.zeros((n + 2, n + 2), dtype=np.uint8)
    for x, y in [(2, 1), (3, 2), (1, 3), (2, 3), (3, 3)]:
        a[y + 1, x + 1] = 1
    d0, d1 = ctx.from_numpy(a), ctx.array(a.shape, hip.U8)
    ctx.gol(d0, d1, n, n, 4 * n, hip.CHOSEN)
    got = d0.download()
    assert np.array_equal(got[1:-1, 1:-1], a[1:-1, 1:-1])
    ctx.gol(d0, d1, n, n, 2 * n, hip.CHOSEN)
    assert not np.array_equal(d0.download()[1:-1, 1:-1], a[1:-1, 1:-1])


def test_gol_chosen_full_config_prefix(ctx):
    # 8192^2 config: 12 generations of the bit-packed form against the oracle, and
    # 100 generations against the byte kernel
    nx = ny = 8192
    a, d0, d1 = _gol_pair(ctx, nx, ny)
    ctx.gol(d0, d1, nx, ny, 12, hip.CHOSEN)
    ref = oracle.gol_u8(a.copy(), np.zeros_like(a), nx, ny, 12)
    assert np.array_equal(d0.download(), ref)
    del ref
    d0.upload(a)
    ctx.gol(d0, d1, nx, ny, 100, hip.CHOSEN)
    c0, c1 = d0.download(), d1.download()
    d0.upload(a)
    ctx.gol(d0, d1, nx, ny, 100, hip.PLAIN)
    assert np.array_equal(d0.download(), c0) and np.array_equal(d1.download(), c1)


def test_gol_full_config_every_100th_generation(ctx):
    # the 8192^2 config for 1000 generations, checked at every 100th generation against
    # the oracle: tests/golden/gol_8192.json holds the SHA-256 of the oracle's interior and
    # its live count at generations 100..1000 (tests/golden/make_gol_golden.py); both the
    # byte kernel and the bit-packed kernel must hit every digest, and agree on both
    # buffers (the other one holds the previous generation with its ring)
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gol_8192.json")))
    nx, ny = gold["nx"], gold["ny"]
    a = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
    oracle.fill_bernoulli(a, 1, gold["seed"], gold["p"])
    p0, p1 = ctx.from_numpy(a), ctx.array(a.shape, hip.U8)
    c0, c1 = ctx.from_numpy(a), ctx.array(a.shape, hip.U8)
    for rec in gold["records"]:
        ctx.gol(p0, p1, nx, ny, 100, hip.PLAIN)
        ctx.gol(c0, c1, nx, ny, 100, hip.CHOSEN)
        got_p, got_c = p0.download(), c0.download()
        gen = rec["generation"]
        for form, got in (("plain", got_p), ("chosen", got_c)):
            interior = np.ascontiguousarray(got[1:-1, 1:-1])
            assert int(interior.sum(dtype=np.int64)) == rec["live"], (form, gen)
            assert hashlib.sha256(interior.tobytes()).hexdigest() == rec["sha256"], (form, gen)
        assert np.array_equal(got_p, got_c), gen
        assert np.array_equal(p1.download(), c1.download()), gen
        print(f"[gol] generation {gen}: live {rec['live']} (both forms match the oracle)", flush=True)


def _fdtd_pair(ctx, n, dt, seed=1004):
    npdt = {hip.F32: np.float32, hip.F64: np.float64, hip.F16: np.float16, hip.BF16: np.uint16}[dt]
    shape = tuple(x + 8 for x in n[::-1])
    a = np.zeros(shape, dtype=npdt)
    oracle.fill_uniform(a, 4, seed, bf16=(dt == hip.BF16))
    return a, ctx.from_numpy(a, dt), ctx.array(shape, dt)


@pytest.mark.parametrize("variant", [hip.PLAIN, hip.CHOSEN])
@pytest.mark.parametrize("dt", [hip.F32, hip.F64, hip.F16, hip.BF16])
@pytest.mark.parametrize("n,steps", [((37, 29, 23), 4), ((64, 64, 64), 3), ((70, 17, 130), 2), ((130, 66, 70), 3)])
def test_fdtd3d_bit_exact(ctx, dt, n, steps, variant):
    nx, ny, nz = n
    coeff = FDTD_COEFF if dt not in (hip.F16,) else [c / 2.0625 for c in FDTD_COEFF]
    a, d0, d1 = _fdtd_pair(ctx, n, dt)
    ctx.fdtd3d(d0, d1, nx, ny, nz, coeff, steps, variant)
    ref = oracle.fdtd3d(a.copy(), np.zeros_like(a), nx, ny, nz, coeff, steps, bf16=(dt == hip.BF16))
    got = (d0, d1)[steps % 2].download()
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("dt", [hip.F32, hip.F64, hip.F16, hip.BF16])
@pytest.mark.parametrize("n", [(70, 33, 40), (68, 21, 37), (71, 30, 19), (68, 9, 259), (68, 1, 1), (68, 2, 3)])
def test_fdtd3d_every_candidate_bit_exact(ctx, dt, n):
    # every tiling the autotuner may pick, forced one at a time; nx = 70, 68, 71 covers
    # x points per lane (1, 2, 4) that do and do not divide nx, nz = 259 gives the
    # 64- to 256-plane z-chunks a last chunk of 3 planes (fewer than a prefetch ring),
    # and one plane of one row / three planes of two rows leave every prefetch short
    coeff = FDTD_COEFF if dt != hip.F16 else [c / 2.0625 for c in FDTD_COEFF]
    a, d0, d1 = _fdtd_pair(ctx, n, dt)
    ref = oracle.fdtd3d(a.copy(), np.zeros_like(a), *n, coeff, 3, bf16=(dt == hip.BF16))
    names = hip.fdtd3d_candidates(dt)
    assert len(names) >= 2
    ran = 0
    for i, name in enumerate(names):
        d0.upload(a)
        try:
            ctx.fdtd3d(d0, d1, *n, coeff, 3, hip.CANDIDATE0 + i)
        except hip.AppfisError as e:
            if name.startswith("dma "):  # direct-to-LDS rows: whole 16-byte chunks per row
                assert "needs nx + 8 divisible" in str(e) and (n[0] + 8) * np.dtype(a.dtype).itemsize % 16, e
            else:  # a candidate whose x points per lane do not divide nx
                assert "needs nx" in str(e) and int(name.split()[0][2:]) > 1 and n[0] % int(name.split()[0][2:]), e
            continue
        ran += 1
        assert np.array_equal(d1.download().view(np.uint8), ref.view(np.uint8)), name
    assert ran >= 2


@pytest.mark.parametrize("dt,off", [(hip.F32, 4), (hip.F64, 8), (hip.F16, 2), (hip.BF16, 2)])
def test_fdtd3d_on_unaligned_buffers(ctx, dt, off):
    # buffers that start `off` bytes into an allocation, so that the preferred tilings'
    # alignment fails (16 bytes for the direct-to-LDS rows, 4 for the 16-bit pairs):
    # PLAIN and CHOSEN fall back to tilings that fit and stay bit-exact (nx + 8 = 72 and
    # an even nx would otherwise admit them)
    from types import SimpleNamespace
    n = (64, 20, 17)
    coeff = FDTD_COEFF if dt != hip.F16 else [c / 2.0625 for c in FDTD_COEFF]
    a, d0, d1 = _fdtd_pair(ctx, n, dt)
    ref = oracle.fdtd3d(a.copy(), np.zeros_like(a), *n, coeff, 3, bf16=(dt == hip.BF16))
    big = [ctx.array((a.nbytes + 64,), hip.U8) for _ in range(2)]
    bufs = [SimpleNamespace(ptr=b.ptr + off, dtype=dt) for b in big]
    for variant in (hip.PLAIN, hip.CHOSEN):
        print(f"[fdtd unaligned] dtype {dt} offset {off} variant {variant}", flush=True)
        hip.check(hip.lib().appfis_memcpy_h2d(ctx.handle, bufs[0].ptr, a.ctypes.data, a.nbytes), "h2d")
        ctx.fdtd3d(bufs[0], bufs[1], *n, coeff, 3, variant)
        got = np.empty_like(a)
        hip.check(hip.lib().appfis_memcpy_d2h(ctx.handle, got.ctypes.data, bufs[1].ptr, a.nbytes), "d2h")
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), variant


# unit roundoff of the storage formats: every sweep rounds each stored value once
FDTD_UNIT_ROUNDOFF = {hip.F32: 2.0 ** -24, hip.F16: 2.0 ** -11, hip.BF16: 2.0 ** -8}


def test_fdtd3d_full_config(ctx):
    """512^3, 100 sweeps.  FP32 two sweeps bit-exact against the oracle; FP64 (the
    reference dtype) all 100 sweeps bit-exact against the oracle's FP64; FP32, FP16 and
    BF16 storage (CHOSEN tiling) against that FP64 field, on normalised fields
    u / 2.0625^n (FP16 runs with the coefficients divided by 2.0625, so its field is
    already normalised).

    Bound: the seeded input is rounded to the storage type once and every sweep rounds
    each stored value once (relative error <= u, the unit roundoff, each time); the
    normalised operator has non-negative weights summing to 1 with a zero halo, so it
    does not amplify earlier errors in the max norm or the L2 norm.  After n sweeps the
    normalised max-abs error is at most (n + 1) * u * max|u0| (max|u0| < 1) and the
    relative L2 error at most (n + 1) * u * |u0| / |u_n|, asserted as 2 (n + 1) u after
    checking |u_n| / |u0| > 0.5 (the mean 0.5 of the seeded field survives 100 sweeps in
    the interior)."""
    n = (512, 512, 512)
    steps = 100
    scale = 2.0625 ** steps
    a, d0, d1 = _fdtd_pair(ctx, n, hip.F32)
    ctx.fdtd3d(d0, d1, *n, FDTD_COEFF, 2)
    ref = oracle.fdtd3d(a.copy(), np.zeros_like(a), *n, FDTD_COEFF, 2)
    assert np.array_equal(d0.download(), ref)
    del ref, d0, d1
    # FP64: the oracle for all 100 sweeps, and the GPU bit-identical to it
    a64 = np.zeros(a.shape)
    oracle.fill_uniform(a64, 4, 1004)
    f64 = oracle.fdtd3d(a64.copy(), np.zeros_like(a64), *n, FDTD_COEFF, steps)
    assert np.linalg.norm(f64.ravel()) / np.linalg.norm(a64.ravel()) / scale > 0.5
    e0, e1 = ctx.from_numpy(a64), ctx.array(a64.shape, hip.F64)
    ctx.fdtd3d(e0, e1, *n, FDTD_COEFF, steps, hip.CHOSEN)
    assert np.array_equal(e0.download(), f64)
    print(f"[fdtd3d] FP64 512^3 x {steps}: bit-identical to the oracle", flush=True)
    del e0, e1
    f64 /= scale
    norm64 = np.linalg.norm(f64.ravel())
    for dt in (hip.F32, hip.F16, hip.BF16):
        coeff = FDTD_COEFF if dt != hip.F16 else [c / 2.0625 for c in FDTD_COEFF]
        b, b0, b1 = _fdtd_pair(ctx, n, dt)
        ctx.fdtd3d(b0, b1, *n, coeff, steps, hip.CHOSEN)
        raw = b0.download()
        del b, b0, b1
        if dt == hip.BF16:
            got = (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        else:
            got = raw.astype(np.float64)
        if dt != hip.F16:
            got /= scale
        diff = got - f64
        max_abs = float(np.abs(diff).max())
        rel_l2 = float(np.linalg.norm(diff.ravel()) / norm64)
        bound = (steps + 1) * FDTD_UNIT_ROUNDOFF[dt]
        print(f"[fdtd3d] dtype {dt} vs FP64 after {steps} sweeps (normalised): max-abs {max_abs:.3e} "
              f"rel-L2 {rel_l2:.3e} (bound {bound:.3e})", flush=True)
        assert max_abs <= bound and rel_l2 <= 2 * bound
        del got, diff, raw
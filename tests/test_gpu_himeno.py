"""GPU parity of the Himeno 19-point Jacobi sweep with the gosa residual.

PLAIN F32 / F64: the field p is bit-exact against the oracle (same operation order, no
FMA contraction on either side); gosa agrees to 1e-12 relative (the FP64 sums run in a
different order).  CHOSEN (defect correction: FP32 inner sweeps on the correction, FP64
rebase every `outer` sweeps) is checked against the FP64 oracle with the stated
tolerances HIMENO_DC_GOSA_RTOL and HIMENO_DC_P_ATOL.
"""
import numpy as np
import pytest

import oracle
from appfis_amd import hip

pytestmark = pytest.mark.gpu

# stated accuracy of the defect-correction form against the FP64 oracle
HIMENO_DC_GOSA_RTOL = 2e-6
HIMENO_DC_P_ATOL = 3e-8


def himeno_arrays(nx, ny, nz, dtype, seed=None):
    """Config init: p = (k/(nz-1))^2 on every cell (both p buffers), a0..a2 = 1
    (or U[0.9,1.1] from seeds seed..seed+2 on the interior), a3 = 1/6, b = 0, c = 1,
    bnd = 1, wrk1 = 0."""
    shape = (nz, ny, nx)
    k = np.arange(nz, dtype=np.float64).reshape(nz, 1, 1)
    p = np.broadcast_to((k / (nz - 1)) ** 2, shape).astype(dtype)
    arrs = [p.copy(), p.copy()]
    for idx, v in enumerate([1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]):
        a = np.full(shape, v, dtype=dtype)
        if seed is not None and idx < 3:
            oracle.fill_uniform(a, 1, seed + idx, 0.9, 1.1)
        arrs.append(a)
    return arrs


def run_gpu(ctx, arrs, nx, ny, nz, iters, variant=hip.PLAIN, outer=0):
    dt = hip.F64 if arrs[0].dtype == np.float64 else hip.F32
    dev = [ctx.from_numpy(a, dt) for a in arrs]
    gosa = ctx.himeno(dev, nx, ny, nz, 0.8, iters, variant, outer)
    return dev[iters % 2].download(), gosa


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,iters", [((33, 29, 37), 7), ((65, 65, 129), 12), ((300, 20, 9), 3)])
@pytest.mark.parametrize("seed", [None, 2001])
def test_himeno_plain_bit_exact(ctx, dtype, shape, iters, seed):
    nx, ny, nz = shape
    arrs = himeno_arrays(nx, ny, nz, dtype, seed)
    p, gosa = run_gpu(ctx, arrs, nx, ny, nz, iters)
    ref_arrs = [a.copy() for a in arrs]
    ref_gosa = oracle.himeno(ref_arrs, nx, ny, nz, 0.8, iters)
    assert np.array_equal(p, ref_arrs[iters % 2])
    assert np.allclose(gosa, ref_gosa, rtol=1e-12, atol=0)


def test_himeno_closed_form_first_gosa(ctx):
    # config init: every interior ss = 1/(3 (nz-1)^2) on the first FP64 sweep
    nx, ny, nz = 65, 65, 129
    arrs = himeno_arrays(nx, ny, nz, np.float64)
    _, gosa = run_gpu(ctx, arrs, nx, ny, nz, 1)
    ss = 1.0 / (3 * (nz - 1) ** 2)
    assert abs(gosa[0] - (nx - 2) * (ny - 2) * (nz - 2) * ss * ss) < 1e-10 * gosa[0]


@pytest.mark.parametrize("seed", [None, 2001])
@pytest.mark.parametrize("outer", [0, 4, 16])
def test_himeno_defect_correction_accuracy(ctx, seed, outer):
    nx, ny, nz, iters = 65, 65, 129, 200
    arrs = himeno_arrays(nx, ny, nz, np.float64, seed)
    p, gosa = run_gpu(ctx, arrs, nx, ny, nz, iters, hip.CHOSEN, outer)
    ref_arrs = [a.copy() for a in arrs]
    ref_gosa = oracle.himeno(ref_arrs, nx, ny, nz, 0.8, iters)
    ref_p = ref_arrs[iters % 2]
    gerr = np.max(np.abs(gosa - ref_gosa) / ref_gosa)
    perr = np.max(np.abs(p - ref_p))
    # plain FP32 for comparison (what the mixed form is buying back)
    arrs32 = himeno_arrays(nx, ny, nz, np.float32, seed)
    p32, gosa32 = run_gpu(ctx, arrs32, nx, ny, nz, iters)
    gerr32 = np.max(np.abs(gosa32 - ref_gosa) / ref_gosa)
    perr32 = np.max(np.abs(p32.astype(np.float64) - ref_p))
    print(f"[himeno dc outer={outer} seed={seed}] gosa rel err {gerr:.3e} (FP32 {gerr32:.3e}); "
          f"p max-abs {perr:.3e} (FP32 {perr32:.3e})")
    assert gerr < HIMENO_DC_GOSA_RTOL
    assert perr < HIMENO_DC_P_ATOL


def test_himeno_full_config(ctx):
    # 257 x 257 x 513 with the boundary shell: FP32 plain 200 sweeps bit-exact against the
    # oracle; FP64 first gosa against the closed form; defect correction against FP64
    nx, ny, nz, iters = 257, 257, 513, 200
    arrs = himeno_arrays(nx, ny, nz, np.float32)
    p, gosa = run_gpu(ctx, arrs, nx, ny, nz, iters)
    ref_arrs = [a.copy() for a in arrs]
    ref_gosa = oracle.himeno(ref_arrs, nx, ny, nz, 0.8, iters)
    assert np.array_equal(p, ref_arrs[iters % 2])
    assert np.allclose(gosa, ref_gosa, rtol=1e-12)
    del ref_arrs, arrs
    a64 = himeno_arrays(nx, ny, nz, np.float64)
    p64, g64 = run_gpu(ctx, a64, nx, ny, nz, iters)
    ss = 1.0 / (3 * (nz - 1) ** 2)
    assert abs(g64[0] - 255 * 255 * 511 * ss * ss) < 1e-9 * g64[0]  # FP64 ss cancels ~6 digits
    pdc, gdc = run_gpu(ctx, a64, nx, ny, nz, iters, hip.CHOSEN)
    gerr = np.max(np.abs(gdc - g64) / g64)
    perr = np.max(np.abs(pdc - p64))
    gerr32 = np.max(np.abs(gosa - g64) / g64)
    print(f"[himeno full] gosa[199] FP64 {g64[-1]:.15e}; defect-correction rel err {gerr:.3e}, "
          f"p max-abs {perr:.3e}; plain FP32 gosa rel err {gerr32:.3e}")
    assert gerr < HIMENO_DC_GOSA_RTOL and perr < HIMENO_DC_P_ATOL

"""ORACLE -- test infrastructure only.

ctypes binding of liborc.so, the CPU restatement of the APPFIS reference runtime
(Appfis/Runtime.hpp, Grid.hpp, Communication.hpp) driving the config stencils.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker or the timed CPU baseline.  The product path
(appfis_amd/libappfis_hip.so and the Appfis/ headers) never loads it.

Array convention: numpy arrays in (z, y, x) order, dense and padded exactly like the
reference APPFIS::Array (x fastest).  dims passed to C are (x, y, z).

Pinning of each function (which checks fix it) is listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")
REF_DIR = os.path.join(_HERE, "_ref")

U8, I32, F32, F64, F16, BF16, I64 = 0, 1, 2, 3, 4, 5, 6
MAX, MIN, SUM = 0, 1, 2

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "liborc.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        vp, i, d, u64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
        ip = C.POINTER(C.c_int)
        dp = C.POINTER(C.c_double)
        sigs = {
            "orc_rng_u64": ([u64, u64], u64),
            "orc_fill_uniform": ([vp, i, i, ip, i, u64, d, d], i),
            "orc_fill_bernoulli_u8": ([vp, i, ip, i, u64, d], i),
            "orc_set_boundary": ([vp, i, i, ip, i, i, vp], i),
            "orc_set_boundary_copy": ([vp, i, i, ip, i, i], i),
            "orc_exchange": ([vp, i, i, ip, i, i], i),
            "orc_reduce": ([vp, i, i, ip, ip, i, vp], i),
            "orc_reduce_sum_wide": ([vp, i, i, ip, ip, vp], i),
            "orc_heat2d_f64": ([vp, vp, i, i, d, i], None),
            "orc_heat2d_f32": ([vp, vp, i, i, C.c_float, i], None),
            "orc_heat3d_ref": ([vp, vp, i, i, i, d, d, d, i], None),
            "orc_gol_u8": ([vp, vp, i, i, i], None),
            "orc_gol_ref_i32": ([vp, vp, i, i, i, i], None),
            "orc_fdtd3d": ([i, vp, vp, i, i, i, dp, i], i),
            "orc_fdtd3d_ref": ([vp, vp, i, i, i, C.c_float, i, i], None),
            "orc_himeno": ([i, C.POINTER(vp), i, i, i, d, i, dp], i),
            "orc_flood_run_f64": ([vp, vp, vp, vp, vp, i, i, d, d, d, d, d, d, d, i, dp, dp, dp], i),
            "orc_flood_run_f32": ([vp, vp, vp, vp, vp, i, i, d, d, d, d, d, d, d, i, dp, dp, dp], i),
            "orc_flood_run_src_f64": ([vp, vp, vp, vp, vp, i, i, d, d, d, d, d, d, d, i, i, vp, i, vp, vp, i,
                                       dp, dp, dp], i),
            "orc_flood_run_src_f32": ([vp, vp, vp, vp, vp, i, i, d, d, d, d, d, d, d, i, i, vp, i, vp, vp, i,
                                       dp, dp, dp], i),
            "orc_flood_dem_f64": ([vp, i, i, d, u64, i, d, d], i),
            "orc_omp_threads": ([], i),
            "orc_f32_to_f16_bits": ([C.c_float], C.c_uint16),
            "orc_f16_bits_to_f32": ([C.c_uint16], C.c_float),
            "orc_f32_to_bf16_bits": ([C.c_float], C.c_uint16),
            "orc_bf16_bits_to_f32": ([C.c_uint16], C.c_float),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


_CODE = {np.dtype(np.uint8): U8, np.dtype(np.int32): I32, np.dtype(np.float32): F32,
         np.dtype(np.float64): F64, np.dtype(np.float16): F16, np.dtype(np.int64): I64}


def _code(a: np.ndarray, bf16: bool = False) -> int:
    if bf16:
        assert a.dtype == np.uint16
        return BF16
    return _CODE[a.dtype]


def _dims(a: np.ndarray):
    d = list(reversed(a.shape))
    return (C.c_int * len(d))(*d), len(d)


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _chk(st, what):
    if st != 0:
        raise RuntimeError(f"oracle {what} failed ({st})")


def omp_threads() -> int:
    return int(lib().orc_omp_threads())


# ---- inputs -----------------------------------------------------------------------
def rng_u64(seed: int, n: int) -> int:
    return int(lib().orc_rng_u64(seed, n))


def fill_uniform(a: np.ndarray, ghost: int, seed: int, lo=0.0, hi=1.0, bf16=False):
    dims, nd = _dims(a)
    _chk(lib().orc_fill_uniform(_ptr(a), _code(a, bf16), nd, dims, ghost, seed, lo, hi), "fill_uniform")
    return a


def fill_bernoulli(a: np.ndarray, ghost: int, seed: int, p: float):
    dims, nd = _dims(a)
    _chk(lib().orc_fill_bernoulli_u8(_ptr(a), nd, dims, ghost, seed, p), "fill_bernoulli")
    return a


# ---- runtime semantics --------------------------------------------------------------
def set_boundary(a: np.ndarray, ghost: int, face: int, value, bf16=False):
    dims, nd = _dims(a)
    v = np.array([value], dtype=a.dtype)
    _chk(lib().orc_set_boundary(_ptr(a), _code(a, bf16), nd, dims, ghost, face, v.ctypes.data), "set_boundary")
    return a


def set_boundary_copy(a: np.ndarray, ghost: int, face: int, bf16=False):
    dims, nd = _dims(a)
    _chk(lib().orc_set_boundary_copy(_ptr(a), _code(a, bf16), nd, dims, ghost, face), "set_boundary_copy")
    return a


def exchange(a: np.ndarray, ghost: int, periodic: bool, bf16=False):
    dims, nd = _dims(a)
    _chk(lib().orc_exchange(_ptr(a), _code(a, bf16), nd, dims, ghost, int(bool(periodic))), "exchange")
    return a


def reduce(a: np.ndarray, box, op: int):
    dims, nd = _dims(a)
    b = (C.c_int * len(box))(*box)
    out = np.zeros(1, dtype=a.dtype)
    _chk(lib().orc_reduce(_ptr(a), _code(a), nd, dims, b, op, out.ctypes.data), "reduce")
    return out[0]


def reduce_sum_wide(a: np.ndarray, box):
    dims, nd = _dims(a)
    b = (C.c_int * len(box))(*box)
    out = np.zeros(1, dtype=np.int64 if a.dtype in (np.uint8, np.int32) else np.float64)
    _chk(lib().orc_reduce_sum_wide(_ptr(a), _code(a), nd, dims, b, out.ctypes.data), "reduce_sum_wide")
    return out[0]


# ---- stencils (in place on the pair (b0, b1); result in [b0, b1][n % 2]) ----------
def heat2d(b0, b1, nx, ny, c, iters):
    if b0.dtype == np.float64:
        lib().orc_heat2d_f64(_ptr(b0), _ptr(b1), nx, ny, c, iters)
    else:
        lib().orc_heat2d_f32(_ptr(b0), _ptr(b1), nx, ny, c, iters)
    return (b0, b1)[iters % 2]


def heat3d_ref(b0, b1, nx, ny, nz, cx, cy, cz, iters):
    lib().orc_heat3d_ref(_ptr(b0), _ptr(b1), nx, ny, nz, cx, cy, cz, iters)
    return (b0, b1)[iters % 2]


def gol_u8(b0, b1, nx, ny, gens):
    lib().orc_gol_u8(_ptr(b0), _ptr(b1), nx, ny, gens)
    return (b0, b1)[gens % 2]


def gol_ref_i32(b0, b1, nx, ny, gens, periodic=True):
    lib().orc_gol_ref_i32(_ptr(b0), _ptr(b1), nx, ny, gens, int(periodic))
    return (b0, b1)[gens % 2]


def fdtd3d(b0, b1, nx, ny, nz, coeff5, steps, bf16=False):
    cf = (C.c_double * 5)(*coeff5)
    _chk(lib().orc_fdtd3d(_code(b0, bf16), _ptr(b0), _ptr(b1), nx, ny, nz, cf, steps), "fdtd3d")
    return (b0, b1)[steps % 2]


def fdtd3d_ref(b0, b1, nx, ny, nz, coeff, radius, steps):
    lib().orc_fdtd3d_ref(_ptr(b0), _ptr(b1), nx, ny, nz, coeff, radius, steps)
    return (b0, b1)[steps % 2]


def himeno(arrays, nx, ny, nz, omega, iters):
    ptrs = (C.c_void_p * 14)(*[_ptr(a) for a in arrays])
    gosa = np.zeros(max(iters, 1), dtype=np.float64)
    _chk(lib().orc_himeno(_code(arrays[0]), ptrs, nx, ny, nz, omega, iters,
                          gosa.ctypes.data_as(C.POINTER(C.c_double))), "himeno")
    return gosa[:iters]


def flood_run(h, qx, qy, z, n2, nx, ny, dx, courant, hextra, dt_fallback, print_interval, rain, start_time,
              steps):
    t = np.zeros(max(steps, 1))
    v = np.zeros(max(steps, 1))
    dt = np.zeros(max(steps, 1))
    fn = lib().orc_flood_run_f64 if h.dtype == np.float64 else lib().orc_flood_run_f32
    dpp = C.POINTER(C.c_double)
    _chk(fn(_ptr(h), _ptr(qx), _ptr(qy), _ptr(z), _ptr(n2), nx, ny, dx, courant, hextra, dt_fallback,
            print_interval, rain, start_time, steps, t.ctypes.data_as(dpp), v.ctypes.data_as(dpp),
            dt.ctypes.data_as(dpp)), "flood_run")
    return t[:steps], v[:steps], dt[:steps]


def flood_run_src(h, qx, qy, z, n2, nx, ny, dx, courant, hextra, dt_fallback, print_interval, rain, start_time,
                  steps, fixed_dt=False, src_cell=(), src_time=(), src_flow=None):
    """flood_run with the reference app's point sources and fixed-increment mode:
    src_cell are padded linear indices, src_time the hydrograph times (s) and src_flow a
    (rows, nsrc) table of flows (m^3/s)."""
    t = np.zeros(max(steps, 1))
    v = np.zeros(max(steps, 1))
    dt = np.zeros(max(steps, 1))
    cells = np.ascontiguousarray(src_cell, dtype=np.int32)
    times = np.ascontiguousarray(src_time, dtype=np.float64)
    flows = np.ascontiguousarray(src_flow if src_flow is not None else np.zeros((0, 0)), dtype=np.float64)
    nsrc = len(cells)
    if nsrc:
        assert flows.shape == (len(times), nsrc), (flows.shape, len(times), nsrc)
    fn = lib().orc_flood_run_src_f64 if h.dtype == np.float64 else lib().orc_flood_run_src_f32
    dpp = C.POINTER(C.c_double)
    _chk(fn(_ptr(h), _ptr(qx), _ptr(qy), _ptr(z), _ptr(n2), nx, ny, dx, courant, hextra, dt_fallback,
            print_interval, rain, start_time, int(bool(fixed_dt)), nsrc, _ptr(cells) if nsrc else None,
            len(times), _ptr(times) if nsrc else None, _ptr(flows) if nsrc else None, steps,
            t.ctypes.data_as(dpp), v.ctypes.data_as(dpp), dt.ctypes.data_as(dpp)), "flood_run_src")
    return t[:steps], v[:steps], dt[:steps]


def flood_dem(nx, ny, dx, seed, nhills=20, slope=0.001, wall=1e6):
    z = np.zeros((ny + 2, nx + 2), dtype=np.float64)
    _chk(lib().orc_flood_dem_f64(_ptr(z), nx, ny, dx, seed, nhills, slope, wall), "flood_dem")
    return z


# ---- binary16 / bfloat16 storage conversions (round to nearest even) ----------------
def f32_to_f16_bits(x: float) -> int:
    return int(lib().orc_f32_to_f16_bits(x))


def f32_to_bf16_bits(x: float) -> int:
    return int(lib().orc_f32_to_bf16_bits(x))


def bf16_bits_to_f32(h: int) -> float:
    return float(lib().orc_bf16_bits_to_f32(h))

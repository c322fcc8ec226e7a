"""ctypes binding of libappfis_hip.so (the C ABI declared in include/appfis_hip.h).

This is the host-side plumbing used by the parity tests and bench.py; C++ callers use
the Appfis/ headers, which sit on the same C ABI.  Loading fails loudly when the
shared library has not been built: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libappfis_hip.so")

# element types / reductions / variants (include/appfis_hip.h)
U8, I32, F32, F64, F16, BF16, I64 = 0, 1, 2, 3, 4, 5, 6
MAX, MIN, SUM = 0, 1, 2
PLAIN, CHOSEN = 0, 1
CANDIDATE0 = 2  # CANDIDATE0 + i: candidate i of an autotuned kernel's table
# faces (APPFIS::DIRECTION)
LEFT, RIGHT, UP, DOWN, FRONT, BACK = 0, 1, 2, 3, 4, 5

NP_DTYPE = {
    U8: np.uint8,
    I32: np.int32,
    F32: np.float32,
    F64: np.float64,
    F16: np.float16,
    BF16: np.uint16,  # raw bfloat16 bits
    I64: np.int64,
}


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    for k, v in NP_DTYPE.items():
        if k == BF16:
            continue
        if np.dtype(v) == dt:
            return k
    raise ValueError(f"no APPFIS dtype for {dt}")


class AppfisError(RuntimeError):
    pass


class FloodParams(C.Structure):
    _fields_ = [
        ("dx", C.c_double),
        ("courant", C.c_double),
        ("hextra", C.c_double),
        ("dt_fallback", C.c_double),
        ("print_interval", C.c_double),
        ("rain_rate", C.c_double),
        ("start_time", C.c_double),
        ("time_increment_fixed", C.c_int),
        ("end_time", C.c_double),
    ]


class FloodSources(C.Structure):
    _fields_ = [
        ("count", C.c_int),
        ("cell", C.POINTER(C.c_int)),
        ("rows", C.c_int),
        ("time", C.POINTER(C.c_double)),
        ("flow", C.POINTER(C.c_double)),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile libappfis_hip.so in-tree with hipcc (gfx950)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AppfisError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (or make -C appfis_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, sz, u64 = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_uint64
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    sigs = {
        "appfis_abi_version": ([], i),
        "appfis_status_string": ([i], C.c_char_p),
        "appfis_last_error": ([], C.c_char_p),
        "appfis_device_count": ([ip], i),
        "appfis_ctx_create": ([i, C.POINTER(vp)], i),
        "appfis_ctx_destroy": ([vp], i),
        "appfis_ctx_device": ([vp, ip], i),
        "appfis_ctx_stream": ([vp], vp),
        "appfis_sync": ([vp], i),
        "appfis_device_name": ([vp, C.c_char_p, i], i),
        "appfis_tune_report": ([vp, C.c_char_p, i], i),
        "appfis_malloc": ([vp, sz, C.POINTER(vp)], i),
        "appfis_free": ([vp, vp], i),
        "appfis_memcpy_h2d": ([vp, vp, vp, sz], i),
        "appfis_memcpy_d2h": ([vp, vp, vp, sz], i),
        "appfis_memcpy_d2d": ([vp, vp, vp, sz], i),
        "appfis_memset": ([vp, vp, i, sz], i),
        "appfis_timer_start": ([vp], i),
        "appfis_timer_stop": ([vp, C.POINTER(C.c_float)], i),
        "appfis_graph_begin": ([vp], i),
        "appfis_graph_end": ([vp, C.POINTER(vp)], i),
        "appfis_graph_launch": ([vp, vp, i], i),
        "appfis_graph_destroy": ([vp, vp], i),
        "appfis_fill_uniform": ([vp, vp, i, i, ip, i, u64, d, d], i),
        "appfis_fill_bernoulli_u8": ([vp, vp, i, ip, i, u64, d], i),
        "appfis_fill_value": ([vp, vp, i, sz, d], i),
        "appfis_set_boundary": ([vp, vp, i, i, ip, i, i, vp], i),
        "appfis_set_boundary_copy": ([vp, vp, i, i, ip, i, i], i),
        "appfis_exchange": ([vp, vp, i, i, ip, i, i], i),
        "appfis_reduce": ([vp, vp, i, i, ip, ip, i, vp], i),
        "appfis_reduce_sum_wide": ([vp, vp, i, i, ip, ip, vp], i),
        "appfis_reduce_device": ([vp, vp, i, i, ip, ip, i, vp], i),
        "appfis_reduce_device_noalloc": ([vp, vp, i, i, ip, ip, i, vp], i),
        "appfis_heat2d": ([vp, i, vp, vp, i, i, d, i, i, i], i),
        "appfis_gol": ([vp, vp, vp, i, i, i, i, i], i),
        "appfis_kernel_timing": ([vp, i], i),
        "appfis_kernel_time": ([vp, dp, ip, ip], i),
        "appfis_fdtd3d": ([vp, i, vp, vp, i, i, i, dp, i, i], i),
        "appfis_fdtd3d_candidates": ([i, C.c_char_p, i], i),
        "appfis_heat2d_candidates": ([i, C.c_char_p, i], i),
        "appfis_heat2d_ensemble": ([vp, i, i, ip, ip, C.POINTER(vp), C.POINTER(vp), d, i], i),
        "appfis_gol_candidates": ([C.c_char_p, i], i),
        "appfis_himeno": ([vp, i, C.POINTER(vp), i, i, i, d, i, i, i, dp], i),
        "appfis_himeno_candidates": ([C.c_char_p, i], i),
        "appfis_flood_run": ([vp, i, vp, vp, vp, vp, vp, i, i, C.POINTER(FloodParams), i, i, dp, dp], i),
        "appfis_flood_run_src": ([vp, i, vp, vp, vp, vp, vp, i, i, C.POINTER(FloodParams), C.POINTER(FloodSources),
                                  i, i, dp, dp, ip], i),
        "appfis_flood_ensemble_run": ([vp, i, i, ip, ip, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                       C.POINTER(vp), C.POINTER(vp), C.POINTER(FloodParams), i, dp, dp], i),
        "appfis_flood_dem": ([vp, i, vp, i, i, d, u64, i, d, d], i),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _candidates(fn, *args):
    buf = C.create_string_buffer(4096)
    n = fn(*args, buf, 4096)
    if n < 0:
        raise AppfisError(lib().appfis_last_error().decode(errors="replace"))
    return buf.value.decode().splitlines()[:n]


def fdtd3d_candidates(dtype: int):
    """Names of the FDTD3d candidate tilings (variant CANDIDATE0 + index)."""
    return _candidates(lib().appfis_fdtd3d_candidates, dtype)


def heat2d_candidates(dtype: int):
    """Names of the Heat2D temporal-blocking candidates (variant CANDIDATE0 + index)."""
    return _candidates(lib().appfis_heat2d_candidates, dtype)


def gol_candidates():
    """Names of the bit-packed GoL blocked kernels (variant CANDIDATE0 + index)."""
    return _candidates(lib().appfis_gol_candidates)


def himeno_candidates():
    """Names of the Himeno sweep shapes, PLAIN then defect correction (variant
    CANDIDATE0 + index)."""
    return _candidates(lib().appfis_himeno_candidates)


def check(status: int, what: str = "") -> None:
    if status != 0:
        L = lib()
        msg = L.appfis_last_error().decode(errors="replace")
        kind = L.appfis_status_string(status).decode()
        raise AppfisError(f"{what}: {kind} ({status}): {msg}")


def _ints(vals):
    arr = (C.c_int * len(vals))(*[int(v) for v in vals])
    return arr


class DeviceArray:
    """A device allocation holding a dense APPFIS array of `shape` (z, y, x) order in
    numpy terms; the C ABI sees dims = (x, y, z) = reversed(shape)."""

    def __init__(self, ctx: "Context", shape, dtype: int):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.np_dtype = np.dtype(NP_DTYPE[dtype])
        self.nbytes = int(np.prod(self.shape)) * self.np_dtype.itemsize
        p = C.c_void_p()
        check(lib().appfis_malloc(ctx.handle, self.nbytes, C.byref(p)), "appfis_malloc")
        self.ptr = p.value

    @property
    def dims(self):
        return list(reversed(self.shape))

    @property
    def ndim(self):
        return len(self.shape)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.np_dtype)
        assert a.shape == self.shape, (a.shape, self.shape)
        check(lib().appfis_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes), "h2d")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.np_dtype)
        check(lib().appfis_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            check(lib().appfis_free(self.ctx.handle, self.ptr), "appfis_free")
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                lib().appfis_free(self.ctx.handle, self.ptr)
        except Exception:
            pass


class Context:
    """One HIP device + one stream (Topology's device/stream, one per host thread)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().appfis_ctx_create(device, C.byref(h)), "appfis_ctx_create")
        self.handle = h.value

    def close(self):
        if self.handle:
            lib().appfis_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def name(self) -> str:
        buf = C.create_string_buffer(256)
        check(lib().appfis_device_name(self.handle, buf, 256), "device_name")
        return buf.value.decode()

    def sync(self):
        check(lib().appfis_sync(self.handle), "sync")

    def tune_report(self) -> str:
        buf = C.create_string_buffer(65536)
        check(lib().appfis_tune_report(self.handle, buf, 65536), "tune_report")
        return buf.value.decode()

    def array(self, shape, dtype: int) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def from_numpy(self, a: np.ndarray, dtype: int | None = None) -> DeviceArray:
        dt = dtype if dtype is not None else dtype_code(a.dtype)
        return DeviceArray(self, a.shape, dt).upload(a)

    # ---- timing / graphs -------------------------------------------------------
    def kernel_timing(self, on: bool):
        """Record events around the generation kernel's launches of later gol calls."""
        check(lib().appfis_kernel_timing(self.handle, 1 if on else 0), "kernel_timing")

    def kernel_time(self):
        """(ms, launches, generations) of the generation kernel in the last timed gol call."""
        ms, n, g = C.c_double(), C.c_int(), C.c_int()
        check(lib().appfis_kernel_time(self.handle, C.byref(ms), C.byref(n), C.byref(g)), "kernel_time")
        return ms.value, n.value, g.value

    def timer_start(self):
        check(lib().appfis_timer_start(self.handle), "timer_start")

    def timer_stop(self) -> float:
        ms = C.c_float()
        check(lib().appfis_timer_stop(self.handle, C.byref(ms)), "timer_stop")
        return float(ms.value)

    def graph_begin(self):
        check(lib().appfis_graph_begin(self.handle), "graph_begin")

    def graph_end(self):
        g = C.c_void_p()
        check(lib().appfis_graph_end(self.handle, C.byref(g)), "graph_end")
        return g.value

    def graph_launch(self, g, times: int = 1):
        check(lib().appfis_graph_launch(self.handle, g, times), "graph_launch")

    def graph_destroy(self, g):
        check(lib().appfis_graph_destroy(self.handle, g), "graph_destroy")

    # ---- runtime ops -----------------------------------------------------------
    def fill_uniform(self, arr: DeviceArray, ghost: int, seed: int, lo: float = 0.0, hi: float = 1.0):
        check(lib().appfis_fill_uniform(self.handle, arr.ptr, arr.dtype, arr.ndim, _ints(arr.dims), ghost,
                                        seed, lo, hi), "fill_uniform")

    def fill_bernoulli(self, arr: DeviceArray, ghost: int, seed: int, p: float):
        check(lib().appfis_fill_bernoulli_u8(self.handle, arr.ptr, arr.ndim, _ints(arr.dims), ghost, seed, p),
              "fill_bernoulli")

    def fill_value(self, arr: DeviceArray, value: float):
        check(lib().appfis_fill_value(self.handle, arr.ptr, arr.dtype, int(np.prod(arr.shape)), value),
              "fill_value")

    def set_boundary(self, arr: DeviceArray, ghost: int, face: int, value):
        v = np.array([value], dtype=arr.np_dtype)
        check(lib().appfis_set_boundary(self.handle, arr.ptr, arr.dtype, arr.ndim, _ints(arr.dims), ghost, face,
                                        v.ctypes.data), "set_boundary")

    def set_boundary_copy(self, arr: DeviceArray, ghost: int, face: int):
        check(lib().appfis_set_boundary_copy(self.handle, arr.ptr, arr.dtype, arr.ndim, _ints(arr.dims), ghost,
                                             face), "set_boundary_copy")

    def exchange(self, arr: DeviceArray, ghost: int, periodic: bool):
        check(lib().appfis_exchange(self.handle, arr.ptr, arr.dtype, arr.ndim, _ints(arr.dims), ghost,
                                    int(bool(periodic))), "exchange")

    def reduce(self, arr: DeviceArray, box, op: int):
        out = np.zeros(1, dtype=arr.np_dtype)
        check(lib().appfis_reduce(self.handle, arr.ptr, arr.dtype, arr.ndim, _ints(arr.dims), _ints(box), op,
                                  out.ctypes.data), "reduce")
        return out[0]

    def reduce_sum_wide(self, arr: DeviceArray, box):
        wide = np.int64 if arr.dtype in (U8, I32) else np.float64
        out = np.zeros(1, dtype=wide)
        check(lib().appfis_reduce_sum_wide(self.handle, arr.ptr, arr.dtype, arr.ndim, _ints(arr.dims), _ints(box),
                                           out.ctypes.data), "reduce_sum_wide")
        return out[0]

    # ---- config kernels --------------------------------------------------------
    def heat2d(self, b0: DeviceArray, b1: DeviceArray, nx, ny, c, iters, variant=PLAIN, tsteps=0):
        check(lib().appfis_heat2d(self.handle, b0.dtype, b0.ptr, b1.ptr, nx, ny, c, iters, variant, tsteps),
              "heat2d")

    def heat2d_ensemble(self, grids, c, iters):
        """grids: list of (buf0, buf1, nx, ny); all of one dtype."""
        n = len(grids)
        dt = grids[0][0].dtype if grids else F64  # an empty list is refused by the library
        b0 = (C.c_void_p * n)(*[g[0].ptr for g in grids])
        b1 = (C.c_void_p * n)(*[g[1].ptr for g in grids])
        check(lib().appfis_heat2d_ensemble(self.handle, dt, n, _ints([g[2] for g in grids]),
                                           _ints([g[3] for g in grids]), b0, b1, c, iters), "heat2d_ensemble")

    def gol(self, b0: DeviceArray, b1: DeviceArray, nx, ny, gens, variant=PLAIN, tsteps=0):
        check(lib().appfis_gol(self.handle, b0.ptr, b1.ptr, nx, ny, gens, variant, tsteps), "gol")

    def fdtd3d(self, b0: DeviceArray, b1: DeviceArray, nx, ny, nz, coeff5, steps, variant=PLAIN):
        cf = (C.c_double * 5)(*[float(x) for x in coeff5])
        check(lib().appfis_fdtd3d(self.handle, b0.dtype, b0.ptr, b1.ptr, nx, ny, nz, cf, steps, variant),
              "fdtd3d")

    def himeno(self, arrays, nx, ny, nz, omega, iters, variant=PLAIN, outer=0, dtype=None):
        ptrs = (C.c_void_p * 14)(*[a.ptr for a in arrays])
        gosa = np.zeros(max(iters, 1), dtype=np.float64)
        dt = arrays[0].dtype if dtype is None else dtype
        check(lib().appfis_himeno(self.handle, dt, ptrs, nx, ny, nz, omega, iters, variant, outer,
                                  gosa.ctypes.data_as(C.POINTER(C.c_double))), "himeno")
        return gosa[:iters]

    def flood_run(self, h, qx, qy, z, n2, nx, ny, params: FloodParams, steps, variant=PLAIN):
        t = np.zeros(max(steps, 1), dtype=np.float64)
        v = np.zeros(max(steps, 1), dtype=np.float64)
        check(lib().appfis_flood_run(self.handle, h.dtype, h.ptr, qx.ptr, qy.ptr, z.ptr, n2.ptr, nx, ny,
                                     C.byref(params), steps, variant,
                                     t.ctypes.data_as(C.POINTER(C.c_double)),
                                     v.ctypes.data_as(C.POINTER(C.c_double))), "flood_run")
        return t[:steps], v[:steps]

    def flood_run_src(self, h, qx, qy, z, n2, nx, ny, params: FloodParams, steps, variant=PLAIN, src_cell=(),
                      src_time=(), src_flow=None):
        """flood_run with point sources: src_cell padded linear indices, src_time (s) and
        src_flow a (rows, count) table (m^3/s).  Returns (time, volume) of the steps run
        (fewer than `steps` when params.end_time stopped the loop)."""
        cells = np.ascontiguousarray(src_cell, dtype=np.int32)
        times = np.ascontiguousarray(src_time, dtype=np.float64)
        flows = np.ascontiguousarray(src_flow if src_flow is not None else np.zeros((0, 0)), dtype=np.float64)
        src = FloodSources(len(cells), cells.ctypes.data_as(C.POINTER(C.c_int)), len(times),
                           times.ctypes.data_as(C.POINTER(C.c_double)), flows.ctypes.data_as(C.POINTER(C.c_double)))
        t = np.zeros(max(steps, 1), dtype=np.float64)
        v = np.zeros(max(steps, 1), dtype=np.float64)
        done = C.c_int(0)
        check(lib().appfis_flood_run_src(self.handle, h.dtype, h.ptr, qx.ptr, qy.ptr, z.ptr, n2.ptr, nx, ny,
                                         C.byref(params), C.byref(src), steps, variant,
                                         t.ctypes.data_as(C.POINTER(C.c_double)),
                                         v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done)), "flood_run_src")
        return t[:done.value], v[:done.value]

    def flood_ensemble_run(self, scen, params: FloodParams, steps, dtype):
        n = len(scen)
        nx = _ints([s["nx"] for s in scen])
        ny = _ints([s["ny"] for s in scen])

        def ptrs(key):
            return (C.c_void_p * n)(*[s[key].ptr for s in scen])

        t = np.zeros(n * max(steps, 1), dtype=np.float64)
        v = np.zeros(n * max(steps, 1), dtype=np.float64)
        check(lib().appfis_flood_ensemble_run(self.handle, dtype, n, nx, ny, ptrs("h"), ptrs("qx"), ptrs("qy"),
                                              ptrs("z"), ptrs("n2"), C.byref(params), steps,
                                              t.ctypes.data_as(C.POINTER(C.c_double)),
                                              v.ctypes.data_as(C.POINTER(C.c_double))), "flood_ensemble_run")
        return t.reshape(n, -1)[:, :steps], v.reshape(n, -1)[:, :steps]

    def flood_dem(self, z: DeviceArray, nx, ny, dx, seed, nhills=20, slope=0.001, wall=1e6):
        check(lib().appfis_flood_dem(self.handle, z.dtype, z.ptr, nx, ny, dx, seed, nhills, slope, wall),
              "flood_dem")


def device_count() -> int:
    n = C.c_int(0)
    check(lib().appfis_device_count(C.byref(n)), "device_count")
    return n.value

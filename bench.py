#!/usr/bin/env python3
"""APPFIS iterative-stencil benchmark on one MI355X.

Headline (the printed JSON line's `value`): GameOfLife 2D, 8192 x 8192 uint8 cells,
periodic, rule B3/S23, seeded Bernoulli(0.25); one step = one generation over the whole
grid, `--steps 1000` is the config's 1000 generations.  Inputs are generated on the
device before the timed region (resident in HBM).  Each timed region is bracketed by a
device synchronisation on both sides and measured with HIP events recorded on the
library's stream (the stream every kernel of the call is launched on).

The same line carries `suite`: the other configs measured the same way (Heat2D FP64, a
64-grid Heat2D ensemble, FDTD3d FP32/FP64/FP16/BF16, Himeno FP32/FP64/defect-correction,
the flood ensemble in
plain FP64 one scenario at a time and in the batched FP32 form), each with the
roofline of its dominant kernel; `--headline-only` skips them.

Roofline: achieved = algorithmic bytes per launch of the dominant kernel / its average
launch time (for GameOfLife: HIP events that the library records around the generation
kernel's launches inside the timed call, over the launch count; for the suite: timed
region / launches); peak = 8 TB/s HBM.  `traffic` is the HBM bytes
per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, calibrated on a copy
kernel), read from profiles/traffic.json when present.  `valu` is the vector-ALU issue
rate: the wave instructions rocprofv3 counted per dispatch (SQ_INSTS_VALU, same file)
over the launch time measured here, against 256 CUs x 64 lanes x 2.4 GHz; the
VALU-bound kernels (bit-packed GoL, flood flux) are bounded by it.

cpu_baseline: the oracle (OpenMP C restatement of the reference CPU path) timed on this
host on a bounded sample of the headline workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Nominal VALU issue rate: 256 CUs x 64 lanes x 2.4 GHz lane-instructions per second (T),
# one wave64 instruction per 4 cycles per SIMD; FP64 add/mul/FMA issue at this rate too
# on MI355X (78.6 TFLOP/s = this x 2 for FMA).  Measured rates can exceed it (rocprofv3's
# VALUBusy reads above 100% for the same kernels), so `frac` is against this nominal
# single-issue rate, not a hard ceiling.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
FDTD_COEFF = [0.5, 0.125, 0.0625, 0.125 / 3, 0.03125]
HIMENO_DC_OUTER = 16  # the library's default rebase interval of the Himeno defect correction


def _traffic(kernels):
    """HBM bytes per launch from profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy kernel).  `kernels` maps kernel names (template form, as the
    summary writes them) to weights: the weighted sum covers one roofline unit."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        tab = json.load(open(path))
        return round(sum(w * float(tab[k]["bytes_per_launch"]) for k, w in kernels.items()), 1)
    except (KeyError, ValueError, TypeError):
        return None


def _valu(kernels, launch_ms):
    """VALU issue rate of one roofline unit: the wave instructions rocprofv3 counted per
    dispatch (SQ_INSTS_VALU, profiles/traffic.json) x 64 lanes, weighted like the bytes,
    over the launch time measured here; plus rocprofv3's derived VALUBusy per kernel (the
    share of the dispatch's time the vector ALUs issue; it includes the dispatch's ramp,
    so it reads low for kernels of a few microseconds).  None when not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        tab = json.load(open(path))
    except ValueError:
        return None
    if not all(isinstance(tab.get(k), dict) and "valu_wave_insts_per_launch" in tab[k] for k in kernels):
        return None
    insts = sum(w * float(tab[k]["valu_wave_insts_per_launch"]) for k, w in kernels.items())
    tops = insts * 64 / (launch_ms * 1e-3) / 1e12
    return {"wave_insts_per_unit": insts, "achieved": round(tops, 2), "peak": VALU_PEAK_TOPS,
            "unit": "T lane-instr/s", "frac": round(tops / VALU_PEAK_TOPS, 4),
            "valu_busy": {k: round(tab[k]["valu_busy_pct"] / 100.0, 4) for k in kernels if "valu_busy_pct" in tab[k]}}


def roofline(alg_bytes_per_launch, launch_ms, kernels):
    """`kernels`: the dominant kernel's name, or {name: launches per roofline unit}."""
    if isinstance(kernels, str):
        kernels = {kernels: 1.0}
    achieved = alg_bytes_per_launch / (launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic(kernels),
            "kernel": " + ".join(kernels), "alg_bytes_per_launch": alg_bytes_per_launch,
            "avg_launch_us": round(launch_ms * 1e3, 3), "valu": _valu(kernels, launch_ms)}


def timed(ctx, fn):
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    fn()
    ms = ctx.timer_stop()
    ctx.sync()
    return ms, time.perf_counter() - t0


# ------------------------------------------------------------------------------------
# GameOfLife (headline)
# ------------------------------------------------------------------------------------
def bench_gol(ctx, hip, steps, warmup, variant, nx=8192, ny=8192):
    """One call of `steps` generations, timed.  Inside that call the library records HIP
    events on its stream before the first and after the last launch of the generation
    kernel (appfis_kernel_timing): their interval over the launch count is the kernel's
    average launch duration.  It leaves out the PLAIN form's initial ring exchange; for
    the bit-packed form with full-row waves it spans the whole call (its first launch
    packs, its last unpacks), otherwise it leaves out the packing and the last generation
    fused with the unpacking of both buffers."""
    d0 = ctx.array((ny + 2, nx + 2), hip.U8)
    d1 = ctx.array((ny + 2, nx + 2), hip.U8)
    ctx.fill_bernoulli(d0, 1, 1002, 0.25)
    if warmup:
        ctx.gol(d0, d1, nx, ny, warmup, variant)
    ctx.fill_bernoulli(d0, 1, 1002, 0.25)
    ctx.kernel_timing(True)
    ms, wall = timed(ctx, lambda: ctx.gol(d0, d1, nx, ny, steps, variant))
    k_ms, k_launches, k_gens = ctx.kernel_time()
    ctx.kernel_timing(False)
    final = (d0, d1)[steps % 2]
    live = int(ctx.reduce_sum_wide(final, [1, nx + 1, 1, ny + 1]))
    res = {"form": "plain" if variant == hip.PLAIN else "chosen", "ms": round(ms, 4), "wall_s": round(wall, 4),
           "glups": round(nx * ny * steps / (ms * 1e-3) / 1e9, 2), "live_count": live,
           "generation_kernel": {"ms": round(k_ms, 4), "launches": k_launches, "generations": k_gens}}
    if k_launches == 0:  # a one-generation call runs no separate generation kernel
        return res
    launch_ms = k_ms / k_launches
    gens_per_launch = k_gens / k_launches
    if variant == hip.PLAIN:
        # uint8 field read + written once per generation
        res["roofline"] = roofline(2.0 * nx * ny, launch_ms, f"k_gol_plain<{(nx + 2) & 15}, 16>")
    else:
        # bit-packed field (1 bit per cell) read + written once per launch; the blocked
        # kernel advances up to G generations per launch (its autotuned name says "gG";
        # the call's last launch may run fewer)
        tname = _tuned(ctx, 4, hip.U8)
        res["tuned"] = tname
        res["autotune"] = [l for l in ctx.tune_report().splitlines() if l.startswith("autotune kernel 4 ")]
        kname = gol_kernel_name(tname)
        if tname and tname.startswith("tw") and k_gens == steps and k_launches >= 2:
            # the full-row form's first launch packs the uint8 field as it loads, its
            # last unpacks both buffers (k_step_tw_io); the ones between are k_step_tw
            io = "k_step_tw_io<" + kname[len("k_step_tw<"):-1]
            kname = {f"{io}, true, false>": 1.0 / k_launches, kname: (k_launches - 2.0) / k_launches,
                     f"{io}, false, true>": 1.0 / k_launches}
        # Algorithmic bytes on the workload's own basis: 2 B per cell update (the uint8
        # field read and written once per generation), the launches' generations each.
        # The blocked kernel moves the field packed (1 bit per cell) and keeps it in the
        # 256 MB last-level cache across generations, so this rate can exceed the HBM
        # peak: HBM does not bind it; integer VALU issue does (roofline.valu, measured
        # from SQ_INSTS_VALU).  The packed field's own bytes per launch are reported apart.
        g = gens_per_launch
        rf = roofline(2.0 * nx * ny * g, launch_ms, kname)
        rf["bound"] = "valu"
        rf["basis"] = ("2 B per cell update (uint8 read + write per generation) x the generations per launch "
                       "(generation_kernel: generations / launches); bound = VALU issue, whose fraction is "
                       "roofline.valu.frac")
        packed = roofline(2.0 * nx * ny / 8, launch_ms, kname)
        rf["packed_field"] = {"alg_bytes_per_launch": packed["alg_bytes_per_launch"],
                              "achieved": packed["achieved"], "unit": "GB/s", "frac": packed["frac"],
                              "note": "the 1-bit field read + written once per launch"}
        res["roofline"] = rf
    d0.free()
    d1.free()
    return res


def heat_kernel_name(tname, dtype="double"):
    """Kernel symbol of a Heat2D temporal-blocking candidate: "tb 64x64 ts12 nt1024" ->
    k_heat2d_tb<double, 64, 64, 12, 1024> (output tile, sweeps, threads); "rt 48x96 ts8 k7
    w16" -> k_heat2d_rt<double, 7, 16, 8, 1> (rows per wave, waves, sweeps, columns per
    lane: "rt2" holds two)."""
    form, tile, ts, a, *rest = tname.split()
    if form.startswith("rt"):
        wpl = int(form[2:] or 1)
        return f"k_heat2d_rt<{dtype}, {int(a[1:])}, {int(rest[0][1:])}, {int(ts[2:])}, {wpl}>"
    tx, ty = tile.split("x")
    return f"k_heat2d_tb<{dtype}, {tx}, {ty}, {int(ts[2:])}, {int(a[2:])}>"


def gol_kernel_name(tname):
    """Kernel symbol of a bit-packed GoL candidate: "tb k12 g8 w16" -> k_step_tb<12, 8, 16>,
    "tw4 k4 g8 w12" -> k_step_tw<4, 4, 8, 12> (words per lane first); no candidate (widths
    the blocked kernels do not take) -> the one-generation k_step<8>."""
    if not tname:
        return "k_step<8>"
    form, *rest = tname.split()
    k, g, w = (int(t[1:]) for t in rest[:3])
    if form.startswith("tw"):
        return f"k_step_tw<{int(form[2:])}, {k}, {g}, {w}>"
    return f"k_step_{form}<{k}, {g}, {w}>"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_gol(nx=8192, ny=8192, target_s=15.0, runs=3):
    """The oracle's OpenMP GoL on the host cores: `runs` timed runs of the same prefix of
    the 1000 generations (about target_s seconds in all), reported as their median."""
    import oracle

    a = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
    oracle.fill_bernoulli(a, 1, 1002, 0.25)
    b = np.zeros_like(a)
    t0 = time.perf_counter()
    oracle.gol_u8(a, b, nx, ny, 2)
    per = (time.perf_counter() - t0) / 2
    gens = max(2, min(1000, int(target_s / runs / max(per, 1e-6))))
    rates = []
    for _ in range(runs):
        oracle.fill_bernoulli(a, 1, 1002, 0.25)
        t0 = time.perf_counter()
        oracle.gol_u8(a, b, nx, ny, gens)
        rates.append(nx * ny * gens / (time.perf_counter() - t0) / 1e9)
    med = float(np.median(rates))
    env = {k: os.environ.get(k, "") for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    return {"value": round(med, 4), "unit": "GLUP/s", "cores": oracle.omp_threads(), "kind": "port",
            "cpu_model": cpu_model(), "runs": [round(r, 4) for r in rates], "omp_env": env,
            "sample": f"GameOfLife {nx}x{ny} uint8, the first {gens} of the 1000 generations, median of {runs} "
                      f"runs of the oracle's OpenMP restatement (schedule(dynamic,1) over rows) on "
                      f"{oracle.omp_threads()} threads of {os.cpu_count()} host CPUs ({cpu_model()})"}


COPY_BYTES = 1 << 30


def calibrate_copy(ctx, hip):
    """A device-to-device copy of COPY_BYTES, run 4 times: it reads and writes exactly
    COPY_BYTES per dispatch, which scales the FETCH_SIZE / WRITE_SIZE counters."""
    a = ctx.array((COPY_BYTES,), hip.U8)
    b = ctx.array((COPY_BYTES,), hip.U8)
    for _ in range(4):
        hip.check(hip.lib().appfis_memcpy_d2d(ctx.handle, b.ptr, a.ptr, COPY_BYTES), "d2d")
    ctx.sync()
    a.free()
    b.free()


# ------------------------------------------------------------------------------------
# suite
# ------------------------------------------------------------------------------------
def bench_heat(ctx, hip, n=1024, iters=500):
    out = {}
    a = ctx.array((n + 2, n + 2), hip.F64)
    b = ctx.array((n + 2, n + 2), hip.F64)
    for name, variant in (("plain", hip.PLAIN), ("chosen", hip.CHOSEN)):
        ctx.fill_uniform(a, 1, 1001)
        ctx.heat2d(a, b, n, n, 0.1, 20, variant)  # warm-up
        ctx.fill_uniform(a, 1, 1001)
        ms, _ = timed(ctx, lambda: ctx.heat2d(a, b, n, n, 0.1, iters, variant))
        chk = float(ctx.reduce_sum_wide(a, [1, n + 1, 1, n + 1]))
        if variant == hip.PLAIN:
            launches = iters
        else:
            tname = _tuned(ctx, 1, hip.F64)  # e.g. "tb 64x64 ts8 nt1024", "rt 48x96 ts8 k7 w16"
            ts = int(tname.split()[2][2:])
            launches = len(_plan(iters, ts))
        per_launch_sweeps = iters / launches
        kname = "k_heat2d_plain<double, 4>" if variant == hip.PLAIN else heat_kernel_name(tname)
        r = {"ms": round(ms, 4), "glups": round(n * n * iters / (ms * 1e-3) / 1e9, 2), "checksum": chk,
             "roofline": roofline(16.0 * n * n * per_launch_sweeps, ms / launches, kname)}
        if variant == hip.CHOSEN:
            r["tuned"] = tname
            r["autotune"] = [l for l in ctx.tune_report().splitlines() if l.startswith("autotune kernel 1 ")]
        out[name] = r
    # mixed precision: the same sweeps in FP32 (temporal blocking), error against FP64
    ctx.fill_uniform(a, 1, 1001)
    init64 = a.download()
    ctx.heat2d(a, b, n, n, 0.1, iters, hip.CHOSEN)
    ref64 = a.download()
    a32 = ctx.from_numpy(init64.astype(np.float32), hip.F32)
    b32 = ctx.array((n + 2, n + 2), hip.F32)
    ctx.heat2d(a32, b32, n, n, 0.1, 20, hip.CHOSEN)  # warm-up (and the one-time autotuning)
    a32.upload(init64.astype(np.float32))
    ms, _ = timed(ctx, lambda: ctx.heat2d(a32, b32, n, n, 0.1, iters, hip.CHOSEN))
    f32 = a32.download().astype(np.float64)
    out["chosen_f32"] = {"ms": round(ms, 4), "glups": round(n * n * iters / (ms * 1e-3) / 1e9, 2),
                         "tuned": _tuned(ctx, 1, hip.F32),
                         "max_abs_err_vs_f64": float(np.max(np.abs(f32 - ref64))),
                         "rel_l2_err_vs_f64": float(np.linalg.norm(f32 - ref64) / np.linalg.norm(ref64)),
                         "speedup_vs_plain_f64": round(out["plain"]["ms"] / ms, 3)}
    a32.free()
    b32.free()
    out["workload"] = f"Heat2D {n}x{n} FP64 5-point, {iters} sweeps (16 B/LUP); chosen_f32 = FP32 storage"
    if CPU_SAMPLES:
        import oracle

        h = init64.copy()
        out["cpu"] = _cpu_rate(lambda k: oracle.heat2d(h, np.zeros_like(h), n, n, 0.1, k), n * n,
                               f"Heat2D {n}x{n} FP64")
    return out


CPU_SAMPLES = True  # per-app CPU oracle samples next to the suite entries (off with --no-cpu)


def _cpu_rate(run, cells_per_step, what, target_s=3.0):
    """GLUP/s of the oracle (OpenMP restatement of the reference CPU path) on a bounded
    number of steps: one step to size the sample, then about target_s seconds."""
    import oracle

    t0 = time.perf_counter()
    run(1)
    one = time.perf_counter() - t0
    k = max(1, min(200, int(target_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    run(k)
    dt = time.perf_counter() - t0
    return {"value": round(cells_per_step * k / dt / 1e9, 4), "unit": "GLUP/s", "cores": oracle.omp_threads(),
            "kind": "port", "sample": f"{what}: {k} steps of the oracle in {dt:.2f} s"}


def _splitmix_sides(n, seed, lo, hi):
    """side_s = lo + floor(u_s * (hi - lo + 1)), u_s = (r(seed, s) >> 11) * 2^-53 (splitmix64)."""
    out = []
    for s in range(n):
        z = (seed + (s + 1) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        z ^= z >> 31
        out.append(lo + int((z >> 11) * (hi - lo + 1) / 2 ** 53))
    return out


def bench_heat_ensemble(ctx, hip, ngrid=64, iters=500):
    """64 FP64 grids, sides uniform in [128, 1024] (splitmix64 seed 1006), 500 sweeps:
    one grid at a time (PLAIN; CHOSEN at depth 8) against the batched ensemble."""
    sides = _splitmix_sides(ngrid, 1006, 128, 1024)
    grids = []
    for k, n in enumerate(sides):
        a, b = ctx.array((n + 2, n + 2), hip.F64), ctx.array((n + 2, n + 2), hip.F64)
        grids.append((a, b, n, n))
    cells = sum(n * n for n in sides)

    def reset():
        for k, (a, b, n, _) in enumerate(grids):
            ctx.fill_uniform(a, 1, 3000 + k)

    out = {"workload": f"Heat2D ensemble: {ngrid} FP64 grids, sides in [128,1024] (sum N^2 = {cells}), "
                       f"{iters} sweeps", "sides": sides}
    res = {}
    for name, fn in (("plain_one_at_a_time", lambda: [ctx.heat2d(a, b, n, n, 0.1, iters, hip.PLAIN)
                                                      for a, b, n, _ in grids]),
                     ("chosen_one_at_a_time", lambda: [ctx.heat2d(a, b, n, n, 0.1, iters, hip.CHOSEN, 8)
                                                       for a, b, n, _ in grids]),
                     ("ensemble", lambda: ctx.heat2d_ensemble(grids, 0.1, iters))):
        reset()
        fn()  # warm-up
        reset()
        ms, _ = timed(ctx, fn)
        res[name] = [(a, b)[iters % 2].download().sum() for a, b, _, _ in grids[:2]]
        r = {"ms": round(ms, 3), "glups": round(cells * iters / (ms * 1e-3) / 1e9, 2)}
        if name == "ensemble":
            # the FP64 ensemble runs 48 x 96 register tiles (7 rows x 16 waves), 8 sweeps per launch
            launches = len(_plan(iters, 8))
            r["roofline"] = roofline(16.0 * cells * iters / launches, ms / launches, "k_heat2d_rt_ens<double, 7, 16, 8, 1>")
        out[name] = r
    assert res["ensemble"] == res["plain_one_at_a_time"] == res["chosen_one_at_a_time"]
    out["ensemble"]["speedup_vs_plain"] = round(out["plain_one_at_a_time"]["ms"] / out["ensemble"]["ms"], 3)
    for a, b, _, _ in grids:
        a.free()
        b.free()
    return out


def _tuned(ctx, kernel, dtype):
    """Name of the variant the autotuner picked last for (kernel, dtype)."""
    pick = None
    for line in ctx.tune_report().splitlines():
        if line.startswith(f"autotune kernel {kernel} dtype {dtype} "):
            pick = line.rsplit("-> ", 1)[1].strip()
    return pick


def _plan(iters, ts):
    n = (iters + ts - 1) // ts
    if (n & 1) != (iters & 1):
        n += 1
    return [0] * min(n, iters)




# the PLAIN form's preferred fixed tilings (the library's rows marked plain), used where
# the shape admits them: the bench's buffers are 256-byte aligned and n = 512
FDTD_PLAIN_PREFERRED = {"f32": "dma ty16 ry2 zc256 d2", "f64": "vx2 ty16 ry1 zc256 d2 u2",
                        "f16": "vx2 ty8 ry2 zc128 d2 u18", "bf16": "vx2 ty8 ry2 zc128 d2 u18"}


def _fdtd_plain(hip, name, dt, n):
    """The PLAIN form's fixed tiling at n^3: the preferred row where it fits (F32: n + 8
    a multiple of 4; the others: n even), else the table's row 0."""
    fits = (n + 8) % 4 == 0 if name == "f32" else n % 2 == 0
    return FDTD_PLAIN_PREFERRED[name] if fits else hip.fdtd3d_candidates(dt)[0]


FDTD_CTYPE = {"f32": "float", "f64": "double", "f16": "__half", "bf16": "__hip_bfloat16"}


def bench_fdtd(ctx, hip, n=512, steps=100):
    out = {"workload": f"FDTD3d r=4 {n}^3, {steps} steps (CHOSEN = autotuned tiling, same values as PLAIN)"}
    finals = {}  # each dtype's field after the timed CHOSEN run, as float64, normalised by 2.0625^steps
    for name, dt, bpl in (("f32", hip.F32, 8), ("f64", hip.F64, 16), ("f16", hip.F16, 4), ("bf16", hip.BF16, 4)):
        shape = (n + 8, n + 8, n + 8)
        a = ctx.array(shape, dt)
        b = ctx.array(shape, dt)
        coeff = FDTD_COEFF if dt != hip.F16 else [c / 2.0625 for c in FDTD_COEFF]
        for form, variant in (("plain", hip.PLAIN), ("chosen", hip.CHOSEN)):
            ctx.fill_uniform(a, 4, 1004)
            ctx.fdtd3d(a, b, n, n, n, coeff, 2, variant)  # warm-up (and the one-time autotuning)
            ctx.fill_uniform(a, 4, 1004)
            ms, _ = timed(ctx, lambda: ctx.fdtd3d(a, b, n, n, n, coeff, steps, variant))
            tname = _fdtd_plain(hip, name, dt, n) if variant == hip.PLAIN else _tuned(ctx, 3, dt)
            if tname.startswith("dma "):  # direct-to-LDS form: "dma ty16 ry2 zc256 d2"
                ty, ry, zc, dd = (int(t[2:]) if t[:2] in ("ty", "ry", "zc") else int(t[1:])
                                  for t in tname.split()[1:])
                kname = f"k_fdtd3d_dma<{FDTD_CTYPE[name]}, {ty}, {ry}, {zc}, {dd}>"
            else:
                vx, ty, ry, zc, dd, uu = (int(t[2:]) if t[:2] in ("vx", "ty", "ry", "zc") else int(t[1:])
                                          for t in tname.split())
                kname = f"k_fdtd3d<{FDTD_CTYPE[name]}, {vx}, {ty}, {ry}, {zc}, {dd}, {uu}>"
            out[f"{name}_{form}"] = {"ms": round(ms, 3), "glups": round(n ** 3 * steps / (ms * 1e-3) / 1e9, 2),
                                     "tiling": tname, "roofline": roofline(float(bpl) * n ** 3, ms / steps, kname)}
            if form == "chosen":
                raw = (a, b)[steps % 2].download()[4:-4, 4:-4, 4:-4]
                if dt == hip.BF16:
                    f = (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
                else:
                    f = raw.astype(np.float64)
                if dt != hip.F16:  # FP16 runs with the coefficients divided by 2.0625
                    f /= 2.0625 ** steps
                finals[name] = f
                del raw
        a.free()
        b.free()
    # error of each storage type against FP64 (same seeded input, normalised fields)
    ref = finals["f64"]
    norm = float(np.linalg.norm(ref.ravel()))
    for name in ("f32", "f16", "bf16"):
        d = finals[name] - ref
        out[f"{name}_chosen"]["error_vs_f64"] = {"max_abs_normalised": float(np.abs(d).max()),
                                                  "rel_l2": float(np.linalg.norm(d.ravel()) / norm)}
        del d
    finals.clear()
    out["autotune"] = [l for l in ctx.tune_report().splitlines() if "kernel 3 " in l]
    if CPU_SAMPLES:
        import oracle

        u = np.zeros((n + 8, n + 8, n + 8), dtype=np.float32)
        oracle.fill_uniform(u, 4, 1004)
        v = np.zeros_like(u)
        out["cpu"] = _cpu_rate(lambda k: oracle.fdtd3d(u, v, n, n, n, FDTD_COEFF, k), n ** 3, f"FDTD3d {n}^3 FP32")
    return out


def _himeno_dev(ctx, hip, nx, ny, nz, dt):
    npdt = np.float64 if dt == hip.F64 else np.float32
    k = np.arange(nz, dtype=np.float64).reshape(nz, 1, 1)
    p = np.ascontiguousarray(np.broadcast_to((k / (nz - 1)) ** 2, (nz, ny, nx)).astype(npdt))
    arrs = [ctx.from_numpy(p, dt), ctx.from_numpy(p, dt)]
    for v in [1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]:
        a = ctx.array((nz, ny, nx), dt)
        ctx.fill_value(a, v)
        arrs.append(a)
    return arrs, p


def bench_himeno(ctx, hip, nx=257, ny=257, nz=513, iters=200):
    out = {"workload": f"Himeno 19-point {nx}x{ny}x{nz} (incl. shell), {iters} sweeps + gosa"}
    lups = (nx - 2) * (ny - 2) * (nz - 2) * iters
    gosas = {}
    for name, dt, variant, bpl in (("f32", hip.F32, hip.PLAIN, 56), ("f64", hip.F64, hip.PLAIN, 112),
                                   ("dc", hip.F64, hip.CHOSEN, None)):
        arrs, p = _himeno_dev(ctx, hip, nx, ny, nz, dt)
        ctx.himeno(arrs, nx, ny, nz, 0.8, 3, variant)  # warm-up
        arrs[0].upload(p)
        arrs[1].upload(p)

        def run():
            gosas[name] = ctx.himeno(arrs, nx, ny, nz, 0.8, iters, variant)

        ms, _ = timed(ctx, run)
        r = {"ms": round(ms, 3), "glups": round(lups / (ms * 1e-3) / 1e9, 2), "gosa_last": float(gosas[name][-1])}
        pts = (nx - 2) * (ny - 2) * (nz - 2)
        # "zc<planes per block> d<coefficient prefetch depth>[ ry<rows per thread>]"
        shape = _tuned(ctx, 5, dt) or "zc128 d1"
        d = int(shape.split()[1][1:])
        ry = int(shape.split()[2][2:]) if len(shape.split()) > 2 else 2
        r["tiling"] = shape
        if bpl:
            r["roofline"] = roofline(float(bpl) * pts, ms / iters,
                                     f"k_himeno<{'float' if dt == hip.F32 else 'double'}, 0, {d}, {ry}>")
        else:
            # defect correction, per sweep on average over one outer cycle of K sweeps
            # (the library's default K = 16): K-1 FP32 inner sweeps (11 coefficients, read
            # from a copy blocked by 64 points, r, q read, q written: 56 B/LUP) and one FP64
            # rebase (11 coefficients, wrk1, B read, q read 4 B, B / r / q written
            # 8 + 4 + 4 B: 124 B/LUP)
            K = HIMENO_DC_OUTER
            r["roofline"] = roofline(((K - 1) * 56.0 + 124.0) / K * pts, ms / iters,
                                     {f"k_himeno<float, 2, {d}, {ry}>": (K - 1) / K,
                                      f"k_himeno<double, 1, {d}, {ry}>": 1 / K})
        out[name] = r
        for a in arrs:
            a.free()
    ref = gosas["f64"]
    for name in ("f32", "dc"):
        out[name]["gosa_max_rel_err_vs_f64"] = float(np.max(np.abs(gosas[name] - ref) / ref))
    out["dc"]["speedup_vs_f64"] = round(out["f64"]["ms"] / out["dc"]["ms"], 3)
    if CPU_SAMPLES:
        import oracle

        k = np.arange(nz, dtype=np.float64).reshape(nz, 1, 1)
        p = np.ascontiguousarray(np.broadcast_to((k / (nz - 1)) ** 2, (nz, ny, nx)))
        harr = [p.copy(), p.copy()] + [np.full((nz, ny, nx), v) for v in
                                       (1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0)]
        out["cpu"] = _cpu_rate(lambda it: oracle.himeno(harr, nx, ny, nz, 0.8, it), (nx - 2) * (ny - 2) * (nz - 2),
                               f"Himeno {nx}x{ny}x{nz} FP64")
    out["autotune"] = [l for l in ctx.tune_report().splitlines() if l.startswith("autotune kernel 5 ")]
    return out


def flood_sizes(nscen=24, seed=1005):
    """N_s = 512 + floor(u_s * 3585), u_s = (r(seed, s) >> 11) * 2^-53 (splitmix64)."""
    sizes = []
    for s in range(nscen):
        z = seed + (s + 1) * 0x9E3779B97F4A7C15
        z &= (1 << 64) - 1
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        z ^= z >> 31
        sizes.append(512 + int((z >> 11) * (4096 - 512 + 1) / 2 ** 53))
    return sizes


def bench_flood(ctx, hip, steps=2000, nscen=24):
    sizes = flood_sizes(nscen)
    cells = sum(n * n for n in sizes)
    prm = hip.FloodParams(10.0, 0.5, 1e-3, 1.0, 0.0, 50e-3 / 3600.0, 0.0)
    out = {"workload": f"Triton flood ensemble: {nscen} scenarios, N in [512,4096] (sum N^2 = {cells}), "
                       f"{steps} steps, 10 m cells, rain 50 mm/h, n = 0.035, CFL 0.5",
           "sizes": sizes}

    def make(dt):
        sc = []
        for s, n in enumerate(sizes):
            d = {"nx": n, "ny": n}
            for key in ("h", "qx", "qy", "z", "n2"):
                d[key] = ctx.array((n + 2, n + 2), dt)
            ctx.flood_dem(d["z"], n, n, 10.0, 1005 + s)
            ctx.fill_value(d["n2"], 0.035 ** 2)
            sc.append(d)
        return sc

    def reset(sc):
        for d in sc:
            for key in ("h", "qx", "qy"):
                ctx.fill_value(d[key], 0.0)

    # plain FP64, one scenario at a time, direct launches
    sc = make(hip.F64)
    vol64 = []

    def run_plain():
        for d in sc:
            t, v = ctx.flood_run(d["h"], d["qx"], d["qy"], d["z"], d["n2"], d["nx"], d["ny"], prm, steps, hip.PLAIN)
            vol64.append(v[-1])

    ctx.flood_run(sc[0]["h"], sc[0]["qx"], sc[0]["qy"], sc[0]["z"], sc[0]["n2"], sizes[0], sizes[0], prm, 4,
                  hip.PLAIN)
    reset(sc)
    ms64, _ = timed(ctx, run_plain)
    h64 = [d["h"].download() for d in sc]
    for d in sc:
        for k in ("h", "qx", "qy", "z", "n2"):
            d[k].free()
    # chosen FP32 bulk, all scenarios batched, graph-captured chunks
    sc = make(hip.F32)
    ctx.flood_ensemble_run(sc, prm, 4, hip.F32)
    reset(sc)
    res = {}

    def run_chosen():
        res["t"], res["v"] = ctx.flood_ensemble_run(sc, prm, steps, hip.F32)

    ms32, _ = timed(ctx, run_chosen)
    # FP32 bulk against the FP64 form on every scenario (the FP64 form itself is checked
    # against the FP64 oracle, within 1e-9, on config scenarios by the GPU tests)
    rels = []
    small = int(np.argmin(sizes))  # the scenario the oracle also runs (CPU leg below)
    keep = {}
    for k, (a, d) in enumerate(zip(h64, sc)):
        b = d["h"].download().astype(np.float64)
        rels.append(float(np.linalg.norm(a - b) / np.linalg.norm(a)))
        if k == small:
            keep = {"f64": a, "f32": b, "t32": float(res["t"][k][-1])}
    del h64
    vrels = np.abs(np.array(vol64) - res["v"][:, -1]) / np.array(vol64)
    lups = cells * steps
    # roofline over one whole step (flux/update, wet-dry/CFL/volume, finalise kernels):
    # 64 B per cell FP64 (read h qx qy z n2, write h qx qy), 32 B FP32
    out["plain_f64"] = {"ms": round(ms64, 2), "glups": round(lups / (ms64 * 1e-3) / 1e9, 2),
                        "roofline": roofline(64.0 * cells / nscen, ms64 / (steps * nscen),
                                             {"k_flux_march<double>": 1, "k_fix_march<double, 1>": 1,
                                              "k_finalize<double, 1>": 1})}
    out["chosen_f32_ensemble"] = {"ms": round(ms32, 2), "glups": round(lups / (ms32 * 1e-3) / 1e9, 2),
                                  "speedup_vs_plain_f64": round(ms64 / ms32, 3),
                                  "h_rel_l2_vs_f64": {"max": max(rels), "median": float(np.median(rels)),
                                                      "per_scenario": [round(r, 9) for r in rels]},
                                  "volume_rel_err_vs_f64": {"max": float(vrels.max()),
                                                            "median": float(np.median(vrels))},
                                  "roofline": roofline(32.0 * cells, ms32 / steps,
                                                       {"k_flux_march<float>": 1, "k_fix_march<float, 1>": 1,
                                                        "k_finalize<float, 1>": 1})}
    for d in sc:
        for k in ("h", "qx", "qy", "z", "n2"):
            d[k].free()
    if CPU_SAMPLES:
        # the oracle (FP64) on the smallest scenario for the whole 2000 steps: the CPU
        # sample, and the reference for the two GPU forms on that scenario
        import oracle

        n0 = sizes[small]
        z = oracle.flood_dem(n0, n0, 10.0, 1005 + small)
        f = [np.zeros_like(z), np.zeros_like(z), np.zeros_like(z), z, np.full_like(z, 0.035 ** 2)]
        t0 = time.perf_counter()
        tc, vc, _ = oracle.flood_run(*f, n0, n0, 10.0, 0.5, 1e-3, 1.0, 0.0, 50e-3 / 3600.0, 0.0, steps)
        dt_cpu = time.perf_counter() - t0
        out["cpu"] = {"value": round(n0 * n0 * steps / dt_cpu / 1e9, 4), "unit": "GLUP/s",
                      "cores": oracle.omp_threads(), "kind": "port",
                      "sample": f"flood scenario {small} ({n0}^2) FP64, all {steps} steps from its dry start, "
                                f"{dt_cpu:.1f} s"}
        ref = f[0]
        nref = float(np.linalg.norm(ref))
        out["error_vs_oracle"] = {
            "scenario": small, "n": n0, "steps": steps,
            "plain_f64_h_max_abs": float(np.abs(keep["f64"] - ref).max()),
            "chosen_f32_h_rel_l2": float(np.linalg.norm(keep["f32"] - ref) / nref),
            "chosen_f32_simulated_time_s": keep["t32"], "oracle_simulated_time_s": float(tc[-1])}
    return out


# ------------------------------------------------------------------------------------
# profile summaries (rocprofv3 rocpd databases -> profiles/*.md, profiles/traffic.json)
# ------------------------------------------------------------------------------------
PROFILE_GROUPS = [
    ("gol_8192x8192", "GameOfLife 8192^2 uint8 (byte kernel, bit-packed kernels, ring exchange)",
     ("k_gol_plain", "k_pack", "k_unpack", "k_step", "k_wrap_axis<unsigned char>")),
    ("heat2d_1024x1024_f64", "Heat2D 1024^2 FP64 and the 64-grid FP64 ensemble", ("k_heat2d",)),
    ("fdtd3d_512cubed", "FDTD3d 512^3 r=4 (FP32/FP64/FP16/BF16, every autotuner candidate)", ("k_fdtd3d",)),
    ("himeno_257x257x513", "Himeno 257x257x513 (FP32, FP64, defect correction)",
     ("k_himeno", "k_add_correction", "k_f64_to_f32")),
    ("flood_ensemble_24", "Triton flood: 24 scenarios N in [512,4096] (plain FP64 one at a time, FP32 ensemble)",
     ("k_flux_march", "k_fix_march", "k_finalize", "k_init_state", "k_dem")),
]


def _short(name):
    """Kernel name in template form: 'k_fdtd3d<float, 1, 8, 4, 64, 2>'."""
    n = name.replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            n = n[:i]
            break
    for ns in ("golbits::", "gol::"):
        n = n.replace(ns, "")
    return n.strip()


def summarize(kt_db, fetch_db, write_db, cmds, out_dir, valu_db=None):
    import sqlite3

    durs = {}
    for name, dur in sqlite3.connect(kt_db).execute("select name, duration from kernels"):
        durs.setdefault(_short(name), []).append(dur / 1000.0)

    def pmc(db, counter):
        acc = {}
        if not db:
            return acc
        q = "select kernel_name, value from counters_collection where counter_name = ?"
        for name, val in sqlite3.connect(db).execute(q, (counter,)):
            acc.setdefault(_short(name), []).append(float(val))
        return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}

    fetch, write = pmc(fetch_db, "FETCH_SIZE"), pmc(write_db, "WRITE_SIZE")
    busy, insts = pmc(valu_db, "VALUBusy"), pmc(valu_db, "SQ_INSTS_VALU")
    traffic = {}
    if "k_copy16" in fetch and "k_copy16" in write:
        cf = COPY_BYTES / (fetch["k_copy16"][0] * 1024.0)
        cw = COPY_BYTES / (write["k_copy16"][0] * 1024.0)
        traffic["_calibration"] = {
            "copy_bytes": COPY_BYTES, "copy_fetch_kib": fetch["k_copy16"][0], "copy_write_kib": write["k_copy16"][0],
            "fetch_scale": round(cf, 4), "write_scale": round(cw, 4),
            "note": "bytes = FETCH_SIZE KiB * 1024 * fetch_scale + WRITE_SIZE KiB * 1024 * write_scale; "
                    "the counters see L2 misses, so MALL hits count as traffic"}
        for k in sorted(set(fetch) & set(write)):
            traffic[k] = {"fetch_kib": round(fetch[k][0], 1), "write_kib": round(write[k][0], 1),
                          "dispatches": fetch[k][1],
                          "bytes_per_launch": round(fetch[k][0] * 1024 * cf + write[k][0] * 1024 * cw, 1)}
    for k in busy:
        if not k.startswith("_"):
            e = traffic.setdefault(k, {})
            e["valu_busy_pct"] = round(busy[k][0], 2)
            if k in insts:
                e["valu_wave_insts_per_launch"] = round(insts[k][0], 0)
    if traffic:
        with open(os.path.join(out_dir, "traffic.json"), "w") as f:
            json.dump(traffic, f, indent=1, sort_keys=True)
    for fname, title, prefixes in PROFILE_GROUPS:
        names = sorted(k for k in durs if k.startswith(prefixes))
        if not names:
            continue
        lines = [f"# {title}", "", "Made by (from the repository root, on one MI355X):", ""]
        lines += [f"    {c}" for c in cmds]
        lines += ["", "Per-dispatch kernel durations (kernel trace) and, where measured, HBM traffic per",
                  "dispatch (FETCH_SIZE x fetch_scale + WRITE_SIZE x write_scale, KiB as read; scales "
                  "from a 1 GiB copy kernel, see traffic.json `_calibration`).", "",
                  "VALU busy: rocprofv3's derived VALUBusy (share of the dispatch's time the vector ALUs",
                  "issue), with SQ_INSTS_VALU (wave instructions per dispatch), from a third counter run.", "",
                  "| kernel | calls | avg us | min us | max us | total ms | FETCH KiB | WRITE KiB | traffic MB "
                  "| VALU busy % | VALU wave-insts |",
                  "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
        for k in names:
            d = durs[k]
            t = traffic.get(k)
            tf = (f"{t['fetch_kib']:.0f} | {t['write_kib']:.0f} | {t['bytes_per_launch'] / 1e6:.2f}"
                  if t and "bytes_per_launch" in t else "- | - | -")
            tf += (f" | {t['valu_busy_pct']:.1f} | {t.get('valu_wave_insts_per_launch', 0):.0f}"
                   if t and "valu_busy_pct" in t else " | - | -")
            lines.append(f"| `{k}` | {len(d)} | {sum(d) / len(d):.2f} | {min(d):.2f} | {max(d):.2f} | "
                         f"{sum(d) / 1e3:.3f} | {tf} |")
        with open(os.path.join(out_dir, fname + ".md"), "w") as f:
            f.write("\n".join(lines) + "\n")
        print(f"wrote {fname}.md ({len(names)} kernels)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--only", default="", help="comma list of suite entries (heat,heat_ensemble,fdtd,himeno,flood)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="suite entries at their config sizes but few sweeps (for profiler runs)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a 1 GiB device-to-device copy 4 times (HBM counter calibration)")
    ap.add_argument("--summarize", metavar="KT_DB", help="write profiles/ summaries from a rocprofv3 "
                    "--kernel-trace database (no GPU work)")
    ap.add_argument("--fetch-db", help="rocprofv3 --pmc FETCH_SIZE database for --summarize")
    ap.add_argument("--write-db", help="rocprofv3 --pmc WRITE_SIZE database for --summarize")
    ap.add_argument("--valu-db", help="rocprofv3 --pmc VALUBusy SQ_INSTS_VALU database for --summarize")
    ap.add_argument("--prof-cmd", action="append", default=[], help="command text recorded in the summaries")
    args = ap.parse_args()
    if args.summarize:
        summarize(args.summarize, args.fetch_db, args.write_db, args.prof_cmd, os.path.join(ROOT, "profiles"),
                  args.valu_db)
        return
    global CPU_SAMPLES
    CPU_SAMPLES = not args.no_cpu
    # OpenMP placement of the CPU baseline (read when the oracle's OpenMP runtime loads)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); "
              "the stencils would shard by slabs along the slowest axis, one rank per GPU.", file=sys.stderr)
        sys.exit(2)

    from appfis_amd import hip

    ctx = hip.Context(0)
    device = ctx.name().strip()
    if args.calibrate:
        calibrate_copy(ctx, hip)
    plain = bench_gol(ctx, hip, args.steps, args.warmup, hip.PLAIN)
    variants = {"plain": plain}
    best = plain
    try:
        chosen = bench_gol(ctx, hip, args.steps, args.warmup, hip.CHOSEN)
        if chosen["live_count"] != plain["live_count"]:
            raise RuntimeError("chosen GoL form disagrees with the plain form")
        variants["chosen"] = chosen
        if chosen["glups"] > plain["glups"]:
            best = chosen
    except hip.AppfisError as e:
        variants["chosen"] = {"error": str(e)}
    suite = {}
    if not args.headline_only:
        only = set(x for x in args.only.split(",") if x) or {"heat", "heat_ensemble", "fdtd", "himeno", "flood"}
        quick = {"heat": {"iters": 40}, "heat_ensemble": {"iters": 40}, "fdtd": {"steps": 4}, "himeno": {"iters": 10},
                 "flood": {"steps": 300}} if args.quick else {}
        for name, fn in (("heat", bench_heat), ("heat_ensemble", bench_heat_ensemble), ("fdtd", bench_fdtd),
                         ("himeno", bench_himeno), ("flood", bench_flood)):
            if name in only:
                print(f"[bench] {name}", file=sys.stderr, flush=True)
                suite[name] = fn(ctx, hip, **quick.get(name, {}))
    cpu = None if args.no_cpu else cpu_baseline_gol()
    line = {
        "metric": "GLUP/s (GameOfLife 8192^2 uint8 B3/S23 periodic); suite: GLUP/s and % of 8 TB/s per app",
        "value": best["glups"],
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["ms"] / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (splitmix64 Bernoulli p=0.25, seed 1002, generated on device)",
        "config": {"workload": "GameOfLife 2D 8192x8192 uint8, periodic, B3/S23, 1 step = 1 generation",
                   "nx": 8192, "ny": 8192, "form": best["form"], "device": device},
        "roofline": best.get("roofline"),
        "cpu_baseline": cpu,
        "variants": variants,
        "suite": suite,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

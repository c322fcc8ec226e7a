#!/usr/bin/env python3
"""APPFIS iterative-stencil benchmark on one MI355X.

Default workload (the headline line): GameOfLife 2D, 8192 x 8192 uint8 cells, periodic,
rule B3/S23, seeded Bernoulli(0.25), one step = one generation over the whole grid;
`--steps 1000` is the config's 1000 generations.  Inputs are generated on the device
before the timed region (resident in HBM).  The timed region is bracketed by a device
synchronisation on both sides and measured with HIP events on the library's stream.

Printed: ONE JSON line with GLUP/s (interior cells x generations / s / 1e9), the
roofline of the dominant kernel and the CPU baseline (the oracle's OpenMP restatement of
the reference CPU path, timed on a bounded sample on this host).

Other workloads (`--workload heat|fdtd|himeno|flood|all`) print their own JSON lines.
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


def _traffic(kernel_key):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) and committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        v = data.get(kernel_key)
        return None if v is None else float(v["bytes_per_launch"])
    except Exception:
        return None


def roofline(alg_bytes_per_launch, launch_ms, kernel_key):
    achieved = alg_bytes_per_launch / (launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic(kernel_key),
            "kernel": kernel_key, "alg_bytes_per_launch": alg_bytes_per_launch,
            "avg_launch_us": round(launch_ms * 1e3, 3)}


# ------------------------------------------------------------------------------------
# GameOfLife (headline)
# ------------------------------------------------------------------------------------
def bench_gol(ctx, hip, steps, warmup, variant, cpu=True, nx=8192, ny=8192):
    d0 = ctx.array((ny + 2, nx + 2), hip.U8)
    d1 = ctx.array((ny + 2, nx + 2), hip.U8)
    ctx.fill_bernoulli(d0, 1, 1002, 0.25)
    if warmup:
        ctx.gol(d0, d1, nx, ny, warmup, variant)
        ctx.sync()
        ctx.fill_bernoulli(d0, 1, 1002, 0.25)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    ctx.gol(d0, d1, nx, ny, steps, variant)
    ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    final = (d0, d1)[steps % 2]
    live = int(ctx.reduce_sum_wide(final, [1, nx + 1, 1, ny + 1]))
    lups = nx * ny * steps
    res = {"variant": "plain" if variant == hip.PLAIN else "chosen", "ms": ms, "wall_s": wall,
           "glups": lups / (ms * 1e-3) / 1e9, "live_count": live}
    if variant == hip.PLAIN:
        launches = steps
        res["roofline"] = roofline(2.0 * nx * ny, ms / launches, "gol_plain")
    return res


def cpu_baseline_gol(nx=8192, ny=8192, target_s=12.0):
    import oracle

    a = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
    oracle.fill_bernoulli(a, 1, 1002, 0.25)
    b = np.zeros_like(a)
    t0 = time.perf_counter()
    oracle.gol_u8(a, b, nx, ny, 2)
    per = (time.perf_counter() - t0) / 2
    gens = max(2, min(200, int(target_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    oracle.gol_u8(a, b, nx, ny, gens)
    dt = time.perf_counter() - t0
    return {"value": round(nx * ny * gens / dt / 1e9, 4), "unit": "GLUP/s", "cores": oracle.omp_threads(),
            "kind": "port", "sample": f"GameOfLife {nx}x{ny} uint8, {gens} generations (of 1000), "
                                       f"oracle OpenMP restatement, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="gol")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "See DESIGN.md for how the stencils would shard.", file=sys.stderr)
        sys.exit(2)

    from appfis_amd import hip

    ctx = hip.Context(0)
    device = ctx.name()
    if args.workload == "gol":
        plain = bench_gol(ctx, hip, args.steps, args.warmup, hip.PLAIN)
        variants = {"plain": plain}
        best = plain
        try:
            chosen = bench_gol(ctx, hip, args.steps, args.warmup, hip.CHOSEN)
            if chosen["live_count"] != plain["live_count"]:
                raise RuntimeError("chosen form disagrees with plain form")
            variants["chosen"] = chosen
            if chosen["glups"] > plain["glups"]:
                best = chosen
        except hip.AppfisError as e:
            variants["chosen"] = {"error": str(e)}
        cpu = None if args.no_cpu else cpu_baseline_gol()
        line = {
            "metric": "GLUP/s (GameOfLife 8192^2 uint8 B3/S23 periodic)",
            "value": round(best["glups"], 2),
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
                       "nx": 8192, "ny": 8192, "form": best["variant"], "device": device},
            "roofline": plain["roofline"],
            "cpu_baseline": cpu,
            "variants": {k: {kk: vv for kk, vv in v.items() if kk != "roofline"} for k, v in variants.items()},
        }
        print(json.dumps(line), flush=True)
    else:
        print(json.dumps({"error": f"unknown workload {args.workload}"}))
        sys.exit(2)
    ctx.close()


if __name__ == "__main__":
    main()

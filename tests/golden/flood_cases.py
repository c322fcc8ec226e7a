"""Small Triton flood cases, written as the reference Triton app's input files.

Each case is a cfg file plus its DEM, initial fields, Manning map, source locations and
hydrograph, in the formats the reference app reads (Triton/src/config_utils.hpp:106-412,
triton_appfis.hpp:422-552, Appfis/Array.hpp:377-443, inflow.hpp:93-136).  The same
case also yields the padded arrays and the point-source table that the oracle and the
HIP kernels take, derived the way the reference app derives them
(triton_appfis.hpp:130-290: DEM wall 1e6 in the ghost ring, n squared with its ring
copied inward, sources mapped by calc_src_col / calc_src_row).

make_flood_golden.py runs the reference app (built from its sources into oracle/_ref)
on every case and commits its printed outputs as tests/golden/flood_ref.npz; the tests
regenerate the inputs from this module and compare the oracle and the GPU against them.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

WALL = 1e6  # dem.SetBoundary(1e6), triton_appfis.hpp:195


def _grid(nx, ny):
    j, i = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    return i, j


def _bump(i, j, ci, cj, w):
    return np.exp(-((i - ci) ** 2 + (j - cj) ** 2) / w)


def case_mound_asc():
    """ASC inputs: a sloping valley with a hill, a lake filling the low corner and a water
    mound; uniform Manning n = 0.05; dt clamped to the print interval."""
    nx, ny, dx = 48, 36, 10.0
    i, j = _grid(nx, ny)
    dem = 1.0 + 0.8 * _bump(i, j, nx / 3, ny / 2, 40.0) + 0.004 * i - 0.002 * j
    h0 = np.maximum(1.35 - dem, 0.0) + 0.5 * _bump(i, j, 2 * nx / 3, ny / 3, 12.0)
    return dict(name="mound_asc", nx=nx, ny=ny, dx=dx, fmt="ASC", dem=dem, h=h0, const_mann=0.05,
                time_step=0.5, fixed=0, courant=0.5, hextra=0.001, duration=150.0, print_interval=50.0)


def case_bin_varying_n():
    """Binary inputs (DEM with its 6-double header, h, qx, qy and a spatially varying
    Manning n without header), wet everywhere, initial momentum in both directions."""
    nx, ny, dx = 40, 52, 5.0
    i, j = _grid(nx, ny)
    dem = 0.5 + 0.6 * _bump(i, j, 12, 30, 30.0) + 0.4 * _bump(i, j, 30, 12, 50.0) - 0.003 * j
    h0 = np.maximum(1.4 - dem, 0.05) + 0.2 * _bump(i, j, 20, 40, 20.0)
    qx0 = 0.05 * np.sin(2 * math.pi * j / ny) * h0
    qy0 = -0.04 * np.cos(2 * math.pi * i / nx) * h0
    n = 0.02 + 0.08 * (0.5 + 0.5 * np.sin(i / 5.0) * np.cos(j / 7.0))
    return dict(name="bin_varying_n", nx=nx, ny=ny, dx=dx, fmt="BIN", dem=dem, h=h0, qx=qx0, qy=qy0, n=n,
                const_mann=0.0, time_step=0.5, fixed=0, courant=0.5, hextra=0.001, duration=60.0,
                print_interval=20.0, xll=1000.0, yll=2000.0)


def case_sources_dry():
    """Dry start with two point sources fed by a hydrograph (times in hours, a comment
    line in both the hydrograph and the location file); the run outlasts the last
    hydrograph row."""
    nx, ny, dx = 44, 30, 10.0
    i, j = _grid(nx, ny)
    dem = 2.0 + 0.01 * i + 0.6 * (1.0 - _bump(i, j, i, ny / 2, 30.0)) + 0.002 * j
    xll, yll = 500.0, 700.0
    # source k at cell (col, row) of the full grid: centre x = xll + (col + 0.5) dx,
    # y measured from the bottom edge (calc_src_row counts rows from the top)
    cells = [(30, 14), (12, 20)]
    src_xy = [(xll + (c + 0.5) * dx, yll + (ny - r - 0.5) * dx) for c, r in cells]
    hyd = [(0.0, 0.0, 0.0), (0.01, 8.0, 3.0), (0.03, 15.0, 10.0), (0.05, 2.0, 12.0)]
    # 190 s: past the last hydrograph row (180 s), and a state in which no branch of the
    # solver sits within rounding of its threshold (runs to 210 s or 240 s have such a
    # near-tie: a perturbation of one ulp in n^2 moves qx by 1e-5 there)
    return dict(name="sources_dry", nx=nx, ny=ny, dx=dx, fmt="ASC", dem=dem, const_mann=0.04, time_step=1.0,
                fixed=0, courant=0.5, hextra=0.001, duration=190.0, print_interval=95.0, xll=xll, yll=yll,
                sources=src_xy, hydrograph=hyd)


def case_dambreak_frictionless():
    """Dam break on a flat dry bed without friction (const_mann = 0 gives n^2 = 0)."""
    nx, ny, dx = 64, 20, 1.0
    i, j = _grid(nx, ny)
    dem = np.zeros((ny, nx))
    h0 = np.where(i < 24, 1.0, 0.0)
    return dict(name="dambreak_frictionless", nx=nx, ny=ny, dx=dx, fmt="ASC", dem=dem, h=h0, const_mann=0.0,
                time_step=0.1, fixed=0, courant=0.45, hextra=0.001, duration=6.0, print_interval=3.0)


def case_fixed_dt():
    """time_increment_fixed = 1: every step takes time_step, no CFL and no clamp."""
    nx, ny, dx = 30, 42, 10.0
    i, j = _grid(nx, ny)
    dem = 1.0 + 0.005 * j + 0.3 * _bump(i, j, 10, 30, 25.0)
    h0 = 0.4 + 0.6 * _bump(i, j, 18, 14, 15.0)
    return dict(name="fixed_dt", nx=nx, ny=ny, dx=dx, fmt="ASC", dem=dem, h=h0, const_mann=0.03,
                time_step=0.25, fixed=1, courant=0.5, hextra=0.001, duration=20.0, print_interval=10.0)


CASES = [case_mound_asc, case_bin_varying_n, case_sources_dry, case_dambreak_frictionless, case_fixed_dt]


# ---- file writers (the reference app's input formats) --------------------------------
def _write_asc(path, a, header=None):
    with open(path, "w") as f:
        if header:
            for k, v in header:
                f.write(f"{k} {v}\n")
        for row in a:  # single spaces: Array::LoadAsciiFile splits on ' '
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


def _write_bin(path, a, header=None):
    with open(path, "wb") as f:
        if header:
            f.write(struct.pack("<6d", *header))
        f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def write_case(c: dict, d: str) -> str:
    """Writes the case's files into directory d; returns the cfg path."""
    os.makedirs(d, exist_ok=True)
    nx, ny, dx = c["nx"], c["ny"], c["dx"]
    xll, yll = c.get("xll", 0.0), c.get("yll", 0.0)
    ext = "asc" if c["fmt"] == "ASC" else "bin"
    cfg = {"input_format": c["fmt"], "const_mann": c["const_mann"], "time_step": c["time_step"],
           "time_increment_fixed": c["fixed"], "courant": c["courant"], "hextra": c["hextra"],
           "sim_start_time": 0, "sim_duration": c["duration"], "print_interval": c["print_interval"],
           "num_sources": len(c.get("sources", []))}
    dem_path = os.path.join(d, f"dem.{ext}")
    if c["fmt"] == "ASC":
        _write_asc(dem_path, c["dem"], [("ncols", nx), ("nrows", ny), ("xllcorner", xll), ("yllcorner", yll),
                                        ("cellsize", dx), ("NODATA_value", -9999)])
    else:
        _write_bin(dem_path, c["dem"], (nx, ny, xll, yll, dx, -9999))
    cfg["dem_filename"] = dem_path
    for key, field in (("h_infile", "h"), ("qx_infile", "qx"), ("qy_infile", "qy"), ("n_infile", "n")):
        if field in c:
            p = os.path.join(d, f"{field}.{ext}")
            (_write_asc if c["fmt"] == "ASC" else _write_bin)(p, c[field])
            cfg[key] = p
    if c.get("sources"):
        loc = os.path.join(d, "sources.txt")
        with open(loc, "w") as f:
            f.write("% x, y\n")
            for x, y in c["sources"]:
                f.write(f"{x!r},{y!r}\n")
        hyd = os.path.join(d, "hydrograph.txt")
        with open(hyd, "w") as f:
            f.write("% time (h), flow of each source (m3/s)\n")
            for row in c["hydrograph"]:
                f.write(",".join(repr(float(v)) for v in row) + "\n")
        cfg["src_loc_file"] = loc
        cfg["hydrograph_filename"] = hyd
    path = os.path.join(d, "case.cfg")
    with open(path, "w") as f:
        f.write("# generated case\n")
        for k, v in cfg.items():
            f.write(f"{k}={v}\n")
    return path


# ---- the arrays the reference app builds from those files ----------------------------
def padded_inputs(c: dict):
    """(h, qx, qy, z, n2) padded (ny+2, nx+2) FP64 arrays as the reference app holds them
    at the start of its step loop."""
    nx, ny = c["nx"], c["ny"]

    def pad(a, ring=0.0):
        p = np.full((ny + 2, nx + 2), ring)
        if a is not None:
            p[1:-1, 1:-1] = a
        return p

    z = pad(c["dem"], WALL)
    h = pad(c.get("h"))
    qx = pad(c.get("qx"))
    qy = pad(c.get("qy"))
    n = c["n"] if "n" in c else np.full((ny, nx), c["const_mann"])
    n2 = pad(n * n)
    # nin.SetBoundaryCopy(): each face's ghost copies the adjacent inner layer, faces in
    # the order LEFT, RIGHT, UP, DOWN (Grid.hpp:682-830), so corners take the y faces'
    n2[:, 0] = n2[:, 1]
    n2[:, -1] = n2[:, -2]
    n2[0, :] = n2[1, :]
    n2[-1, :] = n2[-2, :]
    return h, qx, qy, z, n2


def source_table(c: dict):
    """(cells, times_s, flows) as the reference app builds them (triton_appfis.hpp:
    214-290): cells are padded linear indices in increasing order, flows[r, k] the flow of
    the k-th of those cells at hydrograph row r, times converted from hours."""
    if not c.get("sources"):
        return np.zeros(0, np.int32), np.zeros(0), np.zeros((0, 0))
    nx, ny, dx = c["nx"], c["ny"], c["dx"]
    xll, yll = c.get("xll", 0.0), c.get("yll", 0.0)
    full = {}
    for k, (x, y) in enumerate(c["sources"]):
        col = math.ceil((x - xll) / dx) - 1
        row = math.ceil(ny - (y - yll) / dx) - 1
        full[row * nx + col] = k + 1  # a later source on the same cell replaces the earlier
    cells, number = [], []
    for idx in sorted(full):
        row, col = divmod(idx, nx)
        cells.append((col + 1) + (nx + 2) * (row + 1))
        number.append(full[idx])
    hyd = np.array(c["hydrograph"], dtype=np.float64)
    times = hyd[:, 0] * 3600.0
    flows = hyd[:, number]
    return np.array(cells, np.int32), times, np.ascontiguousarray(flows)

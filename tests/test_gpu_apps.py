"""The reference's five sample apps, built UNCHANGED against this repository's Appfis/
headers (make -C apps ref), run on the MI355X through the generic Runtime_HIP path, and
their printed results are checked against the reference's own serial programs (golden
fixtures) or the oracle:
  GameOfLife  live count            == reference serial program (exact)
  Heat3D      Max / Min             == reference serial program (6 printed digits)
  Himeno      Sum of p              ~= oracle (float sum, 1e-5 relative)
  FDTD3d      output/output.txt     ~= oracle reference functor (6 printed decimals)
  Triton      output/h.txt, q       ~= oracle step chain (6 printed decimals)
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "apps", "bin")
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_serial.json")))


def app(name):
    path = os.path.join(BIN, f"ref_{name}")
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (the reference sources were absent when apps were built)")
    return path


def run(args, cwd, timeout=120):
    print(f"[apps] {' '.join(args)}", flush=True)
    out = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.parametrize("case", GOLDEN["gol"], ids=lambda c: f"{c['nx']}x{c['ny']}x{c['iterations']}")
def test_reference_gol_app(tmp_path, case):
    out = run([app("GameOfLife"), f"-dimx={case['nx']}", f"-dimy={case['ny']}", f"-iteration={case['iterations']}",
               "-output"], tmp_path)
    assert int(re.search(r"Live Count: (\d+)", out).group(1)) == case["live_count"]


@pytest.mark.parametrize("case", GOLDEN["heat3d"], ids=lambda c: f"{c['nx']}x{c['ny']}x{c['nz']}x{c['iterations']}")
def test_reference_heat_app(tmp_path, case):
    out = run([app("HeatDiffusion"), f"-dimx={case['nx']}", f"-dimy={case['ny']}", f"-dimz={case['nz']}",
               f"-iteration={case['iterations']}", "-output"], tmp_path)
    assert float(re.search(r"Max: (\S+)", out).group(1)) == case["max"]
    assert float(re.search(r"Min: (\S+)", out).group(1)) == case["min"]


def test_reference_himeno_app(tmp_path):
    nx, ny, nz, it = 33, 17, 25, 12
    out = run([app("Himeno"), f"-dimx={nx}", f"-dimy={ny}", f"-dimz={nz}", f"-iteration={it}", "-output"], tmp_path)
    got = float(re.search(r"Sum: (\S+)", out).group(1))
    # himeno_appfis.hpp:142-155: p = (i+1)^2 / dimX^2 on the full grid, zero halo
    W, H, D = nx + 2, ny + 2, nz + 2
    p = np.zeros((D, H, W), np.float32)
    i = np.arange(nx)
    p[1:-1, 1:-1, 1:-1] = (((i + 1) * (i + 1)).astype(np.float32) / np.float32(nx * nx))[None, None, :]
    arrs = [p, np.zeros_like(p)]
    for v in [1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]:
        arrs.append(np.full_like(p, np.float32(v)))
    oracle.himeno(arrs, W, H, D, 0.8, it)
    ref = float(np.sum(arrs[it % 2][1:-1, 1:-1, 1:-1].astype(np.float64)))
    assert abs(got - ref) <= 1e-5 * abs(ref)


def test_reference_fdtd3d_app(tmp_path):
    nx, ny, nz, it = 12, 10, 9, 3
    run([app("FDTD3d"), f"-dimx={nx}", f"-dimy={ny}", f"-dimz={nz}", f"-iteration={it}", "-output"], tmp_path)
    got = np.loadtxt(tmp_path / "output" / "output.txt").reshape(nz, ny, nx)
    libc = ctypes.CDLL(None)
    libc.srand(0)
    vals = np.array([libc.rand() for _ in range(nx * ny * nz)], dtype=np.float32)
    init = (np.float32(0.0) + (vals / np.float32(2147483647.0)) * np.float32(1.0)).reshape(nz, ny, nx)
    a = np.zeros((nz + 8, ny + 8, nx + 8), np.float32)
    a[4:-4, 4:-4, 4:-4] = init
    ref = oracle.fdtd3d_ref(a, np.zeros_like(a), nx, ny, nz, 0.1, 4, it)[4:-4, 4:-4, 4:-4]
    assert np.max(np.abs(got - ref)) <= 5.1e-7 + 1e-6 * np.max(np.abs(ref))


def _write_triton_case(d, n, dx, sim_duration):
    z = np.zeros((n + 2, n + 2))
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dem = 1.0 + 0.5 * np.exp(-((i - n / 3) ** 2 + (j - n / 2) ** 2) / 30.0) + 0.01 * i
    h0 = np.maximum(1.6 - dem, 0.0) + 0.4 * np.exp(-((i - 2 * n / 3) ** 2 + (j - n / 3) ** 2) / 10.0)
    with open(d / "dem.asc", "w") as f:
        f.write(f"ncols {n}\nnrows {n}\nxllcorner 0\nyllcorner 0\ncellsize {dx}\nNODATA_value -9999\n")
        for row in dem:
            f.write(" ".join(f"{v:.17g}" for v in row) + "\n")
    with open(d / "h0.asc", "w") as f:
        for row in h0:
            f.write(" ".join(f"{v:.17g}" for v in row) + "\n")
    cfg = (f"dem_filename={d / 'dem.asc'}\nh_infile={d / 'h0.asc'}\ninput_format=ASC\nconst_mann=0.035\n"
           f"time_step=0.5\ntime_increment_fixed=0\ncourant=0.5\nhextra=0.001\nnum_sources=0\n"
           f"sim_start_time=0\nsim_duration={sim_duration}\nprint_interval={sim_duration / 2}\n")
    (d / "case.cfg").write_text(cfg)
    z[1:-1, 1:-1] = dem
    z[0, :] = z[-1, :] = z[:, 0] = z[:, -1] = 1e6
    h = np.zeros_like(z)
    h[1:-1, 1:-1] = h0
    return z, h


def test_reference_triton_app(tmp_path):
    n, dx, dur = 40, 10.0, 60.0
    z, h = _write_triton_case(tmp_path, n, dx, dur)
    out = run([app("Triton"), str(tmp_path / "case.cfg"), "-output"], tmp_path)
    steps = int(re.search(r"Number of iteration: (\d+)", out).group(1))
    qx, qy = np.zeros_like(h), np.zeros_like(h)
    n2 = np.full_like(h, 0.035 * 0.035)
    t, _, _ = oracle.flood_run(h, qx, qy, z, n2, n, n, dx, 0.5, 0.001, 0.5, dur / 2, 0.0, 0.0, steps)
    assert abs(t[-1] - dur) < 1e-9
    got_h = np.loadtxt(tmp_path / "output" / "h.txt")
    assert np.max(np.abs(got_h - h[1:-1, 1:-1])) <= 1.5e-6
    got_u = np.loadtxt(tmp_path / "output" / "u.txt")
    assert np.max(np.abs(got_u - qx[1:-1, 1:-1])) <= 1.5e-6
    assert np.max(np.abs(qx)) > 1e-3  # the water moved


@pytest.mark.parametrize("name,args", [
    ("heat2d", ["-dimx=300", "-dimy=211", "-iteration=37"]),
    ("gol2d", ["-dimx=510", "-dimy=257", "-iteration=41"]),
    ("fdtd3d", ["-dimx=70", "-dimy=33", "-dimz=41", "-iteration=5"]),
])
def test_api_apps_generic_path_matches_hand_kernels(tmp_path, name, args):
    """Config stencils written as user functors on the Appfis/ API (generic Execute
    path, device exchange + swap) against the hand-written kernels on the same grids."""
    path = os.path.join(BIN, name)
    assert os.path.exists(path), f"{path} not built (make -C apps)"
    out = run([path] + args, tmp_path)
    res = json.loads(out.strip().splitlines()[-1])
    assert res["bit_identical"] is True, res
    if name == "gol2d":
        nx, ny, gens = 510, 257, 41
        a = np.zeros((ny + 2, nx + 2), np.uint8)
        oracle.fill_bernoulli(a, 1, 1002, 0.25)
        ref = oracle.gol_u8(a, np.zeros_like(a), nx, ny, gens)
        assert res["live_count"] == int(ref[1:-1, 1:-1].sum())
    if name == "heat2d":
        nx, ny, it = 300, 211, 37
        a = np.zeros((ny + 2, nx + 2))
        oracle.fill_uniform(a, 1, 1001)
        ref = oracle.heat2d(a, np.zeros_like(a), nx, ny, 0.1, it)
        assert abs(res["checksum"] - ref[1:-1, 1:-1].sum()) <= 1e-9 * abs(ref.sum())

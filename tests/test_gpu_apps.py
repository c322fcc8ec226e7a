"""Apps written on the APPFIS API (apps/src, this repository's own code) run on the
MI355X:

  heat2d, gol2d, fdtd3d  the config stencils as user functors through the generic
                         Execute path (device exchange + swap), against the hand-written
                         kernels on the same grids and against the oracle;
  flood                  a Triton-format case (cfg, ASCII or binary DEM with its header,
                         initial fields, Manning map, point sources and hydrograph) read
                         through the Appfis Array loaders and run by the hand-written
                         flood kernels while simtime < sim_duration; its printed outputs
                         against the reference Triton app's own (tests/golden/flood_ref.npz).

The reference's own apps are only compiled unchanged against Appfis/, where the
reference sources are (tests/test_apps_build.py): nothing built from the reference runs here.
"""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "apps", "bin")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import flood_cases  # noqa: E402

FLOOD_REF = np.load(os.path.join(ROOT, "tests", "golden", "flood_ref.npz"))


def run(args, cwd, timeout=120):
    print(f"[apps] {' '.join(args)}", flush=True)
    out = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.parametrize("variant", ["chosen", "plain"])
@pytest.mark.parametrize("make", flood_cases.CASES, ids=lambda m: m.__name__[5:])
def test_flood_app_reproduces_reference_triton_app(tmp_path, make, variant):
    c = make()
    cfg = flood_cases.write_case(c, str(tmp_path))
    exe = os.path.join(BIN, "flood")
    assert os.path.exists(exe), f"{exe} not built (make -C apps)"
    out = run([exe, cfg, "-output"] + (["-plain"] if variant == "plain" else []), tmp_path)
    assert int(re.search(r"Number of iteration: (\d+)", out).group(1)) == int(FLOOD_REF[c["name"] + "_steps"])
    for key in ("h", "u", "v"):
        with open(tmp_path / "output" / f"{key}.txt") as f:
            rows = [line.split() for line in f if line.strip()]
        got = np.array([[int(t.replace(".", "")) for t in row] for row in rows], dtype=np.int64)
        ref = FLOOD_REF[f"{c['name']}_{key}"]
        assert got.shape == ref.shape, key
        # printed with six decimals like the reference; the FP64 device result is within
        # 1e-9 of the reference's, so a printed digit can differ by one where the value
        # sits on a rounding boundary
        diff = np.abs(got - ref)
        assert diff.max() <= 1, (key, int(diff.max()))
        assert np.count_nonzero(diff) <= max(2, diff.size // 1000), (key, int(np.count_nonzero(diff)))


def _oracle_f32_h(c):
    """The oracle's FP32 restatement (correctly rounded FP32 operations) of a case: the
    error that FP32 arithmetic itself makes there."""
    h, qx, qy, z, n2 = (a.astype(np.float32) for a in flood_cases.padded_inputs(c))
    cells, times, flows = flood_cases.source_table(c)
    oracle.flood_run_src(h, qx, qy, z, n2, c["nx"], c["ny"], c["dx"], c["courant"], c["hextra"], c["time_step"],
                         c["print_interval"], 0.0, 0.0, int(FLOOD_REF[c["name"] + "_steps"]), fixed_dt=c["fixed"],
                         src_cell=cells, src_time=times, src_flow=flows)
    return h[1:-1, 1:-1].astype(np.float64)


@pytest.mark.parametrize("make", flood_cases.CASES, ids=lambda m: m.__name__[5:])
def test_flood_app_fp32_against_reference_triton_app(tmp_path, make):
    # the FP32 bulk form of the same driver (binary inputs are doubles, rounded on load)
    # against the reference app's printed FP64 depths.  Every value is finite, and the
    # relative L2 error of h is at most the FP32 form's stated 1e-3 or twice what the
    # oracle's correctly rounded FP32 restatement makes on the case, whichever is larger
    # (FP32 itself is off by 2e-2 on the dry start fed by point sources, where depths
    # hover around the 1 mm wet/dry threshold)
    c = make()
    cfg = flood_cases.write_case(c, str(tmp_path))
    exe = os.path.join(BIN, "flood")
    assert os.path.exists(exe), f"{exe} not built (make -C apps)"
    run([exe, cfg, "-output", "-fp32"], tmp_path)
    with open(tmp_path / "output" / "h.txt") as f:
        got = np.array([[float(t) for t in line.split()] for line in f if line.strip()])
    ref = FLOOD_REF[f"{c['name']}_h"] * 1e-6
    assert got.shape == ref.shape and np.isfinite(got).all()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    rel_f32 = np.linalg.norm(_oracle_f32_h(c) - ref) / np.linalg.norm(ref)
    print(f"[apps] {c['name']} fp32 h rel-L2 {rel:.2e} (oracle FP32 {rel_f32:.2e})")
    assert rel <= max(1e-3, 2.0 * rel_f32), (rel, rel_f32)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("make", [m for m in flood_cases.CASES if not m().get("sources")],
                         ids=lambda m: m.__name__[5:])
def test_flood_app_ensemble_copies_match_single_run(tmp_path, make, precision):
    # APPFIS::kernels::FloodEnsemble: three copies of a case advanced together end
    # identical to the single run of the same case (the driver compares the grids)
    c = make()
    cfg = flood_cases.write_case(c, str(tmp_path))
    exe = os.path.join(BIN, "flood")
    assert os.path.exists(exe), f"{exe} not built (make -C apps)"
    out = run([exe, cfg, "-ensemble=3"] + (["-fp32"] if precision == "fp32" else []), tmp_path)
    assert "Ensemble of 3 copies: identical to the single run" in out, out


@pytest.mark.parametrize("name,args", [
    ("heat2d", ["-dimx=300", "-dimy=211", "-iteration=37"]),
    ("gol2d", ["-dimx=510", "-dimy=257", "-iteration=41", "-output"]),
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
        # uint8 Gather / Scatter and the binary file round trip (single rank): the file
        # written from the gathered full grid is the oracle's interior
        assert res["scatter_gather_u8"] is True, res
        got = np.fromfile(tmp_path / "output" / "gol.bin", dtype=np.uint8).reshape(ny, nx)
        assert np.array_equal(got, ref[1:-1, 1:-1])
    if name == "heat2d":
        nx, ny, it = 300, 211, 37
        a = np.zeros((ny + 2, nx + 2))
        oracle.fill_uniform(a, 1, 1001)
        ref = oracle.heat2d(a, np.zeros_like(a), nx, ny, 0.1, it)
        assert abs(res["checksum"] - ref[1:-1, 1:-1].sum()) <= 1e-9 * abs(ref.sum())

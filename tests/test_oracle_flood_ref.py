"""The flood oracle against the reference Triton app itself, and its Manning friction
against the law it implements.

tests/golden/flood_ref.npz holds the reference app's printed outputs (six decimals) on
the cases of tests/golden/flood_cases.py: ASC and binary inputs, a spatially varying
Manning map, initial momentum, wet/dry fronts, two hydrograph point sources, a
frictionless dam break and the fixed-increment mode (tests/golden/make_flood_golden.py
wrote it from the reference sources built in oracle/_ref).  The oracle, run for the
printed number of steps, must land on the printed time and reproduce every printed value
to its last digit; a 1 % error in the friction coefficient planted in the oracle's input
must fail that comparison.

The friction law: Manning's friction slope S_f = n^2 |u| u / h^(4/3) gives the momentum
sink g h S_f = g n^2 |q| q / h^(7/3), and the reference solves it implicitly (backward
Euler, triton_appfis.hpp:1046-1092): q1 + dt g n^2 |q1| q1 / h^(7/3) = m, with m the
momentum after the fluxes.  On a flat bed with a uniform state the fluxes cancel, m is the
old momentum, and q1 must solve that equation, parallel to m, for any depth: a wrong
exponent, coefficient or norm fails it.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import flood_cases  # noqa: E402

REF = np.load(os.path.join(HERE, "golden", "flood_ref.npz"))
PRINT_HALF_ULP = 5e-7  # the printed value is the true value rounded to six decimals
G = 9.81


def run_oracle(c, n2_scale=1.0):
    h, qx, qy, z, n2 = flood_cases.padded_inputs(c)
    cells, times, flows = flood_cases.source_table(c)
    steps = int(REF[c["name"] + "_steps"])
    t, _, _ = oracle.flood_run_src(h, qx, qy, z, n2 * n2_scale, c["nx"], c["ny"], c["dx"], c["courant"], c["hextra"],
                                   c["time_step"], c["print_interval"], 0.0, 0.0, steps, fixed_dt=c["fixed"],
                                   src_cell=cells, src_time=times, src_flow=flows)
    return t, {"h": h, "u": qx, "v": qy}


def worst(c, fields):
    """Largest distance of a field from the printed reference value, in units of the
    printing half-ulp (<= 1 means every value prints as the reference printed it)."""
    return max(np.abs(fields[k][1:-1, 1:-1] - REF[f"{c['name']}_{k}"] * 1e-6).max() for k in ("h", "u", "v")) \
        / PRINT_HALF_ULP


@pytest.mark.parametrize("make", flood_cases.CASES, ids=lambda m: m.__name__[5:])
def test_oracle_reproduces_reference_triton_app(make):
    c = make()
    t, fields = run_oracle(c)
    # the reference loops while simtime < sim_duration; its step count lands exactly there
    assert t[-1] == pytest.approx(c["duration"], abs=1e-9)
    if len(t) > 1:
        assert t[-2] < c["duration"]
    assert worst(c, fields) <= 1.0 + 1e-6
    # the case really moves water (nothing trivially zero)
    assert np.abs(fields["h"]).max() > 0.1 and np.abs(fields["u"]).max() > 1e-3


@pytest.mark.parametrize("make", [flood_cases.case_mound_asc, flood_cases.case_bin_varying_n,
                                  flood_cases.case_sources_dry, flood_cases.case_fixed_dt],
                         ids=lambda m: m.__name__[5:])
def test_reference_fixtures_detect_a_friction_error(make):
    """A 1 % error in n^2, planted in the oracle's input, moves values far past the
    printing precision: the fixtures pin the friction term."""
    c = make()
    _, fields = run_oracle(c, n2_scale=1.01)
    assert worst(c, fields) > 10.0


@pytest.mark.parametrize("make", flood_cases.CASES, ids=lambda m: m.__name__[5:])
def test_reference_cases_are_well_conditioned(make):
    """Relative perturbations of a few ulps in n^2 and the DEM (what a device cube root or
    a different rounding does to the run) move no value by more than 1e-12: no branch of
    the solver sits within rounding of its threshold, so agreement to the printed digit
    and to 1e-9 is a fair bar for the GPU."""
    c = make()
    steps = int(REF[c["name"] + "_steps"])
    cells, times, flows = flood_cases.source_table(c)
    runs = []
    for eps in (0.0, 1e-15, -1e-15, 3e-15, 1e-14, -1e-14, 1e-13):
        h, qx, qy, z, n2 = flood_cases.padded_inputs(c)
        z[1:-1, 1:-1] *= 1.0 + eps
        oracle.flood_run_src(h, qx, qy, z, n2 * (1.0 + eps), c["nx"], c["ny"], c["dx"], c["courant"], c["hextra"],
                             c["time_step"], c["print_interval"], 0.0, 0.0, steps, fixed_dt=c["fixed"],
                             src_cell=cells, src_time=times, src_flow=flows)
        runs.append((h, qx, qy))
    worst = max(float(np.abs(a - b).max()) for r in runs[1:] for a, b in zip(runs[0], r))
    assert worst <= 1e-12


def test_fixture_case_inputs_round_trip_through_the_file_formats(tmp_path):
    """The cases' files encode what padded_inputs / source_table hand to the oracle: the
    ASC and binary DEM headers and values, and the source cells of calc_src_col/row."""
    c = flood_cases.case_bin_varying_n()
    flood_cases.write_case(c, str(tmp_path))
    raw = np.fromfile(tmp_path / "dem.bin", dtype="<f8")
    assert list(raw[:6]) == [c["nx"], c["ny"], 1000.0, 2000.0, c["dx"], -9999]
    assert np.array_equal(raw[6:].reshape(c["ny"], c["nx"]), c["dem"])
    assert np.array_equal(np.fromfile(tmp_path / "n.bin", dtype="<f8").reshape(c["ny"], c["nx"]), c["n"])
    s = flood_cases.case_sources_dry()
    cells, times, flows = flood_cases.source_table(s)
    W = s["nx"] + 2
    # sources placed at full-grid (col, row) = (30, 14) and (12, 20): padded (31, 15), (13, 21)
    assert sorted(cells.tolist()) == sorted([31 + W * 15, 13 + W * 21])
    assert cells.tolist() == sorted(cells.tolist())
    assert times.tolist() == [0.0, 36.0, 108.0, 180.0]
    # flows follow the cell order: cell (31, 15) is source 1, (13, 21) source 2
    assert flows[:, 0].tolist() == [0.0, 8.0, 15.0, 2.0] and flows[:, 1].tolist() == [0.0, 3.0, 10.0, 12.0]


def _uniform_sheet(H, Qx, Qy, n, nx=24, ny=16):
    z = np.zeros((ny + 2, nx + 2))
    z[0, :] = z[-1, :] = z[:, 0] = z[:, -1] = flood_cases.WALL
    h = np.zeros_like(z)
    h[1:-1, 1:-1] = H
    qx = np.zeros_like(z)
    qx[1:-1, 1:-1] = Qx
    qy = np.zeros_like(z)
    qy[1:-1, 1:-1] = Qy
    n2 = np.full_like(z, n * n)
    return h, qx, qy, z, n2


MANNING_CASES = [(0.2, 0.05, -0.03), (0.7, 0.4, 0.25), (2.5, -1.5, 2.0)]


def manning_residuals(H, Qx, Qy, n, dt, h1, qx1, qy1):
    """Relative residuals of the implicit Manning equation and of parallelism on the
    cells whose four neighbours hold the same state (two cells from every wall)."""
    sl = (slice(2, -2), slice(2, -2))
    q0 = math.hypot(Qx, Qy)
    a = dt * G * n * n / H ** (7.0 / 3.0)
    q1x, q1y = qx1[sl], qy1[sl]
    q1 = np.hypot(q1x, q1y)
    law = np.abs(q1 + a * q1 * q1 - q0) / q0
    par = np.abs(q1x * Qy - q1y * Qx) / (q0 * q0)
    depth = np.abs(h1[sl] - H) / H
    return law.max(), par.max(), depth.max(), float(np.median(q1 / q0))


@pytest.mark.parametrize("H,Qx,Qy", MANNING_CASES)
def test_oracle_manning_friction_solves_the_implicit_law(H, Qx, Qy):
    n, dt = 0.05, 0.5
    h, qx, qy, z, n2 = _uniform_sheet(H, Qx, Qy, n)
    oracle.flood_run_src(h, qx, qy, z, n2, 24, 16, 10.0, 0.5, 1e-3, dt, 0.0, 0.0, 0.0, 1, fixed_dt=True)
    law, par, depth, ratio = manning_residuals(H, Qx, Qy, n, dt, h, qx, qy)
    assert law <= 1e-13 and par <= 1e-14 and depth <= 1e-14
    assert ratio < 0.999  # the friction acted (the check is not vacuous)
    # a planted wrong exponent (h^2 instead of h^(7/3)) fails the same check
    a_wrong = dt * G * n * n / H ** 2
    q0 = math.hypot(Qx, Qy)
    q1 = np.hypot(qx[2:-2, 2:-2], qy[2:-2, 2:-2])
    assert np.abs(q1 + a_wrong * q1 * q1 - q0).max() / q0 > 1e-6

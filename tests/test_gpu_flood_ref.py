"""The HIP flood kernels against the reference Triton app's own outputs, and the Manning
friction of the device against the implicit law (see tests/test_oracle_flood_ref.py for
the fixtures and the law).

appfis_flood_run_src in FP64 (the reference dtype), with direct launches (PLAIN) and
graph-captured chunks (CHOSEN), must reproduce every value the reference app printed to
its last digit on every case -- ASC and binary inputs, varying Manning map, point
sources, dam break, fixed increment -- and agree with the oracle to 1e-9 of the field's
scale (the device's cube root may differ from libm's by an ulp; without friction, where
only correctly rounded operations remain, it must be bit-identical).
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle
from appfis_amd import hip

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
sys.path.insert(0, HERE)
import flood_cases  # noqa: E402
from test_oracle_flood_ref import MANNING_CASES, PRINT_HALF_ULP, REF, _uniform_sheet, manning_residuals  # noqa: E402

pytestmark = pytest.mark.gpu


def params(c, until=False):
    return hip.FloodParams(c["dx"], c["courant"], c["hextra"], c["time_step"], c["print_interval"], 0.0, 0.0,
                           int(c["fixed"]), c["duration"] if until else 0.0)


def run_gpu(ctx, c, variant, until):
    """until: the reference's loop (while simtime < sim_duration) through params.end_time,
    with room for ten times the steps it needs; else exactly the reference's step count."""
    h, qx, qy, z, n2 = flood_cases.padded_inputs(c)
    cells, times, flows = flood_cases.source_table(c)
    steps = int(REF[c["name"] + "_steps"])
    d = [ctx.from_numpy(a) for a in (h, qx, qy, z, n2)]
    print(f"[flood-ref] {c['name']} {c['nx']}x{c['ny']} steps {steps} variant {variant} sources {len(cells)} "
          f"until {until}", flush=True)
    t, _ = ctx.flood_run_src(*d, c["nx"], c["ny"], params(c, until), 10 * steps if until else steps, variant,
                             cells, times, flows)
    assert len(t) == steps
    return t, {"h": d[0].download(), "u": d[1].download(), "v": d[2].download()}


@pytest.mark.parametrize("until", [False, True], ids=["steps", "until"])
@pytest.mark.parametrize("variant", [hip.PLAIN, hip.CHOSEN], ids=["plain", "graph"])
@pytest.mark.parametrize("make", flood_cases.CASES, ids=lambda m: m.__name__[5:])
def test_gpu_fp64_reproduces_reference_triton_app(ctx, make, variant, until):
    c = make()
    t, got = run_gpu(ctx, c, variant, until)
    assert t[-1] == pytest.approx(c["duration"], abs=1e-9)
    for k in ("h", "u", "v"):
        printed = REF[f"{c['name']}_{k}"] * 1e-6
        assert np.abs(got[k][1:-1, 1:-1] - printed).max() <= PRINT_HALF_ULP + 1e-9, k
    # and the oracle, unrounded
    h, qx, qy, z, n2 = flood_cases.padded_inputs(c)
    cells, times, flows = flood_cases.source_table(c)
    oracle.flood_run_src(h, qx, qy, z, n2, c["nx"], c["ny"], c["dx"], c["courant"], c["hextra"], c["time_step"],
                         c["print_interval"], 0.0, 0.0, len(t), fixed_dt=c["fixed"], src_cell=cells, src_time=times,
                         src_flow=flows)
    ref = {"h": h, "u": qx, "v": qy}
    for k in ("h", "u", "v"):
        if c["const_mann"] == 0.0 and "n" not in c:
            assert np.array_equal(got[k], ref[k]), k  # no cube root anywhere: bit-identical
        else:
            scale = max(1.0, float(np.abs(ref[k]).max()))
            assert np.abs(got[k] - ref[k]).max() <= 1e-9 * scale, k


@pytest.mark.parametrize("H,Qx,Qy", MANNING_CASES)
def test_gpu_manning_friction_solves_the_implicit_law(ctx, H, Qx, Qy):
    n, dt = 0.05, 0.5
    arrs = _uniform_sheet(H, Qx, Qy, n)
    d = [ctx.from_numpy(a) for a in arrs]
    prm = hip.FloodParams(10.0, 0.5, 1e-3, dt, 0.0, 0.0, 0.0, 1)
    ctx.flood_run_src(*d, 24, 16, prm, 1, hip.PLAIN)
    law, par, depth, ratio = manning_residuals(H, Qx, Qy, n, dt, d[0].download(), d[1].download(), d[2].download())
    print(f"[flood-ref] Manning H={H}: law {law:.2e} parallel {par:.2e} depth {depth:.2e} q1/q0 {ratio:.6f}",
          flush=True)
    assert law <= 1e-13 and par <= 1e-14 and depth <= 1e-14
    assert ratio < 0.999
    assert math.isfinite(ratio)


def test_gpu_sources_reject_bad_cells(ctx):
    c = flood_cases.case_sources_dry()
    h, qx, qy, z, n2 = flood_cases.padded_inputs(c)
    d = [ctx.from_numpy(a) for a in (h, qx, qy, z, n2)]
    _, times, flows = flood_cases.source_table(c)
    W = c["nx"] + 2
    for cells in ([0, 5 + W * 5], [5 + W * 5, 5 + W * 5]):  # a ghost cell; a repeated cell
        with pytest.raises(hip.AppfisError):
            ctx.flood_run_src(*d, c["nx"], c["ny"], params(c), 2, hip.PLAIN, cells, times, flows)

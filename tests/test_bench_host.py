"""Host logic of bench.py (no GPU): the seeded workload sizes, the roofline arithmetic,
the kernel-name normalisation of the profile summaries, and the one-GPU refusal."""
import json
import os
import subprocess
import sys

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_flood_sizes_follow_the_oracle_generator():
    # N_s = 512 + floor(u_s * 3585), u_s = (r(1005, s) >> 11) * 2^-53
    sizes = bench.flood_sizes(24, 1005)
    assert len(sizes) == 24 and all(512 <= n <= 4096 for n in sizes)
    for s, n in enumerate(sizes):
        u = (oracle.rng_u64(1005, s) >> 11) * 2.0 ** -53
        assert n == 512 + int(u * 3585)
    assert bench._splitmix_sides(24, 1005, 512, 4096) == sizes


def test_heat_ensemble_sides_in_range_and_seeded():
    a = bench._splitmix_sides(64, 1006, 128, 1024)
    assert a == bench._splitmix_sides(64, 1006, 128, 1024)
    assert min(a) >= 128 and max(a) <= 1024 and len(set(a)) > 50


def test_roofline_arithmetic():
    r = bench.roofline(8e9, 2.0, "k")  # 8 GB in 2 ms = 4000 GB/s
    assert r["achieved"] == 4000.0 and r["frac"] == 0.5 and r["peak"] == 8000.0
    assert r["unit"] == "GB/s" and r["bound"] == "hbm" and r["kernel"] == "k"


def test_kernel_name_normalisation():
    s = bench._short
    assert s("void (anonymous namespace)::k_fdtd3d<float, 1, 8, 4, 256, 2, 2>(float const*, float*, int)") == \
        "k_fdtd3d<float, 1, 8, 4, 256, 2, 2>"
    assert s("void gol::k_gol_plain<2, 16>(unsigned char const*, unsigned char*, int, int, int, int)") == \
        "k_gol_plain<2, 16>"
    assert s("void golbits::k_step_tb<10, 4, 16>(unsigned int const*, unsigned int*, int)") == "k_step_tb<10, 4, 16>"
    assert s("void (anonymous namespace)::k_himeno<double, 1, 2, 1>(double const*, double*, (anonymous "
             "namespace)::Coef, int)") == "k_himeno<double, 1, 2, 1>"
    assert s("(anonymous namespace)::k_add_correction(double const*, float const*)") == "k_add_correction"
    assert s("k_copy16(uint4 const*, uint4*, unsigned long)") == "k_copy16"


def test_traffic_lookup_weights(tmp_path, monkeypatch):
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "traffic.json").write_text(json.dumps({"a": {"bytes_per_launch": 100.0},
                                                   "b": {"bytes_per_launch": 10.0}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench._traffic({"a": 1.0}) == 100.0
    assert bench._traffic({"a": 0.5, "b": 2.0}) == 70.0
    assert bench._traffic({"missing": 1.0}) is None


def test_valu_lookup(tmp_path, monkeypatch):
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "traffic.json").write_text(json.dumps({
        "_calibration": {"fetch_scale": 2.0},
        "a": {"bytes_per_launch": 1.0, "valu_busy_pct": 75.5, "valu_wave_insts_per_launch": 1.0e6},
        "b": {"bytes_per_launch": 1.0, "valu_wave_insts_per_launch": 2.0e6},
        "c": {"bytes_per_launch": 1.0}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    # (1e6 + 2 * 2e6) wave instructions x 64 lanes in 0.1 ms = 3.2 T lane-instr/s
    v = bench._valu({"a": 1, "b": 2}, 0.1)
    assert v["wave_insts_per_unit"] == 5.0e6 and v["achieved"] == 3.2 and v["valu_busy"] == {"a": 0.755}
    assert v["frac"] == round(3.2 / bench.VALU_PEAK_TOPS, 4)
    assert bench._valu({"a": 1, "c": 1}, 0.1) is None
    r = bench.roofline(1e9, 1.0, {"a": 1})
    assert r["valu"]["achieved"] == 0.06 and r["traffic"] == 1.0


def test_gol_kernel_names_match_the_profiled_symbols():
    names = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    for tname, want in (("tw4 k4 g8 w12", "k_step_tw<4, 4, 8, 12>"), ("tb1 k12 g8 w16", "k_step_tb1<12, 8, 16>"),
                        ("tb k10 g4 w16", "k_step_tb<10, 4, 16>")):
        assert bench.gol_kernel_name(tname) == want
        assert want in names, want
    # widths without a blocked candidate run the one-generation kernel (not profiled at 8192)
    assert bench.gol_kernel_name(None) == "k_step<8>"


def test_cpu_baseline_reports_median_of_runs_and_the_machine():
    r = bench.cpu_baseline_gol(nx=128, ny=96, target_s=0.2, runs=3)
    assert len(r["runs"]) == 3 and r["value"] == round(sorted(r["runs"])[1], 4)
    assert r["kind"] == "port" and r["cores"] >= 1 and r["cpu_model"]
    assert set(r["omp_env"]) == {"OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES"}


def test_heat_kernel_names():
    assert bench.heat_kernel_name("tb 64x64 ts12 nt1024") == "k_heat2d_tb<double, 64, 64, 12, 1024>"
    assert bench.heat_kernel_name("rt 48x96 ts8 k7 w16") == "k_heat2d_rt<double, 7, 16, 8, 1>"
    assert bench.heat_kernel_name("rt2 112x48 ts8 k8 w8") == "k_heat2d_rt<double, 8, 8, 8, 2>"
    assert bench._short("void (anonymous namespace)::k_heat2d_rt<double, 7, 16, 8, 1>(double const*, double*, int, "
                        "int, double, int)") == "k_heat2d_rt<double, 7, 16, 8, 1>"

"""CPU checks of the C ABI library: it loads and exports every entry point that
include/appfis_hip.h declares (no compute call is made without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "appfis_hip.h")
LIB = os.path.join(ROOT, "appfis_amd", "libappfis_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(appfis_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("appfis_ctx_create", "appfis_heat2d", "appfis_gol", "appfis_fdtd3d", "appfis_himeno",
                 "appfis_flood_run", "appfis_flood_ensemble_run", "appfis_reduce", "appfis_exchange",
                 "appfis_set_boundary", "appfis_set_boundary_copy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (appfis_[a-z0-9_]+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_version():
    from appfis_amd import hip

    L = hip.lib()
    assert L.appfis_abi_version() == 2
    assert L.appfis_status_string(0) == b"ok"
    assert L.appfis_status_string(1) == b"invalid argument"


def test_invalid_arguments_are_reported_not_fatal():
    import ctypes as C

    from appfis_amd import hip

    L = hip.lib()
    # NULL context / bad shapes return APPFIS_ERR_INVALID with a message; nothing exits
    assert L.appfis_sync(None) == 1
    assert b"ctx" in L.appfis_last_error()
    dims = (C.c_int * 2)(4, 4)
    assert L.appfis_exchange(None, None, 0, 2, dims, 1, 1) == 1
    assert L.appfis_heat2d(None, 3, None, None, 4, 4, 0.1, 1, 0, 0) == 1


def test_header_is_plain_c_and_links(tmp_path):
    # the boundary is usable from C (and so from any C FFI): the header compiles as C99
    # with warnings as errors, and a C program links against the library and calls the
    # context-free entry points (no GPU needed)
    src = tmp_path / "abi_c.c"
    src.write_text(r'''
#include <stdio.h>
#include <string.h>
#include "appfis_hip.h"
int main(void) {
    char names[4096];
    if (appfis_abi_version() != APPFIS_ABI_VERSION) return 1;
    if (strcmp(appfis_status_string(APPFIS_OK), "ok") != 0) return 2;
    if (appfis_fdtd3d_candidates(APPFIS_F32, names, (int)sizeof names) < 2) return 3;
    if (appfis_heat2d_candidates(APPFIS_F64, names, (int)sizeof names) < 2) return 4;
    if (appfis_gol_candidates(names, (int)sizeof names) < 2) return 5;
    if (appfis_sync(NULL) != APPFIS_ERR_INVALID) return 6;
    printf("ok %s\n", appfis_last_error());
    return 0;
}
''')
    exe = tmp_path / "abi_c"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.dirname(HEADER),
                    str(src), "-o", str(exe), "-L", os.path.dirname(LIB), "-lappfis_hip",
                    "-Wl,-rpath," + os.path.dirname(LIB)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.startswith("ok")


def test_build_keeps_fp_contraction_off_except_the_fp32_flood_form():
    # the parity contract: every HIP object is compiled without FMA contraction (the
    # reference functors' rounding), except the FP32 bulk flood form, whose results are
    # held to a tolerance against FP64 and which contracts its multiply-adds
    out = subprocess.run(["make", "-n", "-B", "-C", os.path.join(ROOT, "appfis_amd", "csrc")], capture_output=True,
                         text=True, check=True).stdout
    compiles = [l for l in out.splitlines() if "hipcc" in l and " -c " in l]
    srcs = {l.split(" -c ")[1].split()[0]: l for l in compiles}
    assert "flood.hip" in srcs and "flood_f32.hip" in srcs and len(srcs) >= 10, sorted(srcs)
    for src, line in srcs.items():
        assert "-Xclang -target-feature -Xclang -packed-fp32-ops" in line, src
        if src == "flood_f32.hip":
            assert "-ffp-contract=fast" in line and "-ffp-contract=off" not in line, line
        else:
            assert "-ffp-contract=off" in line and "-ffp-contract=fast" not in line, line

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
    assert L.appfis_abi_version() == 1
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

"""Compile check of the drop-in claim: the reference's five sample apps build UNCHANGED
against this repository's Appfis/ headers (make -C apps ref: their sources are linked
from the reference tree, never copied, into a temporary directory, and compiled there
with hipcc for gfx950).  Runs only where the reference sources are present (the build
container); nothing it builds is kept or sent to the GPU machine.  On the GPU the
generic Execute path is checked with this repository's own functor apps
(tests/test_gpu_apps.py)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
APPS = ["HeatDiffusion", "GameOfLife", "FDTD3d", "Himeno", "Triton"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not on this machine")
def test_reference_apps_build_unchanged_against_appfis_headers(tmp_path):
    lib = os.path.join(ROOT, "appfis_amd", "libappfis_hip.so")
    if not os.path.exists(lib):
        pytest.fail(f"{lib} not built (run __graft_entry__.build())")
    r = subprocess.run(["make", "-s", "-j5", "-C", os.path.join(ROOT, "apps"), "ref", f"REFBUILD={tmp_path}/src",
                        f"REFBIN={tmp_path}/bin"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    for app in APPS:
        exe = tmp_path / "bin" / f"ref_{app}"
        assert exe.exists() and os.access(exe, os.X_OK), app
        # the sources compiled are the reference's own files, linked, not copies
        main = tmp_path / "src" / app / "src" / "main.cpp"
        assert main.is_symlink() and os.path.realpath(main) == os.path.join(REF, app, "src", "main.cpp")

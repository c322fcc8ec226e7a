"""Writes tests/golden/flood_ref.npz: the reference Triton app's own outputs on the small
cases of flood_cases.py.

The app is the reference's Triton/src/main.cpp on the reference's APPFIS OpenMP runtime,
compiled from its sources by `make -C oracle ref` into oracle/_ref/triton_acpu (so this
runs only where the reference sources are; the fixture it writes is what the tests
read).  It runs as an MPI singleton with `-output`, which writes h, u (= qx) and v (= qy)
of the interior with
six decimals (Array::SaveAsciiFile, Appfis/Array.hpp:447-471) and prints the number of
steps.  The fixture stores those printed values as integers in units of 1e-6 (exactly
what was printed) and the step count per case.

    python tests/golden/make_flood_golden.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import flood_cases  # noqa: E402

APP = os.path.join(ROOT, "oracle", "_ref", "triton_acpu")
OUT = os.path.join(HERE, "flood_ref.npz")


def run_case(c, d):
    cfg = flood_cases.write_case(c, d)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([APP, cfg, "-output", "-thread=4"], cwd=d, capture_output=True, text=True, timeout=600,
                       env=env)
    if r.returncode != 0:
        raise RuntimeError(f"{c['name']}: reference app failed\n{r.stdout}\n{r.stderr}")
    steps = int(re.search(r"Number of iteration: (\d+)", r.stdout).group(1))
    fields = {}
    for key in ("h", "u", "v"):
        with open(os.path.join(d, "output", f"{key}.txt")) as f:
            rows = [line.split() for line in f if line.strip()]
        assert len(rows) == c["ny"] and all(len(row) == c["nx"] for row in rows), key
        # each token has exactly six decimals: its value in micro-units is an integer
        fields[key] = np.array([[int(t.replace(".", "")) for t in row] for row in rows], dtype=np.int64)
    return steps, fields


def main():
    if not os.path.exists(APP):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    data = {}
    with tempfile.TemporaryDirectory() as tmp:
        for make in flood_cases.CASES:
            c = make()
            steps, fields = run_case(c, os.path.join(tmp, c["name"]))
            data[f"{c['name']}_steps"] = np.array(steps, dtype=np.int64)
            for key, a in fields.items():
                data[f"{c['name']}_{key}"] = a
            print(f"{c['name']}: {steps} steps, max |h| {np.abs(fields['h']).max() * 1e-6:.6f}, "
                  f"max |qx| {np.abs(fields['u']).max() * 1e-6:.6f}, max |qy| {np.abs(fields['v']).max() * 1e-6:.6f}")
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()

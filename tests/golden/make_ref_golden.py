"""Writes tests/golden/ref_serial.json from the reference's own serial programs.

The programs are compiled from the reference sources by `make -C oracle ref`
(outputs in oracle/_ref/, never committed); this script runs them and records what
they print:
  - GameOfLife/src/game_serial.hpp : "Live Count: N" (srand(0); rand()%2 init,
    periodic, the reference rule with stale writes)
  - HeatDiffusion/src/heat_serial.hpp : "Max: / Min:" (Gaussian init, 7-point FTCS,
    cx = cy = cz = 0.1, Dirichlet 0), printed with 6 significant digits

Run: python tests/golden/make_ref_golden.py   (needs the reference at /root/reference)
"""
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_serial.json")

GOL_CASES = [(64, 64, 10), (512, 512, 50), (100, 60, 37), (33, 17, 101), (256, 128, 200)]
HEAT_CASES = [(16, 16, 16, 10), (24, 20, 12, 25), (32, 32, 32, 40)]


def run(args):
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    return out


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    data = {"source": "reference *_serial.hpp programs compiled from source (oracle/_ref)", "gol": [], "heat3d": []}
    for nx, ny, it in GOL_CASES:
        out = run([os.path.join(REF_BIN, "game2d_serial"), f"-dimx={nx}", f"-dimy={ny}", f"-iteration={it}",
                   "-output"])
        live = int(re.search(r"Live Count: (\d+)", out).group(1))
        data["gol"].append({"nx": nx, "ny": ny, "iterations": it, "live_count": live})
    for nx, ny, nz, it in HEAT_CASES:
        out = run([os.path.join(REF_BIN, "heat3d_serial"), f"-dimx={nx}", f"-dimy={ny}", f"-dimz={nz}",
                   f"-iteration={it}", "-output"])
        mx = float(re.search(r"Max: (\S+)", out).group(1))
        mn = float(re.search(r"Min: (\S+)", out).group(1))
        data["heat3d"].append({"nx": nx, "ny": ny, "nz": nz, "iterations": it, "max": mx, "min": mn})
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps(data, indent=1))


if __name__ == "__main__":
    sys.exit(main())

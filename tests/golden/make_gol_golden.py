"""Writes tests/golden/gol_8192.json: the oracle's GameOfLife on the full 8192 x 8192
workload (uint8 cells, Bernoulli p = 0.25 from splitmix64 seed 1002, periodic, B3/S23)
at every 100th generation up to 1000 -- the SHA-256 of the interior bytes (row-major,
x fastest) and the live-cell count.

The GPU test runs both HIP forms (byte kernel and bit-packed kernel) in chunks of 100
generations and checks each chunk's interior against these digests, so the whole
1000-generation run is compared with the CPU path, not only with itself.  The oracle
takes about ten minutes on eight CPU cores:

    python tests/golden/make_gol_golden.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

NX = NY = 8192
SEED, P = 1002, 0.25
OUT = os.path.join(HERE, "gol_8192.json")


def main():
    a = np.zeros((NY + 2, NX + 2), dtype=np.uint8)
    oracle.fill_bernoulli(a, 1, SEED, P)
    b = np.zeros_like(a)
    records = []
    t0 = time.time()
    for chunk in range(1, 11):
        cur = oracle.gol_u8(a, b, NX, NY, 100)  # 100 is even: generation 100*chunk is in a
        assert cur is a
        interior = np.ascontiguousarray(a[1:-1, 1:-1])
        records.append({"generation": 100 * chunk, "sha256": hashlib.sha256(interior.tobytes()).hexdigest(),
                        "live": int(interior.sum(dtype=np.int64))})
        print(f"generation {100 * chunk}: live {records[-1]['live']} ({time.time() - t0:.0f} s)", flush=True)
    with open(OUT, "w") as f:
        json.dump({"nx": NX, "ny": NY, "seed": SEED, "p": P, "rule": "B3/S23", "periodic": True,
                   "records": records}, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()

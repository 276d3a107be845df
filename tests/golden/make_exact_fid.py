"""Write tests/golden/exact_fid.json: exact FIDs of the two full-size single-FID workloads.

* ring12: N=12 periodic ring, unit couplings, t = k * 0.01 for k = 0..200.
* j1j2_n16: N=16 J1-J2 ring with J2_i = 0.1 + 0.3 u_i (u_i from the seeded input stream 7),
  t = k * 0.005 for k = 0..1000.

Both by total-Sz sector diagonalisation (tests/exact_fid.py); values stored as hex bit patterns.
The N=16 case diagonalises blocks up to 12870 x 12870 and takes a few minutes on 8 cores.

    python tests/golden/make_exact_fid.py
"""
from __future__ import annotations

import json
import os
import struct
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import exact_fid as X  # noqa: E402
from ddevol_amd import hamiltonians as H  # noqa: E402

CASES = {
    "ring12": dict(n=12, terms=lambda: H.ring(12), dt=0.01, steps=200),
    "j1j2_n16": dict(n=16, terms=lambda: H.j1j2_n16_terms(16), dt=0.005, steps=1000),
}


def hexd(x):
    return struct.pack(">d", float(x)).hex()


def main():
    out = {"source": "tests/golden/make_exact_fid.py (total-Sz sector diagonalisation)"}
    for name, c in CASES.items():
        t0 = time.time()
        times = [k * c["dt"] for k in range(c["steps"] + 1)]
        fid = X.sector_fid(c["n"], c["terms"](), times,
                           progress=lambda m: print(f"{name}: {m} ({time.time() - t0:.0f} s)",
                                                    flush=True))
        out[name] = {"n": c["n"], "dt": hexd(c["dt"]), "steps": c["steps"],
                     "fid": [hexd(v) for v in fid]}
        print(f"{name}: done in {time.time() - t0:.0f} s, F(end) = {fid[-1]:.15f}", flush=True)
    with open(os.path.join(HERE, "exact_fid.json"), "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()

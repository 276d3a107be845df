#!/usr/bin/env python3
"""Write tests/golden/golden.json from the reference CPU engine.

The reference (engine/src/tpgp_rev.c and friends) is compiled from its own sources by
``make -C oracle ref`` into oracle/_ref/libtpgp_ref.so; this script drives it through ctypes on
small configurations and records, as bit patterns, what the GPU engine and the oracle must
reproduce:

* j1j2_n6: the setup of engine/tests/test_cuda_trajectory_rev.c:27-89 (N=6, J2=.3, dt 1/64,
  3 steps, thresholds {0, 1e-5} x validated {0, 1}): final map, per-step ledgers and readouts;
* ring12: the plain config (N=12 ring, dt .01, threshold 1e-8), first 20 steps;
* j1j2_n16: the metric's config (N=16, J2 from InputPCG(7), dt .005, threshold 1e-6),
  first 60 steps;
* cluster27: one powder orientation of the 3x3x3 cluster (dt .01, threshold 1e-5), 3 steps;
* ring48 / range64: two-word keys (N=48 ring with NN/NNN couplings, N=64 range-4 ring), 1 step;
* trig: libm cos/sin bits of every rotation angle used above (the GPU box must agree).

Run in the build container (where the reference sources exist):
    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ddevol_amd import hamiltonians as H  # noqa: E402
from oracle.bindings import CpuEngine, terms_array  # noqa: E402


def hexf(x: float) -> str:
    return struct.pack(">d", x).hex()


def map_digest(keys, values) -> str:
    h = hashlib.sha256()
    h.update(keys.astype("<u8").tobytes())
    h.update(values.astype("<f8").tobytes())
    return h.hexdigest()


def ledger_hex(led):
    return [hexf(led.dropped_l1), hexf(led.dropped_l2), hexf(led.dropped_l2_squared),
            hexf(led.arithmetic_l2), led.rotations, led.removal_events, led.peak_terms,
            led.arithmetic_enclosed]


def trajectory(eng, n, terms, dt, steps, threshold, validated=0, keep_map=False, digests=()):
    m = eng.new_map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    o = eng.options(threshold=threshold, validated=validated, max_terms=10 ** 9)
    led = eng.ledger(validated)
    ta = terms_array(terms)
    rec = {"n": n, "dt": hexf(dt), "steps": steps, "threshold": hexf(threshold),
           "validated": validated, "fid": [], "radius": [], "active": [], "ledger": [],
           "digest": {}}
    for k in range(steps + 1):
        st, mid, rad = m.readout(1, n)
        assert st == 0
        rec["fid"].append(hexf(mid))
        rec["radius"].append(hexf(rad))
        rec["active"].append(len(m))
        if k in digests:
            keys, vals = m.arrays()
            rec["digest"][str(k)] = map_digest(keys, vals)
        if k == steps:
            break
        st = m.strang_step(ta, len(terms), dt, o, led)
        assert st == 0, st
        rec["ledger"].append(ledger_hex(led))
    keys, vals = m.arrays()
    rec["final_digest"] = map_digest(keys, vals)
    if keep_map:
        rec["final_keys"] = [[int(w) for w in k] for k in keys]
        rec["final_values"] = [hexf(v) for v in vals]
    return rec


def trig_table(terms, dt):
    half = dt * 0.5
    out = []
    for (_, _, _, _, h) in terms:
        x = 2 * h * half
        out.append([hexf(h), hexf(math.cos(x)), hexf(math.sin(x))])
    return out


def main():
    eng = CpuEngine("ref")
    assert eng.has_mpfr(), "golden vectors are recorded with the MPFR build"
    out = {"source": "reference CPU engine (engine/src/tpgp_rev.c) compiled by make -C oracle ref",
           "mpfr": eng.has_mpfr()}
    j = []
    for thr in (0.0, 1e-5):
        for val in (0, 1):
            j.append(trajectory(eng, 6, H.j1j2(6, 0.3), 1.0 / 64, 3, thr, val, keep_map=True))
    out["j1j2_n6"] = j
    print("j1j2_n6 done", flush=True)
    out["ring12"] = trajectory(eng, 12, H.ring(12), 0.01, 20, 1e-8, digests=(5, 10, 20))
    print("ring12 done", flush=True)
    out["j1j2_n16"] = trajectory(eng, 16, H.j1j2_n16_terms(16), 0.005, 60, 1e-6,
                                 digests=(20, 40, 60))
    print("j1j2_n16 done", flush=True)
    orient = H.powder_orientations(1)[0]
    out["cluster27"] = trajectory(eng, 27, H.powder_terms(orient), 0.01, 3, 1e-5,
                                  digests=(1, 2, 3))
    print("cluster27 done", flush=True)
    g = H.InputPCG(7)
    nn48 = [0.5 + g.uniform() for _ in range(48)]
    out["ring48"] = trajectory(eng, 48, H.serving_ring(48, nn48), 0.01, 1, 1e-5, digests=(1,))
    print("ring48 done", flush=True)
    out["range64"] = trajectory(eng, 64, H.range_ring(64, 4), 0.01, 1, 1e-5, digests=(1,))
    print("range64 done", flush=True)
    out["trig"] = {"ring12": trig_table(H.ring(12), 0.01),
                   "j1j2_n16": trig_table(H.j1j2_n16_terms(16), 0.005)}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()

"""The GPU engine against the reference's own recorded trajectories (tests/golden/golden.json).

The fixtures were written by tests/golden/make_golden.py from the reference CPU engine
(engine/src/tpgp_rev.c compiled from its sources): FID midpoints and radii, sampled lengths,
ledgers after every step and SHA-256 digests of the final maps, all as bit patterns.  The GPU
path (resident evolution, exact ledger) must reproduce every one of them.  Trigonometry is the
host libm's; test_oracle.py checks that this host reproduces the recorded cos/sin bits.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from tests.golden_util import digest
from tests.test_oracle import config_terms, fromhex, ledger_from_hex

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def gpu_traj(ctx, n, terms, dt, steps, thr, validated=0):
    m = E.Map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    o = E.options(threshold=thr, validated=validated, max_terms=10 ** 9, context=ctx)
    led = E.ledger(validated)
    ev = E.Evolution(m, terms, dt, o, exact_ledger=True, context=ctx)
    fids, rads, act, leds = [], [], [], []
    for k in range(steps + 1):
        st, fid, rad, a, acc = ev.run(1 if k < steps else 0, led)
        assert st == 0, (k, st, E.N.engine().tpgp_hip_last_error().decode())
        fids.append(fid[0])
        rads.append(rad[0])
        act.append(int(a[0]))
        if k < steps:
            leds.append(led.bits())
    assert ev.close(m) == 0
    return m, fids, rads, act, leds


def check(rec, m, fids, rads, act, leds):
    assert [struct.pack(">d", x).hex() for x in fids] == rec["fid"]
    assert [struct.pack(">d", x).hex() for x in rads] == rec["radius"]
    assert act == rec["active"]
    assert leds == [ledger_from_hex(e) for e in rec["ledger"]]
    keys, vals = m.arrays()
    assert digest(keys, vals) == rec["final_digest"]


def test_golden_j1j2_n6_gpu(hip_ctx, golden):
    for rec in golden["j1j2_n6"]:
        m, fids, rads, act, leds = gpu_traj(hip_ctx, 6, H.j1j2(6, 0.3), 1.0 / 64, 3,
                                            fromhex(rec["threshold"]), rec["validated"])
        check(rec, m, fids, rads, act, leds)
        keys, vals = m.arrays()
        assert [[int(w) for w in k] for k in keys] == rec["final_keys"]
        assert [struct.pack(">d", v).hex() for v in vals] == rec["final_values"]


@pytest.mark.parametrize("name", ["ring12", "j1j2_n16", "cluster27", "ring48", "range64"])
def test_golden_configs_gpu(hip_ctx, golden, name):
    rec = golden[name]
    print(f"golden {name}: N={rec['n']}, {rec['steps']} steps", flush=True)
    m, fids, rads, act, leds = gpu_traj(hip_ctx, rec["n"], config_terms(name), fromhex(rec["dt"]),
                                        rec["steps"], fromhex(rec["threshold"]))
    check(rec, m, fids, rads, act, leds)


@pytest.mark.parametrize("limit", ["0", "100000000"])
@pytest.mark.parametrize("name", ["ring12", "j1j2_n16", "cluster27", "range64"])
def test_golden_both_paths(hip_ctx, golden, name, limit, monkeypatch):
    """The same golden trajectories with every map forced onto the multi-kernel path (limit 0)
    and onto the small-map persistent path (limit 1e8)."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", limit)
    rec = golden[name]
    m, fids, rads, act, leds = gpu_traj(hip_ctx, rec["n"], config_terms(name), fromhex(rec["dt"]),
                                        rec["steps"], fromhex(rec["threshold"]))
    check(rec, m, fids, rads, act, leds)

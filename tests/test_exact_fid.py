"""The exact FIDs of the full-size single-FID workloads (tests/golden/exact_fid.json), pinned.

* The sector method (tests/exact_fid.py) equals dense density-matrix FIDs on small rings and
  J1-J2 chains, odd and even N (the mirror-pair weighting differs between them).
* Its second moment equals the closed form M2 = (9 / 2N) sum_bonds d^2 (engine/tests/
  check_moments_rev.c:63-66) and the Pauli-algebra commutator sum.
* The committed N=12 series is recomputed here; the N=16 series (minutes to recompute) is pinned
  by its second moment and by the oracle's Strang trajectory over its first steps, which must
  lie within the oracle's own rigorous allowance of it.
"""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np
import pytest

import exact_fid as X
from ddevol_amd import hamiltonians as H
from oracle import bindings as OB
from test_oracle import dense_fid, run_traj

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "exact_fid.json")


def fromhex(h):
    return struct.unpack(">d", bytes.fromhex(h))[0]


@pytest.fixture(scope="module")
def exact():
    with open(FIXTURE) as f:
        data = json.load(f)
    out = {}
    for name in ("ring12", "j1j2_n16"):
        e = data[name]
        out[name] = dict(n=e["n"], dt=fromhex(e["dt"]), steps=e["steps"],
                         fid=np.array([fromhex(h) for h in e["fid"]]))
    return out


@pytest.mark.parametrize("n,model", [(8, "ring"), (7, "ring"), (6, "j1j2"), (7, "j1j2")])
def test_sector_method_equals_dense(n, model):
    terms = H.ring(n) if model == "ring" else H.j1j2(n, [0.1 + 0.05 * i for i in range(n)])
    t = np.array([0.0, 0.03, 0.4, 1.7, 3.2])
    a = X.sector_fid(n, terms, t)
    b = dense_fid(n, terms, t)
    assert a[0] == pytest.approx(1.0, abs=1e-14)
    assert np.max(np.abs(a - b)) < 1e-13


def test_sector_method_rejects_non_conserving():
    with pytest.raises(ValueError):
        X.sector_fid(4, [(0, 1, 1, 1, 0.5)], [0.1])  # XX alone flips two spins the same way
    with pytest.raises(ValueError):
        X.sector_fid(4, [(0, 1, 1, 3, 0.5)], [0.1])


def test_second_moment_closed_form():
    for n in (8, 12, 16):
        assert X.second_moment(n, H.ring(n)) == pytest.approx(4.5, rel=1e-14)
    assert X.second_moment(8, H.j1j2(8, 0.3)) == pytest.approx(981 / 200, rel=1e-14)
    j2 = H.j1j2_n16_couplings()
    m2 = 9 / (2 * 16) * sum(1.0 + x * x for x in j2)
    assert X.second_moment(16, H.j1j2_n16_terms()) == pytest.approx(m2, rel=1e-14)


def test_ring12_fixture_recomputes(exact):
    e = exact["ring12"]
    again = X.sector_fid(12, H.ring(12), [k * e["dt"] for k in range(e["steps"] + 1)])
    assert np.max(np.abs(again - e["fid"])) < 1e-13
    assert e["fid"][0] == pytest.approx(1.0, abs=1e-14)


@pytest.mark.parametrize("name", ["ring12", "j1j2_n16"])
def test_fixture_second_moment(exact, name):
    e = exact[name]
    terms = H.ring(12) if name == "ring12" else H.j1j2_n16_terms()
    t = e["dt"]
    m2 = X.second_moment(e["n"], terms)
    est = 2 * (1 - e["fid"][1]) / t ** 2  # = M2 - M4 t^2 / 12 + ...
    assert abs(est - m2) < 1e-3 * m2
    assert e["fid"][0] == pytest.approx(1.0, abs=1e-14)
    assert np.all(np.abs(e["fid"]) <= 1.0 + 1e-12)


@pytest.mark.parametrize("name,steps", [("ring12", 40), ("j1j2_n16", 12)])
def test_fixture_within_oracle_allowance(oracle, exact, name, steps):
    """The oracle's Strang FID (threshold as the workload) over the first steps stays within
    k * strang_bound + (dropped_l2 + arithmetic_l2) / sqrt(N) of the exact series, and that
    allowance is far from vacuous."""
    e = exact[name]
    n = e["n"]
    terms = H.ring(12) if name == "ring12" else H.j1j2_n16_terms()
    thr = 1e-8 if name == "ring12" else 1e-6
    m = oracle.new_map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    o = oracle.options(threshold=thr, max_terms=10 ** 9)
    led = oracle.ledger(0)
    ta = OB.terms_array(terms)
    per_step = oracle.strang_bound(ta, len(terms), e["dt"])
    worst = 0.0
    for k in range(steps + 1):
        st, mid, _ = m.readout(1, n)
        allow = k * per_step * (1 + 1e-12) + (led.dropped_l2 + led.arithmetic_l2) / math.sqrt(n)
        err = abs(mid - e["fid"][k])
        assert err <= allow + 1e-13, (k, mid, e["fid"][k], allow)
        worst = max(worst, err)
        if k < steps:
            assert m.strang_step(ta, len(terms), e["dt"], o, led) == 0
    assert allow < 1e-3

"""Full-size single-FID workloads on the device against their exact FIDs.

The north-star accuracy statement: device FIDs lie within each workload's truncation allowance
of the exact (total-Sz sector diagonalised) FID, where the allowance is the run driver's
observable bound k * strang_bound + (dropped_l2 + arithmetic_l2) / sqrt(N)
(engine/src/cli_rev.c:584-602, engine/src/tpgp_rev.c:887-961).  The ledger is read after every
step, so each step is held to its own allowance.

* N=12 periodic ring, dt 0.01, threshold 1e-8: all 200 steps of the workload (allowance 5e-3 at
  the end).
* N=16 J1-J2 (seeded J2_i), dt 0.005, threshold 1e-6: the first 500 of its 1000 steps (up to
  ~6e6 terms); the timed window of the benchmark starts at step 601.  The rigorous allowance
  passes 1 (and so says nothing) near step 450: it is held below 0.5 at step 300.

Beside the rigorous statement, an empirical one: the error stays below 1e-3 over these steps
(measured: 9e-6 on the ring, 4e-5 on J1-J2), so a regression that keeps FIDs inside a loose
allowance but moves them is still caught.
"""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "exact_fid.json")


def fromhex(h):
    return struct.unpack(">d", bytes.fromhex(h))[0]


@pytest.mark.parametrize("name,thr,steps", [("ring12", 1e-8, 200), ("j1j2_n16", 1e-6, 500)])
def test_device_fid_within_allowance_of_exact(hip_ctx, name, thr, steps):
    with open(FIXTURE) as f:
        e = json.load(f)[name]
    n, dt = e["n"], fromhex(e["dt"])
    exact = np.array([fromhex(h) for h in e["fid"]])
    terms = H.ring(n) if name == "ring12" else H.j1j2_n16_terms(n)
    print(f"{name}: N={n} dt={dt} threshold={thr} steps={steps}", flush=True)
    mp = E.Map(n)
    for s in range(n):
        assert mp.add_single(s, 1, 1.0) == 0
    o = E.options(threshold=thr, max_terms=10 ** 9, context=hip_ctx)
    led = E.ledger(0)
    per_step = E.strang_bound(terms, dt)
    ev = E.Evolution(mp, terms, dt, o, context=hip_ctx)
    pf, worst, worst_ratio, active = 0.0, 0.0, 0.0, 0
    for k in range(steps + 1):
        if k == 0:
            st, fid, _, act, _ = ev.run(0, led)
            got = fid[0]
        else:
            st, fid, _, act, acc = ev.run(1, led)
            assert st == 0 and acc == 1, (k, st)
            pf = E.N.engine().tpgp_bound_add_up(pf, per_step)
            got = fid[1]
        allow = E.fid_allowance(pf, led, n)
        err = abs(got - exact[k])
        assert err <= allow + 1e-12, (k, got, exact[k], allow)
        worst = max(worst, err)
        if allow > 1e-9:
            worst_ratio = max(worst_ratio, err / allow)
        if k == 300:
            assert allow < 0.5, allow  # a real statement at this size, not a vacuous one
        active = int(act[-1])
        if k % 100 == 0:
            print(f"  step {k}: F {got:.12f} exact {exact[k]:.12f} err {err:.2e} "
                  f"allowance {allow:.2e} terms {active}", flush=True)
    ev.close()
    print(f"{name}: worst |F - exact| {worst:.3e}, worst error/allowance {worst_ratio:.3f}, "
          f"final allowance {allow:.3e}", flush=True)
    assert worst < 1e-3
    if name == "ring12":
        assert allow < 0.05

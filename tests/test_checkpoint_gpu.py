"""Checkpoint and resume of a resident session, and the device l2 norm.

* A session run N steps, snapshotted into a TPGPRVCP record, reloaded and resumed for M steps
  gives bit for bit the FIDs, radii, lengths, final map and exact ledger of N + M uninterrupted
  steps (the record carries everything the loop needs).
* tpgp_evolution_l2 agrees with the host tpgp_map_l2 (long double, map order, as
  tpgp_rev.c:266-271) within 2 units in the last place on maps of 1e5+ terms, where every
  block of the reduction runs several grid strides.
"""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H

pytestmark = pytest.mark.gpu


def sum_x(n):
    m = E.Map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    return m


def bits(a):
    return np.asarray(a, np.float64).view(np.uint64)


def ulps(x, y):
    ix = struct.unpack("<q", struct.pack("<d", x))[0]
    iy = struct.unpack("<q", struct.pack("<d", y))[0]
    return abs(ix - iy)


@pytest.mark.parametrize("n,thr,first,more", [(12, 1e-6, 5, 7), (40, 1e-4, 2, 3)])
def test_snapshot_checkpoint_resume(hip_ctx, tmp_path, n, thr, first, more):
    terms = H.j1j2(n, 0.3) if n <= 32 else H.serving_ring(n, [1.0] * n)
    dt = 0.01
    o = E.options(threshold=thr, max_terms=10 ** 8, context=hip_ctx)
    # uninterrupted
    led_a = E.ledger(0)
    ev = E.Evolution(sum_x(n), terms, dt, o, exact_ledger=True, context=hip_ctx)
    st, fid_a, rad_a, act_a, acc = ev.run(first + more, led_a)
    assert st == 0 and acc == first + more
    end_a = E.Map(n)
    assert ev.close(end_a) == 0
    # interrupted: run, snapshot, save, load, resume
    led_b = E.ledger(0)
    ev = E.Evolution(sum_x(n), terms, dt, o, exact_ledger=True, context=hip_ctx)
    st, fid_b, rad_b, act_b, acc = ev.run(first, led_b)
    assert st == 0 and acc == first
    snap = ev.snapshot()
    ev.close()
    path = tmp_path / "s.ck"
    E.Checkpoint.of(snap, terms, o, led_b, dt=dt, step=first, total_steps=first + more,
                    model="edges").save(path)
    rec = E.Checkpoint.load(path)
    assert rec.step == first and rec.same_configuration(rec)
    led_c = rec.c.ledger
    o2 = rec.c.options
    o2.hip_context = hip_ctx.ptr
    ev = E.Evolution(rec.map(), rec.terms(), rec.dt, o2, exact_ledger=True, context=hip_ctx)
    st, fid_c, rad_c, act_c, acc = ev.run(more, led_c)
    assert st == 0 and acc == more
    end_c = E.Map(n)
    assert ev.close(end_c) == 0
    fid = np.concatenate([fid_b[:first], fid_c])
    rad = np.concatenate([rad_b[:first], rad_c])
    act = np.concatenate([act_b[:first], act_c])
    assert np.array_equal(bits(fid), bits(fid_a)) and np.array_equal(bits(rad), bits(rad_a))
    assert np.array_equal(act, act_a)
    ka, va = end_a.arrays()
    kc, vc = end_c.arrays()
    assert np.array_equal(ka, kc) and np.array_equal(bits(va), bits(vc))
    assert led_a.bits() == led_c.bits()


def exact_l2(values):
    """sqrt(sum of x^2) of binary64 values, the sum taken exactly in integers and the square
    root rounded once (to 110 bits, then to binary64)."""
    parts = []
    for x in values:
        if x == 0:
            continue
        m, e = math.frexp(abs(float(x)))
        parts.append((int(m * (1 << 53)), 2 * (e - 53)))
    if not parts:
        return 0.0
    lo = min(e for _, e in parts)
    total = sum(m * m << (e - lo) for m, e in parts)
    shift = max(0, 220 - total.bit_length())
    shift += (lo - shift) % 2  # even exponent under the root
    r = math.isqrt(total << shift)
    return math.ldexp(float(r), (lo - shift) // 2)


def random_map(n, count, seed):
    """Random canonical map (sorted unique keys) with values over 16 decades of magnitude."""
    rng = np.random.default_rng(seed)
    w = (n + 31) // 32
    digits = rng.integers(0, 4, size=(count, n), dtype=np.uint64)
    keys = np.zeros((count, w), np.uint64)
    for s in range(n):
        keys[:, s // 32] |= digits[:, s] << np.uint64(2 * (s % 32))
    keys = np.unique(keys, axis=0)
    keys = keys[np.lexsort(tuple(keys[:, k] for k in range(w)))]
    values = rng.standard_normal(len(keys)) * 10.0 ** rng.uniform(-8, 8, len(keys))
    values[values == 0] = 1.0
    return E.Map.from_sorted(n, keys, values)


def test_device_l2_matches_host(hip_ctx):
    """Along a J1-J2 N=16 run, and on random maps of 1e6 terms (W = 1 and 2) whose reduction
    runs several grid strides per block."""
    n, dt = 16, 0.005
    o = E.options(threshold=1e-6, max_terms=10 ** 8, context=hip_ctx)
    led = E.ledger(0)
    ev = E.Evolution(sum_x(n), H.j1j2_n16_terms(n), dt, o, context=hip_ctx)
    for chunk in (1, 20, 40):
        assert ev.run(chunk, led)[0] == 0
        dev = ev.l2()
        host = ev.snapshot()
        want = E.N.engine().tpgp_map_l2(E.C.byref(host.m))
        _, v = host.arrays()
        exact = exact_l2(v)
        print(f"len {len(host)}: device {dev!r} host {want!r} exact {exact!r}", flush=True)
        assert ulps(dev, want) <= 2 and ulps(dev, exact) <= 1
    ev.close()
    for nn, seed in ((20, 1), (40, 2)):
        mp = random_map(nn, 1_000_000, seed)
        ev = E.Evolution(mp, H.ring(nn), dt, o, context=hip_ctx)
        dev = ev.l2()
        want = E.N.engine().tpgp_map_l2(E.C.byref(mp.m))
        _, v = mp.arrays()
        exact = exact_l2(v)
        print(f"random N={nn} len {len(mp)}: device {dev!r} host {want!r} exact {exact!r}",
              flush=True)
        assert len(mp) > 4 * 512 * 256
        assert ulps(dev, want) <= 2 and ulps(dev, exact) <= 1
        ev.close()

"""Appending rotations against copying every class, bit for bit.

An appending rotation leaves the classes it does not change where they are in the slot and
writes the classes it rewrites after the slot's used extent.  Every case here runs twice, once
with appending (the default) and once with TPGP_HIP_APPEND=0 (every class copied to another
buffer), with the device layout checked after every pass (TPGP_HIP_VALIDATE=1: order, class
regions, extents inside the slot, no two classes overlapping), and the two runs must agree bit
for bit: records, readouts, sums of squares, downloaded maps, FIDs and ledgers.

Cases: an appending rotation that outgrows its slot (reported as an overflow, rolled back, run
again copying); a committed state whose classes are out of key order in the slot (sum of
squares, re-layout into a larger slot, download); readouts and partitions of such states in
the middle of a chunk; whole resident trajectories on the multi-kernel path.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import native as N
from tests.test_oracle import config_terms, fromhex

pytestmark = pytest.mark.gpu
X, Y, Z = 1, 2, 3


class Batch:
    """A one-map resident batch (W = 1) driven through the C ABI."""

    def __init__(self, ctx, n, cap, append, monkeypatch):
        monkeypatch.setenv("TPGP_HIP_APPEND", "1" if append else "0")
        monkeypatch.setenv("TPGP_HIP_VALIDATE", "1")
        monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", "0")  # the multi-kernel passes
        self.L = N.hip()
        self.err = C.create_string_buffer(512)
        self.b = self.L.tpgp_hip_batch_create(ctx.ptr, 1, 1, (C.c_size_t * 1)(n),
                                              (C.c_size_t * 1)(cap), self.err, 512)
        assert self.b, self.err.value.decode()
        self.seq = 0

    def ok(self, st):
        assert st == 0, (st, self.err.value.decode())

    def close(self):
        self.L.tpgp_hip_batch_destroy(self.b)

    def upload(self, keys, vals):
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self.ok(self.L.tpgp_hip_batch_upload(self.b, 0, k.ctypes.data_as(C.POINTER(C.c_uint64)),
                                             v.ctypes.data_as(C.POINTER(C.c_double)), len(v),
                                             self.err, 512))

    def rotate(self, rots):
        """rots: (i, j, qi, qj, theta, threshold) in order, one call."""
        req = (N.HipRotation * len(rots))()
        for r, (i, j, qi, qj, th, thr) in enumerate(rots):
            req[r] = N.HipRotation(1, qi, qj, 0, 0, i, j, math.cos(th), math.sin(th), 0.0, 0.0,
                                   thr)
        seq = (C.c_uint32 * len(rots))(*range(self.seq, self.seq + len(rots)))
        self.seq += len(rots)
        self.ok(self.L.tpgp_hip_batch_rotate_many(self.b, req, seq, len(rots), self.err, 512))

    def readout(self, pauli=X):
        self.ok(self.L.tpgp_hip_batch_readout(self.b, pauli, None, self.err, 512))

    def sync(self):
        rec = (N.HipRecord * 64)()
        rd = (C.c_double * (64 * 16))()
        nrec, nrd, nlog = C.c_size_t(), C.c_size_t(), C.c_size_t()
        lens = (C.c_uint64 * 1)()
        self.ok(self.L.tpgp_hip_batch_sync(self.b, rec, 64, C.byref(nrec), rd, 16, C.byref(nrd),
                                           lens, None, 0, C.byref(nlog), self.err, 512))
        recs = [(r.len_in, r.len_out, r.dropped, r.l1_up, r.sq_up, r.executed, r.nonfinite,
                 r.overflow, r.rotation) for r in rec[:nrec.value]]
        return recs, [rd[k] for k in range(64 * nrd.value)], int(lens[0])

    def commit(self):
        self.ok(self.L.tpgp_hip_batch_commit(self.b, None, self.err, 512))

    def rollback(self):
        self.ok(self.L.tpgp_hip_batch_rollback(self.b, self.err, 512))

    def sum_squares(self):
        out = (C.c_double * 2)()
        self.ok(self.L.tpgp_hip_batch_sum_squares(self.b, 0, out, self.err, 512))
        return out[0], out[1]

    def reserve(self, cap):
        self.ok(self.L.tpgp_hip_batch_reserve(self.b, 0, cap, self.err, 512))

    def download(self):
        n = C.c_size_t()
        self.ok(self.L.tpgp_hip_batch_download(self.b, 0, None, None, 0, C.byref(n), self.err, 512))
        k = np.zeros(n.value, dtype=np.uint64)
        v = np.zeros(n.value, dtype=np.float64)
        self.ok(self.L.tpgp_hip_batch_download(self.b, 0, k.ctypes.data_as(C.POINTER(C.c_uint64)),
                                               v.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                               C.byref(n), self.err, 512))
        return k, v.view(np.uint64)


def random_map(rng, n, count, low_digits=None):
    """`count` distinct sorted keys of n sites (reference layout) and values in +-[0.5, 1.5];
    low_digits fixes the symbols of sites 0 and 1 (as the two low digits) when given."""
    if low_digits is None:
        keys = rng.choice(4 ** n, count, replace=False)
    else:
        keys = (rng.choice(4 ** (n - 2), count, replace=False) << 4) | low_digits
    keys = np.sort(keys.astype(np.uint64))
    vals = rng.uniform(0.5, 1.5, count) * rng.choice([-1.0, 1.0], count)
    return keys, vals


def first_diff(a, c, path=()):
    """The path and values of the first place where two nested results differ (None: equal)."""
    if isinstance(a, (list, tuple)) and isinstance(c, (list, tuple)):
        if len(a) != len(c):
            return path + ("len",), len(a), len(c)
        for k, (x, y) in enumerate(zip(a, c)):
            d = first_diff(x, y, path + (k,))
            if d:
                return d
        return None
    return None if a == c else (path, a, c)


def same(a, c):
    d = first_diff(a, c)
    assert d is None, f"appending vs copying differ at {d[0]}: {d[1]!r} != {d[2]!r}"


def both(hip_ctx, monkeypatch, n, cap, body):
    """body(batch) with appending on and off; the two result lists must be equal."""
    out = []
    for append in (True, False):
        b = Batch(hip_ctx, n, cap, append, monkeypatch)
        try:
            out.append(body(b))
        finally:
            b.close()
    return out


def test_append_overflow_runs_again_copying(hip_ctx, monkeypatch):
    """Every key in class (Z, I) of sites (0, 1): an XX rotation sends all of them to two
    classes, 2 x 1e5 terms written after an extent of 1e5 in a slot of 2.1e5.  The appending run
    plans an append (bound 0.75 of the map), overflows, is rolled back and runs again copying;
    the accepted result equals the copying run's."""
    rng = np.random.default_rng(5)
    keys, vals = random_map(rng, 12, 100_000, low_digits=Z)
    rot = [(0, 1, X, X, 0.4, 1e-12)]

    def body(b):
        b.upload(keys, vals)
        b.rotate(rot)
        recs, _, _ = b.sync()
        first = recs[0][7]  # overflow flag of the first try
        if first:
            b.rollback()
            b.seq = 0  # the same rotation again, under its own sequence number
            b.rotate(rot)
            recs, _, _ = b.sync()
        b.commit()
        return first, recs, b.sum_squares(), b.download()

    (first_a, *res_a), (first_c, *res_c) = both(hip_ctx, monkeypatch, 12, 210_000, body)
    assert first_a == 1 and first_c == 0
    assert res_a[0][0][7] == 0 and res_a[0][0][1] == 200_000
    assert res_a[0] == res_c[0]
    assert res_a[1] == res_c[1]
    for x, y in zip(res_a[2], res_c[2]):
        assert np.array_equal(x, y)


def test_gapped_committed_state(hip_ctx, monkeypatch):
    """A committed state left by an appending rotation (anticommuting classes after the extent,
    the others in place): its sum of squares, a re-layout into a larger slot, more rotations
    from the re-laid state, and the download all match the copying run."""
    rng = np.random.default_rng(7)
    keys, vals = random_map(rng, 12, 60_000)

    def body(b):
        out = []
        b.upload(keys, vals)
        b.rotate([(0, 1, X, X, 0.3, 1e-9), (0, 1, Y, Y, 0.3, 1e-9)])
        out.append(b.sync()[0])
        b.commit()
        out.append(b.sum_squares())
        b.reserve(2_000_000)
        out.append(b.sum_squares())
        b.rotate([(0, 1, Z, Z, 0.3, 1e-9), (1, 2, X, X, 0.2, 1e-9)])
        out.append(b.sync()[0])
        b.commit()
        out.append(b.sum_squares())
        k, v = b.download()
        out.append((k.tolist(), v.tolist()))
        return out

    a, c = both(hip_ctx, monkeypatch, 12, 1_000_000, body)
    same(a, c)


def run_segment(b, ops):
    """ops (rotations and readouts) up to a sync, run again after a rollback while a rotation
    overflowed its slot (as a caller must: the library then plans more room); the accepted
    records and readouts."""
    seq0 = b.seq
    for attempt in range(4):
        b.seq = seq0
        for op in ops:
            op(b)
        recs, rd, ln = b.sync()
        if not any(r[7] for r in recs):
            b.commit()
            return recs, [struct_bits(x) for x in rd], ln
        b.rollback()
    raise AssertionError("a segment overflowed four times")


def test_readouts_and_partitions_of_gapped_states(hip_ctx, monkeypatch):
    """Readouts between appending rotations and partitions whose input has classes out of
    place, inside one segment and across segments; several thresholds (some rotations
    truncate).  The maps grow ~1.5x per rotation, so some appending segments outgrow the slot
    and run again."""
    rng = np.random.default_rng(11)
    keys, vals = random_map(rng, 10, 120_000)  # at most 4^10 keys: the maps stay in the slot
    bonds = [(0, 1), (1, 2), (2, 3), (0, 3), (5, 9), (9, 8), (4, 7)]

    def body(b):
        out = []
        b.upload(keys, vals)
        ops = []
        for r, (i, j) in enumerate(bonds):
            thr = 5e-2 if r % 3 == 2 else 1e-3
            rots = [(i, j, q, q, 0.05 + 0.07 * r + 0.01 * q, thr) for q in (X, Y, Z)]
            ops += [lambda b, x=rots[:2]: b.rotate(x), lambda b: b.readout(X),
                    lambda b, x=rots[2:]: b.rotate(x), lambda b: b.readout(Z)]
            if r % 2:
                out.append(run_segment(b, ops))
                out.append(b.sum_squares())
                ops = []
        out.append(run_segment(b, ops))
        k, v = b.download()
        out.append((k.tolist(), v.tolist()))
        return out

    a, c = both(hip_ctx, monkeypatch, 10, 2_000_000, body)
    same(a, c)


@pytest.mark.parametrize("cut,limit,lds", [(1.4, 5000, True), (1.35, 8000, False)])
def test_gapped_state_onto_the_small_map_path(hip_ctx, monkeypatch, cut, limit, lds):
    """A committed state left by an appending rotation is the input of the small-map kernels
    (the LDS kernel loads it class by class when it fits the LDS capacity of 4096 one-word
    terms, the global one runs from it when it does not): with a small-path bound of `limit`
    terms, a call that truncates most of the map and then appends, a second call that copies and
    appends (its length estimate falls to the bound), then calls on the small path."""
    rng = np.random.default_rng(13)
    keys, vals = random_map(rng, 12, 20_000)

    def body(b):
        out = []
        b.upload(keys, vals)
        calls = [[(0, 1, Z, Z, 0.3, cut), (0, 1, X, X, 0.3, 1e-12)],
                 [(0, 1, Y, Y, 0.2, 1e-12), (0, 1, Z, Z, 0.2, 1e-12)],
                 [(1, 2, X, X, 0.25, 1e-12), (1, 2, Y, Y, 0.25, 1e-12), (3, 4, Z, Z, 0.1, 1e-12)],
                 [(2, 5, X, X, 0.25, 1e-12), (0, 2, Y, Y, 0.15, 1e-12)]]
        for rots in calls:
            out.append(run_segment(b, [lambda b, x=rots: b.rotate(x), lambda b: b.readout(X)]))
            out.append(b.sum_squares())
        k, v = b.download()
        out.append((k.tolist(), v.tolist()))
        return out

    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", str(limit))
    out = []
    for append in (True, False):
        monkeypatch.setenv("TPGP_HIP_APPEND", "1" if append else "0")
        monkeypatch.setenv("TPGP_HIP_VALIDATE", "1")
        L = N.hip()
        err = C.create_string_buffer(512)
        b = Batch.__new__(Batch)
        b.L, b.err, b.seq = L, err, 0
        b.b = L.tpgp_hip_batch_create(hip_ctx.ptr, 1, 1, (C.c_size_t * 1)(12),
                                      (C.c_size_t * 1)(400_000), err, 512)
        assert b.b, err.value.decode()
        try:
            out.append(body(b))
        finally:
            b.close()
    a, c = out
    lens = [[r[0] for r in seg[0]] for seg in a[0:8:2]]
    print(f"cut {cut}: input lengths per call {lens}", flush=True)
    # the third call starts on the small path from the second call's appended state
    assert lens[2][0] <= limit and (lens[2][0] <= 4096) == lds
    same(a, c)


def struct_bits(x):
    return int(np.float64(x).view(np.uint64))


# (spins, dt, threshold, steps): maps of 1e4-2e5 terms by the last step
TRAJ = {"j1j2_n16": (16, 0.02, 1e-6, 40), "ring12": (12, 0.02, 1e-7, 40),
        "range64": (64, 0.02, 1e-5, 8)}


def trajectory(hip_ctx, monkeypatch, name, append):
    monkeypatch.setenv("TPGP_HIP_APPEND", "1" if append else "0")
    monkeypatch.setenv("TPGP_HIP_VALIDATE", "1")
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", "0")
    n, dt, thr, steps = TRAJ[name]
    m = E.Map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    o = E.options(threshold=thr, max_terms=10 ** 9, context=hip_ctx)
    led = E.ledger(0)
    ev = E.Evolution(m, config_terms(name), dt, o, exact_ledger=True, context=hip_ctx)
    fids, leds = [], []
    for k in range(steps + 1):
        st, fid, rad, a, acc = ev.run(1 if k < steps else 0, led)
        assert st == 0, (k, st, E.N.engine().tpgp_hip_last_error().decode())
        fids.append(struct_bits(fid[0]))
        if k < steps:
            leds.append(led.bits())
    assert ev.close(m) == 0
    keys, vals = m.arrays()
    return fids, leds, [[int(w) for w in k] for k in keys], [struct_bits(v) for v in vals]


@pytest.mark.parametrize("name", sorted(TRAJ))
def test_trajectories_append_equals_copy(hip_ctx, monkeypatch, name):
    """Resident trajectories (exact ledger) on the multi-kernel path: appending and copying
    give the same FIDs, ledgers after every step and final maps."""
    print(f"append trajectory {name}: {TRAJ[name]}", flush=True)
    a = trajectory(hip_ctx, monkeypatch, name, True)
    c = trajectory(hip_ctx, monkeypatch, name, False)
    same(a, c)

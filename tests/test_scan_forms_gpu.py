"""The rotation scan's three forms give the same results, bit for bit.

A rotation's scan turns the count pass's per-tile kept counts and ledger partials into output
offsets, class regions, the map's length and its record.  It has three forms: one launch of one
block per map (k_rot_scan_maps, for small maps), two levels (k_rot_scan1 per chunk of 256
tiles, then k_rot_scan2 per map) and direct (k_rot_scan1 writes every chunk's output offset
itself and the map-level end runs inside the write pass).  The library picks the form by size;
TPGP_HIP_SCAN forces one.  The two-level and direct forms fold the ledger partials in the same
order, so their records agree to the last bit; the one-launch form folds them in another order,
so its records are compared without the two ledger sums, and everything else in full.

Every case runs with the device layout checked after every pass (TPGP_HIP_VALIDATE=1) on the
multi-kernel path.  Cases: an appending rotation that outgrows its slot (in the direct form the
overflow is found inside the write pass, whose tiles then write only where they fit); readouts
and partitions between appending rotations; several maps of different sizes in one fleet; a
resident trajectory with the exact ledger.
"""
from __future__ import annotations

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from tests.test_append_gpu import (TRAJ, Batch, X, Y, Z, first_diff, random_map, run_segment,
                                   trajectory)

pytestmark = pytest.mark.gpu
FORMS = ("levels", "direct", "maps")


def without_ledger_sums(out):
    """Records (9-tuples: len_in, len_out, dropped, l1_up, sq_up, ...) with the two ledger sums
    taken out, anywhere in a nested result."""
    if isinstance(out, tuple) and len(out) == 9 and all(isinstance(x, (int, float)) for x in out):
        return out[:3] + out[5:]
    if isinstance(out, (list, tuple)):
        return type(out)(without_ledger_sums(x) for x in out)
    return out


def agree(res):
    """res: form -> result.  direct equals levels in full, maps equals both without the
    ledger sums."""
    d = first_diff(res["direct"], res["levels"])
    assert d is None, f"direct vs levels differ at {d[0]}: {d[1]!r} != {d[2]!r}"
    d = first_diff(without_ledger_sums(res["maps"]), without_ledger_sums(res["levels"]))
    assert d is None, f"maps vs levels differ at {d[0]}: {d[1]!r} != {d[2]!r}"


def each_form(hip_ctx, monkeypatch, n, cap, body):
    res = {}
    for form in FORMS:
        monkeypatch.setenv("TPGP_HIP_SCAN", form)
        b = Batch(hip_ctx, n, cap, True, monkeypatch)
        try:
            res[form] = body(b)
        finally:
            b.close()
    monkeypatch.delenv("TPGP_HIP_SCAN")
    return res


def test_overflow_found_in_the_write_pass(hip_ctx, monkeypatch):
    """Every key in class (Z, I) of sites (0, 1): an XX rotation sends all 1e5 of them to two
    classes, written after the extent of a slot of 2.1e5.  The appending rotation overflows in
    every form; rolled back, it runs again copying, and the accepted results agree."""
    rng = np.random.default_rng(5)
    keys, vals = random_map(rng, 12, 100_000, low_digits=Z)
    rot = [(0, 1, X, X, 0.4, 1e-12)]

    def body(b):
        b.upload(keys, vals)
        b.rotate(rot)
        first, _, _ = b.sync()
        b.rollback()
        b.seq = 0
        b.rotate(rot)
        recs, _, _ = b.sync()
        b.commit()
        k, v = b.download()
        return first, recs, b.sum_squares(), (k.tolist(), v.tolist())

    res = each_form(hip_ctx, monkeypatch, 12, 210_000, body)
    for form, (first, recs, _, _) in res.items():
        assert first[0][7] == 1 and recs[0][7] == 0 and recs[0][1] == 200_000, form
    agree(res)


def test_readouts_and_partitions(hip_ctx, monkeypatch):
    """Readouts between appending rotations and partitions of states with classes out of place,
    some rotations truncating; maps of 1.2e5 .. 1e6 terms (500 .. 4000 tiles), segments that
    outgrow the slot run again."""
    rng = np.random.default_rng(11)
    keys, vals = random_map(rng, 10, 120_000)
    bonds = [(0, 1), (1, 2), (2, 3), (0, 3), (5, 9), (9, 8)]

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
        k, v = b.download()
        out.append((k.tolist(), v.tolist()))
        return out

    agree(each_form(hip_ctx, monkeypatch, 10, 2_000_000, body))


def test_fleet_of_several_maps(hip_ctx, monkeypatch):
    """Four FIDs of different sizes in one fleet on the multi-kernel path: the write pass's
    first blocks end the scan of one map each.  FIDs, radii, lengths and final maps agree."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", "0")
    g = H.InputPCG(20260711)
    jobs = []
    for k, n in enumerate((12, 16, 10, 14)):
        nn = [0.5 + g.uniform() for _ in range(n)]
        jobs.append((n, H.serving_ring(n, nn), [1e-5, 1e-6][k % 2], 6 + k))
    res = {}
    for form in FORMS:
        monkeypatch.setenv("TPGP_HIP_SCAN", form)
        fl = E.Fleet(1, len(jobs), 1 << 12, context=hip_ctx)
        for s, (n, t, thr, steps) in enumerate(jobs):
            m = E.Map(n)
            for i in range(n):
                assert m.add_single(i, 1, 1.0) == 0
            assert fl.admit(s, m, t, 0.01, steps, E.options(threshold=thr, max_terms=10 ** 8),
                            exact_ledger=True) == 0
        for _ in range(100000):
            st, running = fl.step()
            assert st == 0, (form, st, E.N.engine().tpgp_hip_last_error().decode())
            if running == 0:
                break
        out = []
        for s, (n, *_rest) in enumerate(jobs):
            state, done, status, led = fl.slot(s)
            f, r, a = fl.fid(s, 64)
            fm = E.Map(n)
            assert fl.release(s, fm) == 0
            k, v = fm.arrays()
            out.append((status, np.asarray(f).view(np.uint64).tolist(),
                        np.asarray(r).view(np.uint64).tolist(), np.asarray(a).tolist(),
                        led.bits(), np.asarray(k).tolist(),
                        np.asarray(v, np.float64).view(np.uint64).tolist()))
        fl.close()
        res[form] = out
    monkeypatch.delenv("TPGP_HIP_SCAN")
    agree(res)


def test_trajectory(hip_ctx, monkeypatch):
    """The N=16 J1-J2 resident trajectory (exact ledger, 40 steps, up to ~1e5 terms): FIDs,
    ledgers after every step and the final map agree in every form."""
    res = {}
    for form in FORMS:
        monkeypatch.setenv("TPGP_HIP_SCAN", form)
        print(f"scan form {form}: trajectory j1j2_n16 {TRAJ['j1j2_n16']}", flush=True)
        res[form] = trajectory(hip_ctx, monkeypatch, "j1j2_n16", True)
    monkeypatch.delenv("TPGP_HIP_SCAN")
    agree(res)

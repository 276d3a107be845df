"""Step graphs: an evolution that records its steps as HIP graphs and replays them gives, bit for
bit, what enqueueing every step directly gives.

The resident loop (tpgp_evolution_run) records a step once it starts from a host state it has
started from before (same buffers, digit layout, floor) and replays the recording whenever that
state recurs.  Each case runs the same FID twice, TPGP_HIP_GRAPHS=0 and =1 (read when the
session's batch is created), and compares every FID sample, term count, the exact ledger and
the final map bitwise (the fast ledger's bound sums to 1e-12 relative, its counts exactly); the
recorded run must have replayed steps.  Runs are chunked, with snapshots
and norms of the map between the chunks (eager work and syncs between replays), and one case
starts from a slot too small for the FID, so the slots grow (re-layouts drop the graphs) while
graphs are in use.
"""
import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N

pytestmark = pytest.mark.gpu


def sum_x(n):
    m = E.Map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    return m


def run(ctx, monkeypatch, graphs, n, terms, dt, thr, chunks, exact, validated, cap):
    monkeypatch.setenv("TPGP_HIP_GRAPHS", "1" if graphs else "0")
    o = E.options(threshold=thr, max_terms=10 ** 9, validated=validated, context=ctx)
    led = E.ledger(validated)
    ev = E.Evolution(sum_x(n), terms, dt, o, exact_ledger=exact, initial_capacity=cap,
                     context=ctx)
    fids, acts, norms = [], [], []
    for k, steps in enumerate(chunks):
        st, fid, _, act, acc = ev.run(steps, led)
        assert st == 0 and acc == steps, N.engine().tpgp_hip_last_error().decode()
        fids.append(fid if k == 0 else fid[1:])
        acts.append(act if k == 0 else act[1:])
        snap = ev.snapshot()  # eager work between replays
        norms.append(ev.l2())
        assert len(snap) == int(act[acc])
    counts = ev.graph_counts()
    final = E.Map(n)
    assert ev.close(final) == 0
    return np.concatenate(fids), np.concatenate(acts), norms, led, final, counts


CASES = [
    # N, terms, dt, threshold, chunks, exact ledger, validated, initial capacity
    ("ring12_lds", 12, lambda: H.ring(12), 0.01, 1e-8, [20, 15, 15], False, 0, 0),
    ("ring12_exact", 12, lambda: H.ring(12), 0.01, 1e-8, [25, 25], True, 0, 0),
    ("j1j2_n16_multi", 16, lambda: H.j1j2_n16_terms(16), 0.005, 1e-6, [60, 40, 40], False, 0, 0),
    ("j1j2_n16_grow", 16, lambda: H.j1j2_n16_terms(16), 0.005, 1e-6, [50, 50], True, 0, 256),
    ("j1j2_n40_two_words", 40, lambda: H.j1j2(40, 0.3), 0.01, 1e-5, [12, 12], False, 0, 0),
    ("ring12_validated", 12, lambda: H.ring(12), 0.01, 1e-8, [15, 15], False, 1, 0),
]


@pytest.mark.parametrize("name,n,terms_fn,dt,thr,chunks,exact,validated,cap", CASES,
                         ids=[c[0] for c in CASES])
def test_recorded_steps_match_direct(hip_ctx, monkeypatch, name, n, terms_fn, dt, thr, chunks,
                                     exact, validated, cap):
    terms = terms_fn()
    print(f"{name}: direct", flush=True)
    f0, a0, n0, l0, m0, c0 = run(hip_ctx, monkeypatch, False, n, terms, dt, thr, chunks, exact,
                                 validated, cap)
    print(f"{name}: recorded, {int(a0[-1])} terms at the end", flush=True)
    f1, a1, n1, l1, m1, c1 = run(hip_ctx, monkeypatch, True, n, terms, dt, thr, chunks, exact,
                                 validated, cap)
    print(f"{name}: graphs recorded/replayed {c1}", flush=True)
    assert c0 == (0, 0)
    assert c1[1] > 0
    assert np.array_equal(f0.view(np.uint64), f1.view(np.uint64))
    assert np.array_equal(a0, a1)
    assert np.array_equal(np.array(n0).view(np.uint64), np.array(n1).view(np.uint64))
    if exact:
        assert l0.bits() == l1.bits()
    else:
        # the fast ledger's sums are upper bounds rounded in a tree whose shape follows the
        # device path a step took (the small-map kernel or the multi-kernel passes, one scan
        # kernel or two); the path is chosen from the lengths known when the step is planned,
        # and a recorded step keeps the choice of its recording.  Counts agree exactly.
        b0, b1 = l0.bits(), l1.bits()
        assert b0[4:] == b1[4:]
        for x, y in zip(b0[:4], b1[:4]):
            fx, fy = (np.array([x, y], np.uint64).view(np.float64))
            assert fx == fy or abs(fx - fy) <= 1e-12 * max(abs(fx), abs(fy)), (fx, fy)
    k0, v0 = m0.arrays()
    k1, v1 = m1.arrays()
    assert np.array_equal(k0, k1) and np.array_equal(v0.view(np.uint64), v1.view(np.uint64))


def fleet_run(ctx, monkeypatch, graphs):
    monkeypatch.setenv("TPGP_HIP_GRAPHS", "1" if graphs else "0")
    jobs = [(12, H.ring(12), 0.01, 1e-8, 40, True), (10, H.ring(10), 0.01, 1e-8, 15, False),
            (16, H.j1j2_n16_terms(16), 0.005, 1e-6, 60, False), (8, H.ring(8), 0.01, 1e-8, 30,
                                                                 True)]
    fl = E.Fleet(1, 4, 1024, context=ctx)
    for s, (n, terms, dt, thr, steps, exact) in enumerate(jobs):
        assert fl.admit(s, sum_x(n), terms, dt, steps, E.options(threshold=thr, max_terms=10 ** 8),
                        exact_ledger=exact) == 0
    rounds = 0
    while True:
        st, running = fl.step()
        assert st == 0, N.engine().tpgp_hip_last_error().decode()
        rounds += 1
        if rounds == 20:  # a new job in a freed slot changes the round's slots mid-run
            assert fl.slot(1)[0] == N.SLOT_DONE
            assert fl.release(1) == 0
            assert fl.admit(1, sum_x(9), H.ring(9), 0.01, 30,
                            E.options(threshold=1e-8, max_terms=10 ** 8)) == 0
        if running == 0:
            break
    out = []
    for s in range(4):
        state, done, status, led = fl.slot(s)
        f, _, act = fl.fid(s, 128)
        final = E.Map(jobs[s][0] if s != 1 else 9)
        assert fl.release(s, final) == 0
        out.append((state, done, status, led.bits(), f, act, final.arrays()))
    counts = fl.graph_counts()
    fl.close()
    return out, counts


def test_fleet_rounds_recorded_match_direct(hip_ctx, monkeypatch):
    d, c0 = fleet_run(hip_ctx, monkeypatch, False)
    g, c1 = fleet_run(hip_ctx, monkeypatch, True)
    print(f"fleet graphs recorded/replayed {c1}", flush=True)
    assert c0 == (0, 0) and c1[1] > 0
    for s, (a, b) in enumerate(zip(d, g)):
        assert a[:3] == b[:3] and a[0] == N.SLOT_DONE, s
        if s in (0, 3):  # exact ledgers
            assert a[3] == b[3], s
        else:
            assert a[3][4:] == b[3][4:], s
        assert np.array_equal(a[4].view(np.uint64), b[4].view(np.uint64)), s
        assert np.array_equal(a[5], b[5]), s
        (k0, v0), (k1, v1) = a[6], b[6]
        assert np.array_equal(k0, k1) and np.array_equal(v0.view(np.uint64), v1.view(np.uint64))

"""Batched and served FIDs (tpgp_fleet_*): every FID is bit for bit the same alone and in the mix.

"Alone" is the single-map resident loop (tpgp_hip_evolve), itself pinned to the oracle by
test_rotation_gpu.py; the powder case is also checked against the oracle directly.  Compared
per job: FID samples and their radii (bit patterns), sampled map lengths, the final map and the
ledger (exact mode: bit patterns).
"""
from __future__ import annotations

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from oracle import bindings as OB

pytestmark = pytest.mark.gpu


def sum_x(n):
    m = E.Map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    return m


def alone(ctx, n, terms, dt, steps, thr, exact, max_terms=10 ** 8):
    m = sum_x(n)
    o = E.options(threshold=thr, max_terms=max_terms, context=ctx)
    led = E.ledger(0)
    st, fid, rad, act, acc = E.evolve(m, terms, dt, steps, o, led, exact_ledger=exact)
    assert st in (0, E.N.ECAP), (st, E.N.engine().tpgp_hip_last_error().decode())
    return st, fid[:acc + 1], rad[:acc + 1], act[:acc + 1], m, led


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64),
                          np.asarray(b, np.float64).view(np.uint64))


def check_job(fl, slot, ref, n, exact=True):
    st, fid, rad, act, m, led = ref
    state, done, status, fled = fl.slot(slot)
    f, r, a = fl.fid(slot, len(fid) + 4)
    assert status == st, (slot, status, st)
    assert len(f) == len(fid), (slot, len(f), len(fid))
    assert same_bits(f, fid) and same_bits(r, rad), slot
    assert np.array_equal(a, act), slot
    fm = E.Map(n)
    assert fl.release(slot, fm) == 0
    ak, av = m.arrays()
    bk, bv = fm.arrays()
    assert np.array_equal(ak, bk) and same_bits(av, bv), slot
    if exact:
        assert led.bits() == fled.bits(), slot


def run_all(fl):
    for _ in range(100000):
        st, running = fl.step()
        assert st == 0, (st, E.N.engine().tpgp_hip_last_error().decode())
        if running == 0:
            return
    raise AssertionError("fleet did not finish")


def test_powder_batch_matches_alone_and_oracle(hip_ctx, oracle):
    """Six orientations of the 3x3x3 cluster (dt .01, threshold 1e-5), 3 steps, one fleet."""
    orients = H.powder_orientations(6)
    n, dt, thr, steps = 27, 0.01, 1e-5, 3
    fl = E.Fleet(1, len(orients), 1 << 14, context=hip_ctx)
    refs = []
    for k, o in enumerate(orients):
        terms = H.powder_terms(o)
        init = sum_x(n)
        assert fl.admit(k, init, terms, dt, steps,
                        E.options(threshold=thr, max_terms=10 ** 8), exact_ledger=True) == 0
        refs.append(alone(hip_ctx, n, terms, dt, steps, thr, True))
    run_all(fl)
    for k in range(len(orients)):
        check_job(fl, k, refs[k], n)
    # the first orientation against the oracle directly
    terms = H.powder_terms(orients[0])
    om = oracle.new_map(n)
    for s in range(n):
        om.add_single(s, 1, 1.0)
    oo = oracle.options(threshold=thr, max_terms=10 ** 8)
    ol = oracle.ledger(0)
    ot = OB.terms_array(terms)
    fid_o = []
    for k in range(steps + 1):
        fid_o.append(om.readout(1, n)[1])
        if k < steps:
            assert om.strang_step(ot, len(terms), dt, oo, ol) == 0
    assert same_bits(fid_o, refs[0][1])
    assert ol.bits() == refs[0][5].bits()
    fl.close()


@pytest.mark.parametrize("words,sizes", [(1, [8, 12, 20, 30, 16]), (2, [34, 40, 48])])
def test_continuous_batching_alone_equals_mixed(hip_ctx, words, sizes):
    """Jobs of different sizes, thresholds and lengths admitted at different times into a
    small fleet (slots reused as they free up); each result equals its alone run."""
    g = H.InputPCG(20260711)
    jobs = []
    for k, n in enumerate(sizes * 2):
        nn = [0.5 + g.uniform() for _ in range(n)]
        thr = [1e-4, 1e-5, 1e-6][k % 3]
        steps = 2 + (k * 5) % 7
        jobs.append((n, H.serving_ring(n, nn), thr, steps, k % 2 == 0))
    n_slots = 3
    fl = E.Fleet(words, n_slots, 1 << 12, context=hip_ctx)
    refs = [alone(hip_ctx, n, t, 0.01, st, thr, ex) for (n, t, thr, st, ex) in jobs]
    pending = list(range(len(jobs)))
    slot_job = [None] * n_slots
    finished = 0
    for it in range(10000):
        for s in range(n_slots):  # admit into free slots, one new job per round at most
            if slot_job[s] is None and pending:
                j = pending.pop(0)
                n, t, thr, st, ex = jobs[j]
                assert fl.admit(s, sum_x(n), t, 0.01, st,
                                E.options(threshold=thr, max_terms=10 ** 8),
                                exact_ledger=ex) == 0
                slot_job[s] = j
                break
        st, running = fl.step()
        assert st == 0
        for s in range(n_slots):
            j = slot_job[s]
            if j is None:
                continue
            state = fl.slot(s)[0]
            if state in (E.N.SLOT_DONE, E.N.SLOT_FAILED):
                check_job(fl, s, refs[j], jobs[j][0], exact=jobs[j][4])
                slot_job[s] = None
                finished += 1
        if finished == len(jobs):
            break
    assert finished == len(jobs)
    fl.close()


def test_fleet_growth_and_failure_isolation(hip_ctx):
    """Slots that outgrow a tiny capacity are grown and re-run transparently; a job that hits
    its max_terms cap fails alone (ECAP) while the others finish unchanged."""
    n, dt = 12, 0.01
    terms = H.ring(n)
    fl = E.Fleet(1, 3, 64, context=hip_ctx)
    specs = [(1e-8, 10 ** 8, 6), (1e-8, 1000, 8), (1e-6, 10 ** 8, 5)]
    refs = []
    for k, (thr, cap, steps) in enumerate(specs):
        assert fl.admit(k, sum_x(n), terms, dt, steps,
                        E.options(threshold=thr, max_terms=cap), exact_ledger=True) == 0
        refs.append(alone(hip_ctx, n, terms, dt, steps, thr, True, max_terms=cap))
    assert refs[1][0] == E.N.ECAP
    run_all(fl)
    assert fl.slot(1)[0] == E.N.SLOT_FAILED and fl.slot(1)[2] == E.N.ECAP
    for k in range(3):
        check_job(fl, k, refs[k], n)
    fl.close()


@pytest.mark.parametrize("limit", ["0", "100000000"])
def test_fleet_descriptor_ring_wraps(hip_ctx, limit, monkeypatch):
    """64 powder orientations: one step's 948 rotations span several descriptor chunks, so the
    staging ring wraps within a call (regression: every chunk must carry its own map list).
    Both device paths; every FID equal to its alone run."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", limit)
    K, n, dt, thr, steps = 64, 27, 0.01, 1e-5, 2
    orients = H.powder_orientations(K)
    fl = E.Fleet(1, K, 1 << 12, context=hip_ctx)
    for k, o in enumerate(orients):
        assert fl.admit(k, sum_x(n), H.powder_terms(o), dt, steps,
                        E.options(threshold=thr, max_terms=10 ** 8)) == 0
    run_all(fl)
    for k in (0, 17, 63):
        ref = alone(hip_ctx, n, H.powder_terms(orients[k]), dt, steps, thr, False)
        check_job(fl, k, ref, n, exact=False)
    fl.close()

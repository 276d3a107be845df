"""Serving on the GPU: admission by projected bytes, deferral, refusal, and per-slot failure.

* Server (ddevol_amd/serve.py): a stream under a budget that holds two or three requests at a
  time never holds more device bytes than the budget, every result bit for bit equal to the
  request run alone; requests projected large run in their own fleet beside the small ones; a
  request whose projection exceeds the budget is refused without stopping the others.
* Fleet budget (tpgp_fleet_set_budget): a slot that would have to grow beyond the budget fails
  alone with ENOMEM at its last accepted step; its neighbours finish unchanged.
* Trim (tpgp_fleet_trim): a free slot that grew for a finished job returns to the opening share,
  a running neighbour keeps its map and capacity, and the trimmed slot serves the next job
  exactly as a fresh one.
* Device memory: a re-layout that cannot be allocated (a capacity far beyond the device's memory,
  asked for from the host) returns ENOMEM and leaves every map as it was, and the batch runs on;
  a resize below a map's committed length is rejected with nothing changed.
"""
import ctypes as C

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from ddevol_amd import serve as S

pytestmark = pytest.mark.gpu


def alone_fid(ctx, req):
    m = S.initial_state(req.n)
    st, fid, rad, act, acc = E.evolve(m, req.terms, req.dt, req.steps,
                                      E.options(threshold=req.threshold, max_terms=10 ** 10,
                                                context=ctx), E.ledger(0))
    return st, fid[:acc + 1]


def small_stream():
    g = H.InputPCG(7)
    reqs = []
    for k, (n, thr, steps) in enumerate([(8, 1e-4, 30), (12, 1e-4, 20), (10, 1e-5, 25),
                                         (8, 1e-5, 15), (12, 1e-4, 35), (9, 1e-4, 10)]):
        nn = [0.5 + g.uniform() for _ in range(n)]
        reqs.append(S.Request(k, 0.01 * k, n, H.serving_ring(n, nn), thr, steps))
    return reqs


def test_server_defers_to_the_budget_and_matches_alone(hip_ctx):
    reqs = small_stream()
    slot_terms = 4096
    probe = S.Server(hip_ctx, slots_per_width=8, budget_bytes=1 << 40, slot_terms=slot_terms)
    need = max(probe.projected_bytes(r) for r in reqs)
    key = (1, "small")
    assert all(probe.fleet_of(r) == key for r in reqs)
    # every slot starts with slot_terms terms: the fleets' shares count whether used or not;
    # a running request adds what its reservation holds beyond its slot's share
    floor, share = probe.floor, probe.share[key]
    probe.close()
    extra = need - share
    assert extra > 0
    budget = floor + 2 * extra + extra // 2  # room for two of the largest, never for all six
    srv = S.Server(hip_ctx, slots_per_width=8, budget_bytes=budget, slot_terms=slot_terms)
    running_max = 0
    orig_step = srv.fleets[key].step

    def step():  # at every round: the requests in the fleet, and what they are charged
        nonlocal running_max
        running_max = max(running_max, srv.running(key))
        assert srv.held_bytes() <= budget
        assert sum(fl.held_bytes() for fl in srv.fleets.values()) <= budget
        return orig_step()

    srv.fleets[key].step = step
    print(f"budget {budget} bytes, reservation up to {need}", flush=True)
    srv.run(reqs)
    srv.close()
    assert not srv.refused
    assert 2 <= running_max < len(reqs)
    assert any(r.admitted > r.arrival + 1e-3 for r in reqs), "nothing had to wait"
    for r in reqs:
        st, fid = alone_fid(hip_ctx, r)
        assert r.status == st == 0
        assert np.array_equal(r.fid.view(np.uint64), fid.view(np.uint64)), r.rid


def test_server_runs_large_requests_apart(hip_ctx, monkeypatch):
    """Requests projected past LARGE_TERMS run in the large-map fleet, the rest in the small
    one; both fleets take rounds, and every FID equals the request's run alone."""
    monkeypatch.setattr(S, "LARGE_TERMS", 3000)
    reqs = small_stream()
    big = [r for r in reqs if S.projected_peak_terms(r.n, r.threshold, r.steps) >= 3000]
    assert 0 < len(big) < len(reqs)
    srv = S.Server(hip_ctx, slots_per_width=8, budget_bytes=1 << 34, slot_terms=4096,
                   large_slots=2)
    stepped = {}
    orig = srv._round

    def round_(k, clock):
        stepped[k] = stepped.get(k, 0) + 1
        return orig(k, clock)

    srv._round = round_
    srv.run(reqs)
    srv.close()
    print(f"rounds per fleet: {stepped}", flush=True)
    assert set(stepped) == {(1, "small"), (1, "large")}
    for r in reqs:
        st, fid = alone_fid(hip_ctx, r)
        assert r.status == st == 0
        assert np.array_equal(r.fid.view(np.uint64), fid.view(np.uint64)), r.rid


def test_server_refuses_what_cannot_fit(hip_ctx):
    reqs = small_stream()[:3]
    g = H.InputPCG(20260711)
    nn = [0.5 + g.uniform() for _ in range(48)]
    big = S.Request(99, 0.0, 48, H.serving_ring(48, nn), 1e-6, 1000)  # ~4e11 terms projected
    srv = S.Server(hip_ctx, slots_per_width=4, budget_bytes=4 << 30, slot_terms=4096)
    srv.run(reqs + [big])
    srv.close()
    assert srv.refused == [big] and big.status == -1
    for r in reqs:
        st, fid = alone_fid(hip_ctx, r)
        assert r.status == 0 and np.array_equal(r.fid.view(np.uint64), fid.view(np.uint64))


def test_fleet_budget_fails_only_the_slot_that_must_grow(hip_ctx):
    n_big, dt = 12, 0.01
    fl = E.Fleet(1, 3, 512, context=hip_ctx)  # the N=12 ring passes 512 terms at step 3
    fl.set_budget(fl.held_bytes() + 1024)  # no slot can grow by more than a few terms
    jobs = [(n_big, H.ring(n_big), 1e-8, 12), (4, H.ring(4), 1e-8, 12), (5, H.ring(5), 1e-8, 9)]
    for s, (n, terms, thr, steps) in enumerate(jobs):
        assert fl.admit(s, S.initial_state(n), terms, dt, steps,
                        E.options(threshold=thr, max_terms=10 ** 8)) == 0
    for _ in range(100):
        st, running = fl.step()
        assert st == 0, N.engine().tpgp_hip_last_error().decode()
        if running == 0:
            break
    state, done, status, _ = fl.slot(0)
    print(f"big slot: state {state}, status {status}, {done} steps accepted", flush=True)
    assert state == N.SLOT_FAILED and status == N.ENOMEM and 0 < done < jobs[0][3]
    # its samples up to the last accepted step equal the alone run's
    m = S.initial_state(n_big)
    _, fid, _, _, _ = E.evolve(m, jobs[0][1], dt, done,
                               E.options(threshold=1e-8, max_terms=10 ** 8, context=hip_ctx),
                               E.ledger(0))
    f, _, _ = fl.fid(0, 64)
    assert np.array_equal(f[:done + 1].view(np.uint64), fid.view(np.uint64))
    for s in (1, 2):
        n, terms, thr, steps = jobs[s]
        state, done, status, _ = fl.slot(s)
        assert state == N.SLOT_DONE and status == 0 and done == steps
        m = S.initial_state(n)
        _, fid, _, _, _ = E.evolve(m, terms, dt, steps,
                                   E.options(threshold=thr, max_terms=10 ** 8, context=hip_ctx),
                                   E.ledger(0))
        f, _, _ = fl.fid(s, 64)
        assert np.array_equal(f.view(np.uint64), fid.view(np.uint64))
    fl.close()


def test_trim_returns_free_slots_to_their_share(hip_ctx):
    dt = 0.01
    fl = E.Fleet(1, 2, 256, context=hip_ctx)
    share = fl.slot_bytes(0)
    jobs = [(10, H.ring(10), 1e-8, 6), (12, H.ring(12), 1e-8, 14)]
    for s, (n, terms, thr, steps) in enumerate(jobs):
        assert fl.admit(s, S.initial_state(n), terms, dt, steps,
                        E.options(threshold=thr, max_terms=10 ** 8)) == 0
    while fl.slot(0)[0] == N.SLOT_RUNNING:
        assert fl.step()[0] == 0
    assert fl.slot_bytes(0) > share and fl.slot(1)[0] == N.SLOT_RUNNING
    f0, _, _ = fl.fid(0, 64)
    assert fl.release(0) == 0
    grown1 = fl.slot_bytes(1)
    print(f"trim: slot 0 holds {fl.slot_bytes(0)} bytes, share {share}", flush=True)
    assert fl.trim() == 0, N.engine().tpgp_hip_last_error().decode()
    assert fl.slot_bytes(0) == share and fl.slot_bytes(1) == grown1
    # the trimmed slot runs the first job again while slot 1 finishes its own
    n, terms, thr, steps = jobs[0]
    assert fl.admit(0, S.initial_state(n), terms, dt, steps,
                    E.options(threshold=thr, max_terms=10 ** 8)) == 0
    for _ in range(100):
        st, running = fl.step()
        assert st == 0
        if running == 0:
            break
    g0, _, _ = fl.fid(0, 64)
    assert np.array_equal(f0.view(np.uint64), g0.view(np.uint64))
    for s, (n, terms, thr, steps) in enumerate(jobs):
        m = S.initial_state(n)
        _, fid, _, _, _ = E.evolve(m, terms, dt, steps,
                                   E.options(threshold=thr, max_terms=10 ** 8, context=hip_ctx),
                                   E.ledger(0))
        f, _, _ = fl.fid(s, 64)
        assert fl.slot(s)[:3] == (N.SLOT_DONE, steps, 0)
        assert np.array_equal(f.view(np.uint64), fid.view(np.uint64))
    fl.close()


def test_relayout_beyond_device_memory_keeps_every_map(hip_ctx):
    hip = N.hip()
    err = C.create_string_buffer(512)
    spins = (C.c_size_t * 2)(12, 12)
    caps = (C.c_size_t * 2)(1024, 1024)
    b = hip.tpgp_hip_batch_create(hip_ctx.ptr, 1, 2, spins, caps, err, 512)
    assert b, err.value
    rng = np.random.default_rng(3)
    maps = []
    for k in range(2):
        keys = np.unique(rng.integers(1, 4 ** 12, 500, dtype=np.uint64))
        vals = rng.standard_normal(len(keys))
        assert hip.tpgp_hip_batch_upload(b, k, N.as_u64p(keys), N.as_dp(vals), len(keys), err,
                                         512) == 0, err.value
        maps.append((keys, vals))
    want = (C.c_size_t * 2)(1 << 36, 0)  # 6.9e10 terms x 48 B: far beyond the device
    print("reserving a capacity beyond device memory (expect ENOMEM)", flush=True)
    st = hip.tpgp_hip_batch_reserve_many(b, want, err, 512)
    assert st == N.HIP_ENOMEM, (st, err.value)
    for k, (keys, vals) in enumerate(maps):
        out_k = np.zeros(len(keys), np.uint64)
        out_v = np.zeros(len(keys))
        n = C.c_size_t()
        assert hip.tpgp_hip_batch_download(b, k, N.as_u64p(out_k), N.as_dp(out_v), len(keys),
                                           C.byref(n), err, 512) == 0, err.value
        assert n.value == len(keys)
        assert np.array_equal(out_k, keys) and np.array_equal(out_v.view(np.uint64),
                                                             vals.view(np.uint64))
    cap = C.c_size_t()
    ln = C.c_size_t()
    assert hip.tpgp_hip_batch_info(b, 0, C.byref(ln), C.byref(cap)) == 0
    assert cap.value == 1024 and ln.value == len(maps[0][0])
    # a resize below the committed length is refused before anything moves
    small = (C.c_size_t * 2)(len(maps[0][0]) - 1, 0)
    assert hip.tpgp_hip_batch_resize(b, small, err, 512) == N.HIP_EINVAL, err.value
    assert hip.tpgp_hip_batch_info(b, 0, C.byref(ln), C.byref(cap)) == 0
    assert cap.value == 1024 and ln.value == len(maps[0][0])
    # and it still grows within reason
    want = (C.c_size_t * 2)(4096, 2048)
    assert hip.tpgp_hip_batch_reserve_many(b, want, err, 512) == 0, err.value
    hip.tpgp_hip_batch_destroy(b)
    # the failed allocation is not reported again by the next call's launch check
    st, _, _, _, acc = E.evolve(S.initial_state(8), H.ring(8), 0.01, 3,
                                E.options(threshold=1e-8, max_terms=10 ** 6, context=hip_ctx),
                                E.ledger(0))
    assert st == 0 and acc == 3, N.engine().tpgp_hip_last_error().decode()

#!/usr/bin/env python3
"""Benchmark: FP64 Pauli branch-updates/s on the N=16 J1-J2 FID, one MI355X.

Workload (the metric's configuration): N=16 periodic J1-J2 ring, J1 = 1, J2_i = 0.1 + 0.3 u_i
with u_i the first 16 uniforms of the reference's seeded InputPCG(7); secular dipolar terms
(96 per step order); observable sum_s X_s with coefficients 1; Strang steps dt = 0.005;
truncation |c| <= 1e-6; binary64 throughout.  Synthetic by construction: no dataset.

A "step" is one Strang step: 192 rotations over the resident map.  The headline is quoted at
the FID's operating point: untimed steps 1..600 grow the map from 16 terms to ~1.43e7 (8-9 s on
MI355X), whatever --warmup says; --warmup W adds W further untimed steps there, and the timed
region is the next K steps (601+W .. 600+W+K), bracketed by device synchronisation, with the map
already resident in HBM.  --start-step moves the operating point (0 with --warmup 0 times a whole
FID).  One branch-update = one input term processed by one executed rotation (the reference's
ledger.rotations counts the same rotations, engine/src/tpgp_rev.c:740).

roofline: the dominant kernel is the leaf kernel kind (rotation write, count, splits, scans,
partition histogram and scatter, small-map kernel) with the largest total time over the sampled
launches of the timed region.  Algorithmic bytes per launch, for W key words (16 B per term for
N <= 32): k_rot_write (8W + 8)(len_in + len_out - 2 in_place) -- every input term of the classes
the rotation rewrites read once and every output written once (an appending rotation leaves the
classes it does not change in place: in_place terms neither read nor written); k_rot_count
(8W + 8)(len_in - in_place); k_small (8W + 8)(len_in + len_out) per rotation it runs.
Durations are measured live with HIP events recorded on the batch stream around the launches of
a sample of the timed calls.

cpu_baseline: the reference CPU engine (engine/src/tpgp_rev.c compiled from its own sources into
oracle/_ref) -- or, where that build is absent, the oracle's C restatement -- runs the J1-J2
rotation sequence, repeated, from the same map state (the state at the start of the timed
region) with its OpenMP branch backend for --cpu-seconds (default 20 s) on every core of this
job's CPU share, then again on one thread.  The CPU model and thread counts are reported.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public specification
# HBM bytes per launch of k_rot_write over the default window (untimed steps 1..605, timed steps
# 606..625), from rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a plain copy's readings (factors
# 2.00 and 1.00): profiles/k_rot_write_j1j2_n16_headline_pmc.md.  Used for roofline.traffic when
# the run is that window and no --traffic is given; other windows report null.
MEASURED_TRAFFIC = {("k_rot_write", 605, 20): 251.6e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def host_threads():
    """Threads for the CPU baseline: every core this job may use.  On a shared host the job's
    share is what OMP_NUM_THREADS grants (its CPU allotment); otherwise the cores in this
    process's affinity mask."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return min(int(env), avail or int(env)), avail
    return max(1, avail or 1), avail


def cpu_baseline(keys, values, n_spins, terms, dt, threshold, budget_s, threads):
    """Time the reference (or oracle) CPU path on the same state: the Strang rotation sequence,
    repeated from the given map state (the evolution simply continues), until `budget_s` has
    passed; returns a dict."""
    from oracle import bindings as OB

    kind = "reference" if os.path.exists(OB.REF_SO) else "port"
    if kind == "port" and not os.path.exists(OB.ORACLE_SO):
        OB.build()
    eng = OB.CpuEngine("ref" if kind == "reference" else "oracle")
    m = eng.new_map(n_spins)
    if kind == "reference":
        # install the arrays directly (malloc'ed: the engine frees them on replacement)
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]
        n = len(values)
        kp = libc.malloc(max(keys.nbytes, 8))
        vp = libc.malloc(max(values.nbytes, 8))
        C.memmove(kp, keys.ctypes.data, keys.nbytes)
        C.memmove(vp, values.ctypes.data, values.nbytes)
        mm = m.ptr.contents
        mm.keys = C.cast(kp, C.POINTER(C.c_uint64))
        mm.values = C.cast(vp, C.POINTER(C.c_double))
        mm.len = mm.capacity = n
    else:
        assert m.assign(keys, values) == 0
    half = dt * 0.5
    seq = list(terms) + list(reversed(terms))
    opts = eng.options(threshold=threshold, threads=threads, backend=OB.OPENMP)
    led = eng.ledger(0)
    updates = 0
    rots = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        i, j, qi, qj, h = seq[rots % len(seq)]
        n_in = len(m)
        # tpgp_rotate(theta=h*half) evaluates cos/sin of 2*(h*half)*1, which equals the step's
        # 2*h*half bit for bit (scaling by 2 is exact), so the work is the step's rotation
        st = eng.rotate(m.ptr, i, j, qi, qj, h * half, C.byref(opts), C.byref(led))
        if st:
            raise RuntimeError(f"CPU baseline rotation failed: status {st}")
        updates += n_in
        rots += 1
    el = time.perf_counter() - t0
    return {"value": updates / el, "unit": "branch-updates/s", "cores": threads, "kind": kind,
            "rotations": rots, "seconds": el, "terms_at_start": len(values),
            "terms_at_end": len(m)}


def same_bits(a, b):
    import numpy as np

    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return a.shape == b.shape and bool(np.array_equal(a.view(np.uint64), b.view(np.uint64)))


def bench_powder(args):
    """Batched powder average: N=27 3x3x3 cluster, `orientations` field directions from the
    seeded PCG, dt 0.01, threshold 1e-5, a window of `powder_steps` Strang steps per FID, all
    orientations in one fleet; against the one-FID-at-a-time loop on the first `alone` of them
    (whose FIDs must be bit-identical to the batched ones)."""
    from ddevol_amd import engine as E
    from ddevol_amd import hamiltonians as H

    ctx = E.Context(0)
    K, steps, n, dt, thr = args.orientations, args.powder_steps, 27, 0.01, 1e-5
    orients = H.powder_orientations(K)
    terms = [H.powder_terms(o) for o in orients]
    fl = E.Fleet(1, K, 1 << 15, context=ctx)
    for k in range(K):
        init = E.Map(n)
        for s in range(n):
            init.add_single(s, 1, 1.0)
        st = fl.admit(k, init, terms[k], dt, steps, E.options(threshold=thr, max_terms=10 ** 10))
        assert st == 0, st
    t0 = time.perf_counter()
    rounds = 0
    while True:
        st, running = fl.step()
        assert st == 0, N_last_error()
        rounds += 1
        if running == 0:
            break
        if rounds % 10 == 0:  # a progress line per 10 fleet rounds (Strang steps)
            log(f"powder batched: round {rounds}, {running} running, "
                f"{time.perf_counter() - t0:.1f} s")
    t_batch = time.perf_counter() - t0
    c = fl.counters()
    fids = [fl.fid(k, steps + 1)[0] for k in range(K)]
    peak = max(int(fl.fid(k, steps + 1)[2].max()) for k in range(K))
    fl.close()
    log(f"powder batched: {K} FIDs x {steps} steps in {t_batch:.2f} s ({rounds} rounds), "
        f"{c.updates} updates, largest map {peak} terms")
    # one FID at a time (a resident evolution session per orientation, in chunks of 25 steps
    # with a progress line each: the samples are those of one call)
    import numpy as np

    A = min(args.alone, K)
    t0 = time.perf_counter()
    same = True
    for k in range(A):
        m = E.Map(n)
        for s in range(n):
            m.add_single(s, 1, 1.0)
        led = E.ledger(0)
        ev = E.Evolution(m, terms[k], dt, E.options(threshold=thr, max_terms=10 ** 10,
                                                    context=ctx), context=ctx)
        parts, done = [], 0
        while done < steps:
            st, fid, rad, act, acc = ev.run(min(25, steps - done), led)
            assert st == 0, N_last_error()
            parts.append(fid[:acc + 1] if not parts else fid[1:acc + 1])
            done += acc
            log(f"powder one-at-a-time: orientation {k}, step {done}, {int(act[acc])} terms, "
                f"{time.perf_counter() - t0:.1f} s")
        ev.close()
        same &= same_bits(np.concatenate(parts), fids[k])
    t_alone = time.perf_counter() - t0
    batched, alone = K / t_batch, A / t_alone
    log(f"powder one-at-a-time: {A} FIDs in {t_alone:.2f} s; identical FIDs: {same}")
    return {
        "metric": f"batched powder FIDs/s (N=27 3x3x3 cluster, {steps}-step FID window)",
        "value": batched, "unit": "FIDs/s", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (orientations from InputPCG(20260711); initial state sum_s X_s)",
        "config": {"workload": f"powder N=27, {K} orientations, dt 0.01, threshold 1e-5, "
                               f"steps 0..{steps} per orientation, one fleet",
                   "rounds": rounds, "branch_updates": c.updates,
                   "updates_per_s": c.updates / t_batch, "largest_map_terms": peak},
        "one_at_a_time": {"value": alone, "unit": "FIDs/s", "sample": f"first {A} orientations",
                          "fids_identical": same},
        "batched_over_one_at_a_time": batched / alone,
    }


def bench_residual(args):
    """The independent endpoint (residual) bound of one N=16 J1-J2 Strang step
    (tpgp_residual_bound, residual_rev.c:181-194, incident-terms generator with the Gram
    enclosure, depth --residual-depth) at the operating point --start-step: its interval vectors
    D1, D2 built on the GPU, then the serial Gram and norm sums on the host; against the
    reference library's CPU bound (oracle/_ref, one thread: the reference's residual code has no
    threads) on the same two maps, which must give the same bits."""
    import numpy as np

    from ddevol_amd import engine as E
    from ddevol_amd import hamiltonians as H
    from ddevol_amd import native as N

    ctx = E.Context(0)
    n, dt, thr, depth = 16, 0.005, 1e-6, args.residual_depth
    terms = H.j1j2_n16_terms(n)
    a = E.Map(n)
    for s in range(n):
        a.add_single(s, 1, 1.0)
    o = E.options(threshold=thr, max_terms=10 ** 10, context=ctx)
    st, *_ = E.evolve(a, terms, dt, args.start_step, o, E.ledger(0))
    assert st == 0, N_last_error()
    keys, vals = a.arrays()
    b = E.Map.from_sorted(n, keys, vals)
    st, *_ = E.evolve(b, terms, dt, 1, o, E.ledger(0))
    assert st == 0, N_last_error()
    log(f"residual: maps of {len(a)} -> {len(b)} terms after {args.start_step} steps")
    # the device part alone: D1 and D2
    err = C.create_string_buffer(512)
    out = (N.HipIntervalVector * 2)()
    ta = N.terms_array(terms)
    hip = N.hip()
    t0 = time.perf_counter()
    rc = hip.tpgp_hip_residual_vectors(ctx.ptr, a.m.keys, a.m.values, a.m.len, b.m.keys,
                                       b.m.values, b.m.len, 1, n, ta, len(terms), dt, 1, out,
                                       err, 512)
    t_vec = time.perf_counter() - t0
    assert rc == 0, err.value
    sizes = (out[0].len, out[1].len)
    for d in range(2):
        hip.tpgp_hip_interval_vector_free(C.byref(out[d]))
    # the whole bound, best of --residual-repeats
    times = []
    for _ in range(max(1, args.residual_repeats)):
        t0 = time.perf_counter()
        st, up = E.residual_bound(a, b, terms, dt, depth)
        times.append(time.perf_counter() - t0)
        assert st == 0, N_last_error()
    t_gpu = min(times)
    log(f"residual: D1 {sizes[0]} and D2 {sizes[1]} entries; vectors {t_vec:.3f} s; "
        f"bound {up:.6e} in {t_gpu:.3f} s")
    cpu = None
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libtpgp_ref.so")
    if not args.no_cpu and os.path.exists(ref_so):
        lib = C.CDLL(ref_so)
        fn = lib.tpgp_residual_bound
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double, C.c_uint,
                       C.POINTER(C.c_double)]
        rup = C.c_double(float("nan"))
        t0 = time.perf_counter()
        rst = fn(C.addressof(a.m), C.addressof(b.m), C.cast(ta, C.c_void_p), len(terms), dt,
                 depth, C.byref(rup))
        t_cpu = time.perf_counter() - t0
        same = rst == 0 and np.float64(up).view(np.uint64) == np.float64(rup.value).view(
            np.uint64)
        log(f"residual: reference CPU bound {rup.value:.6e} in {t_cpu:.2f} s; same bits: {same}")
        cpu = {"value": 1.0 / t_cpu, "unit": "bounds/s", "cores": 1, "kind": "reference",
               "seconds": t_cpu, "same_bits": bool(same), "cpu_model": cpu_model(),
               "sample": "one bound on the same two maps (the reference's residual code is "
                         "single-threaded)"}
    return {
        "metric": "residual bounds/s (N=16 J1-J2 Strang step, tpgp_residual_bound)",
        "value": 1.0 / t_gpu, "unit": "bounds/s", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "256-bit intervals (MPFR-exact), f64 result",
        "data": "synthetic (seeded couplings; initial state sum_s X_s)",
        "config": {"workload": f"N=16 J1-J2 ring, dt 0.005, threshold 1e-6: the bound of Strang "
                               f"step {args.start_step + 1} (maps of {len(a)} -> {len(b)} "
                               f"terms), depth {depth}, incident-terms generator with the Gram "
                               f"enclosure",
                   "d1_entries": sizes[0], "d2_entries": sizes[1], "bound": up},
        "seconds": t_gpu, "device_vectors_seconds": t_vec,
        "host_sums_seconds": max(t_gpu - t_vec, 0.0),
        "cpu_baseline": cpu,
        "gpu_over_cpu": (cpu["seconds"] / t_gpu) if cpu else None,
    }


def N_last_error():
    from ddevol_amd import native as N

    return N.engine().tpgp_hip_last_error().decode()


def bench_serve(args):
    """Serving stream with continuous batching (ddevol_amd.serve) against the one-FID-at-a-time
    loop on the same requests; p50/p99 of completion - arrival over the served requests,
    FIDs/s, identical FIDs.  Requests whose projected term-array bytes exceed the budget even
    beside empty slots are refused by admission and counted; the one-at-a-time loop runs the
    served ones.  --serve-phase runs one side only (served: the stream; alone: the loop over the
    requests the stream would serve) so that long streams fit separate runs."""
    import numpy as np

    from ddevol_amd import engine as E
    from ddevol_amd import serve as S

    ctx = E.Context(0)
    lo, hi = args.serve_steps
    reqs = S.poisson_stream(args.requests, args.rate, seed=7, steps_range=(lo, hi))
    srv = S.Server(ctx, slots_per_width=args.slots, budget_bytes=int(args.budget_gb * 2 ** 30))
    t_srv = rounds = None
    if args.serve_phase in ("both", "served"):
        t0 = time.perf_counter()
        srv.run(reqs, log=log)
        t_srv = time.perf_counter() - t0
        rounds = srv.rounds
        refused = {r.rid for r in srv.refused}
    else:  # the admission decision alone: the same requests as the stream would refuse
        refused = {r.rid for r in reqs if not srv.fits(r)}
    srv.close()
    served = [r for r in reqs if r.rid not in refused]
    pct = lambda a, q: float(np.percentile(a, q))  # noqa: E731
    out = {
        "metric": "served FIDs/s with p50/p99 latency (continuous batching)",
        "value": None, "unit": "FIDs/s", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Poisson stream, InputPCG(7))",
        "config": {"workload": f"{len(reqs)} requests, Poisson rate {args.rate}/s, N in 8..48 "
                               f"rings (NN U[0.5,1.5], NNN 1/8), threshold in "
                               f"{{1e-4,1e-5,1e-6}}, steps U{{{lo}..{hi}}}, dt 0.01",
                   "hbm_budget_gb": args.budget_gb, "slots_per_key_width": args.slots,
                   "rounds": rounds, "refused": sorted(refused), "served": len(served),
                   "phase": args.serve_phase},
    }
    if t_srv is not None:
        ok = [r for r in served if r.status == 0]
        lat = np.array([r.latency for r in served])
        out["value"] = len(ok) / t_srv
        out["seconds"] = t_srv
        out["latency_s"] = {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": float(lat.max())}
        out["failed"] = [r.rid for r in served if r.status != 0]
        log(f"served {len(ok)} of {len(reqs)} requests in {t_srv:.2f} s ({rounds} rounds), "
            f"{len(refused)} refused by admission")
    if args.serve_phase in ("both", "alone"):
        alone = [S.Request(r.rid, r.arrival, r.n, r.terms, r.threshold, r.steps, r.dt)
                 for r in served]
        t0 = time.perf_counter()
        S.run_alone(ctx, alone, log=log)
        t_alone = time.perf_counter() - t0
        lat_a = np.array([r.latency for r in alone])
        one = {"value": sum(r.status == 0 for r in alone) / t_alone, "unit": "FIDs/s",
               "seconds": t_alone,
               "latency_s": {"p50": pct(lat_a, 50), "p99": pct(lat_a, 99)}}
        if t_srv is not None:
            one["fids_identical"] = all(same_bits(a.fid, b.fid) for a, b in zip(served, alone))
            out["batched_over_one_at_a_time"] = t_alone / t_srv
        out["one_at_a_time"] = one
        log(f"one at a time: {len(alone)} requests in {t_alone:.2f} s")
    return out


def fid_allowance(E, terms, dt, led, n, steps):
    """The run driver's FID allowance after `steps` Strang steps: the per-step Strang bound summed
    upward, plus the ledger's (dropped + arithmetic) l2 over sqrt(N)."""
    per_step, pf = E.strang_bound(terms, dt), 0.0
    for _ in range(steps):
        pf = E.N.engine().tpgp_bound_add_up(pf, per_step)
    return E.fid_allowance(pf, led, n)


def sampled_calls(steps, call_steps, sample_every):
    """Indices of the calls (of `call_steps` steps) that a window of `steps` steps times with HIP
    events: one in `sample_every`, in pairs mirrored about the window's centre, so the sample's
    mean position is the window's; windows shorter than two periods time every call."""
    n_calls = -(-steps // call_steps)
    if sample_every <= 1 or n_calls < 2 * sample_every:
        return list(range(n_calls))
    return [c for c in range(n_calls)
            if (c if 2 * c < n_calls else n_calls - 1 - c) % sample_every == 1]


def timed_window(ev, led, steps, kinds, sample_every, progress=None):
    """`steps` Strang steps on the resident map, timed end to end, as calls of a few steps (one
    readout sync each).  HIP events around every launch cost a few microseconds each, so the
    kernel durations are sampled: events are on for one call in `sample_every`, chosen in pairs
    mirrored about the window's centre so the sample sees the window's mean map size while the
    map grows (short windows time every call), and the algorithmic counts of the roofline are
    taken over the same calls (engine counters)."""
    from ddevol_amd import native as N

    call_steps = 2 if steps <= 40 else 5 if steps <= 100 else 25
    sampled = set(sampled_calls(steps, call_steps, sample_every))

    mask = sum(1 << k for k in kinds)
    prof = {N.PROFILE_NAMES[k]: [0.0, 0] for k in kinds}
    w = {"s_updates": 0, "s_in_out": 0, "s_in_place": 0, "s_steps": 0, "call_steps": call_steps,
         "sample_every": sample_every, "steps": steps}
    c0 = ev.counters()
    t0 = time.perf_counter()
    done = call = 0
    while done < steps:
        chunk = min(call_steps, steps - done)
        timed = call in sampled
        ev.profile(mask if timed else 0)  # resets the event totals
        ca = ev.counters()
        st, fid, rad, act, acc = ev.run(chunk, led)
        if st:
            raise RuntimeError(f"timed run failed {done + acc} steps into the window: status "
                               f"{st}: {N.engine().tpgp_hip_last_error().decode()}")
        if timed:
            cb = ev.counters()
            for k in kinds:
                ms, cnt = ev.profile_read(k)
                prof[N.PROFILE_NAMES[k]][0] += ms
                prof[N.PROFILE_NAMES[k]][1] += cnt
            w["s_updates"] += cb.updates - ca.updates
            w["s_in_out"] += cb.in_out - ca.in_out
            w["s_in_place"] += cb.in_place - ca.in_place
            w["s_steps"] += acc
        if call == 0:
            w["terms_first"], w["fid_first"] = int(act[0]), float(fid[0])
        done += acc
        call += 1
        if progress and done < steps:
            progress(done, int(act[-1]), time.perf_counter() - t0)
    w["elapsed"] = time.perf_counter() - t0
    c1 = ev.counters()
    ev.profile(0)
    w["prof"] = {k: tuple(v) for k, v in prof.items()}
    w["updates"] = c1.updates - c0.updates
    w["in_out"] = c1.in_out - c0.in_out
    w["in_place"] = c1.in_place - c0.in_place
    w["rotations"] = c1.rotations - c0.rotations
    w["terms_last"], w["fid_last"] = int(act[-1]), float(fid[-1])
    w["sample"] = (f"HIP events on {w['s_steps']} of the {steps} timed steps (calls of "
                   f"{call_steps} steps, 1 in {sample_every}, in pairs mirrored about the window "
                   f"centre); algorithmic bytes over the same launches")
    return w


def bench_single(args, n, terms, dt, thr, label, workload):
    """One large FID: untimed warmup steps, then timed steps on the resident map."""
    from ddevol_amd import engine as E
    from ddevol_amd import native as N

    ctx = E.Context(0)
    mp = E.Map(n)
    for s in range(n):
        mp.add_single(s, 1, 1.0)
    opts = E.options(threshold=thr, max_terms=10 ** 11, context=ctx)
    led = E.ledger(0)
    done = 0
    if args.load_state:  # the map after an earlier run's warmup (--save-state)
        import numpy as np

        st = np.load(args.load_state)
        mp = E.Map.from_sorted(n, st["keys"], st["values"])
        done = int(st["warmup"])
        log(f"{label}: loaded state after {done} steps: {len(mp)} terms")
    grow = (done if args.load_state else args.start_step) + args.warmup
    ev = E.Evolution(mp, terms, dt, opts, context=ctx,
                     initial_capacity=4 * len(mp) + (1 << 20) if args.load_state else 0)
    size = len(mp)
    tw = time.perf_counter()
    # untimed, in chunks with a progress line each; --until-terms: until the map reaches that
    # many terms (the start step then bounds the steps)
    while done < grow and not (args.until_terms and size >= args.until_terms):
        chunk = min(5 if args.until_terms else 25, grow - done)
        st, fid, rad, act, acc = ev.run(chunk, led)
        assert st == 0, N_last_error()
        done += chunk
        size = int(act[-1])
        log(f"{label}: warmup step {done}: {size} terms")
    warm_s = time.perf_counter() - tw
    if args.save_state:
        import numpy as np

        assert ev.close(mp) == 0
        keys, vals = mp.arrays()
        np.savez(args.save_state, keys=keys, values=vals, warmup=done)
        log(f"{label}: saved {len(vals)} terms after {done} steps to {args.save_state}")
        return {"saved": args.save_state, "terms": len(vals), "warmup": done}
    ev.run(0, led)
    kinds = (list(range(len(N.PROFILE_NAMES))) if args.breakdown
             else [N.PROFILE_MERGE, N.PROFILE_COUNT])
    w = timed_window(ev, led, args.steps, kinds, 1 if args.breakdown else args.sample_every)
    ev.close()
    el, prof = w["elapsed"], w["prof"]
    log(f"{label}: steps {done + 1}..{done + args.steps}, {el:.2f} s, "
        f"terms {w['terms_first']} -> {w['terms_last']}")
    term_bytes = 8 * ((n + 31) // 32) + 8
    ms_w, n_w = prof["k_rot_write"]
    roof = None
    if n_w and ms_w > 0:
        # every input term of the rewritten classes read once, every output written once
        # (sampled launches)
        alg = term_bytes * (w["s_in_out"] - w["s_in_place"]) / n_w
        avg_s = ms_w * 1e-3 / n_w
        roof = {"bound": "hbm", "achieved": alg / avg_s / 1e9, "peak": 8000.0, "unit": "GB/s",
                "frac": alg / avg_s / 8e12, "traffic": None, "kernel": "k_rot_write",
                "launches": n_w, "sample": w["sample"], "avg_us": avg_s * 1e6,
                "algorithmic_bytes_per_launch": alg,
                "kernels": {k: {"launches": v[1], "avg_us": v[0] * 1e3 / max(v[1], 1)}
                            for k, v in prof.items()}}
    return {
        "metric": f"FP64 Pauli branch-updates/s ({label})", "value": w["updates"] / el,
        "unit": "branch-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": done,
        "ms_per_step": el / max(args.steps, 1) * 1e3, "higher_is_better": True,
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (initial state sum_s X_s)",
        "config": {"workload": workload, "terms_at_start": w["terms_first"],
                   "terms_at_end": w["terms_last"], "fid": [w["fid_first"], w["fid_last"]],
                   "ledger_dropped_l2": led.dropped_l2, "warmup_seconds": warm_s,
                   "whole_run_seconds": warm_s + el},
        "roofline": roof,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="j1j2",
                    choices=["j1j2", "powder", "serve", "large", "plain", "residual"],
                    help="j1j2: the headline N=16 J1-J2 window (default); powder: batched "
                         "orientations vs one at a time; serve: continuous batching with "
                         "p50/p99; large: N=64 range-4 ring; plain: N=12 ring; residual: the "
                         "residual bound of one J1-J2 step against the reference's CPU bound")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="calibration only: run 5 plain device-to-device copies of this many "
                         "bytes (for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--orientations", type=int, default=512)
    ap.add_argument("--powder-steps", type=int, default=20)
    ap.add_argument("--alone", type=int, default=16,
                    help="powder: orientations run one at a time for the comparison")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rate", type=float, default=16.0)
    ap.add_argument("--serve-steps", type=int, nargs=2, default=[100, 2000],
                    help="serve: the range of Strang steps per request (the config's 100..2000)")
    ap.add_argument("--serve-phase", default="both", choices=["both", "served", "alone"],
                    help="serve: run the stream, the one-at-a-time loop, or both")
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--budget-gb", type=float, default=192.0,
                    help="serve: the HBM budget admission charges projected term arrays to")
    ap.add_argument("--steps", type=int, default=20, help="timed Strang steps")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps at the operating point, after the untimed growth to "
                         "--start-step and before the timed steps")
    ap.add_argument("--start-step", type=int, default=None,
                    help="the operating point: untimed Strang steps that grow the map before "
                         "--warmup (default 600, where the N=16 J1-J2 map holds ~1.4e7 terms; "
                         "475 for --workload large, so that the timed steps are the last 20 of "
                         "its 500-step FID; 0 with --warmup 0 times a whole FID)")
    ap.add_argument("--sample-every", type=int, default=4,
                    help="headline: HIP events on every Nth call of the timed window (kernel "
                         "durations for the roofline); 1 = every call")
    ap.add_argument("--until-terms", type=int, default=0,
                    help="large/plain: warm up until the map holds this many terms (--warmup "
                         "then bounds the warmup steps)")
    ap.add_argument("--residual-depth", type=int, default=1,
                    help="residual: subdivision depth of the bound")
    ap.add_argument("--residual-repeats", type=int, default=3,
                    help="residual: timed bounds (the best is reported)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--breakdown", action="store_true",
                    help="time every kernel kind with HIP events (adds a little per-launch cost)")
    ap.add_argument("--save-state", help="after the warmup, save the map to this .npz and exit")
    ap.add_argument("--traffic", type=float, default=None,
                    help="measured HBM bytes per launch of the dominant kernel (rocprofv3 "
                         "FETCH_SIZE + WRITE_SIZE, calibrated), reported as roofline.traffic")
    ap.add_argument("--load-state", help="start the timed region from a state saved by "
                                         "--save-state (its warmup count is used)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the timed window's kernel table from a rocprofv3 --kernel-trace "
                         "output directory of a bench run (needs --bench-json) and exit")
    ap.add_argument("--summarize-pmc", metavar="DIR", nargs="+",
                    help="print the timed window's counter averages from rocprofv3 --pmc "
                         "output directories of a bench run (needs --bench-json) and exit")
    ap.add_argument("--copy-pmc", metavar="DIR", nargs="*",
                    help="with --summarize-pmc: the same counters for a --copy-bytes run")
    ap.add_argument("--bench-json", help="the JSON line the profiled bench run printed")
    args = ap.parse_args()
    if args.summarize_trace or args.summarize_pmc:
        with open(args.bench_json) as fh:
            line = json.loads([s for s in fh.read().splitlines() if s.startswith("{")][-1])
        rot = line["config"]["rotations"]
        if args.summarize_trace:
            print(summarize_trace(args.summarize_trace, rot, line["steps"]))
        else:
            print(summarize_pmc(args.summarize_pmc, rot, args.copy_pmc))
        return
    if args.start_step is None and args.workload != "j1j2":
        args.start_step = {"large": 475, "residual": RESIDUAL_START_STEP}.get(
            args.workload, HEADLINE_START_STEP)
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one FID per GPU: only --gpus 1 is supported; "
                                   "multi-GPU sharding is a design, not part of this build"}))
        sys.exit(2)
    if args.copy_bytes:
        import torch

        src = torch.ones(args.copy_bytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "copies": 5}), flush=True)
        return
    if args.workload != "j1j2":
        from ddevol_amd import hamiltonians as H

        if args.workload == "powder":
            out = bench_powder(args)
        elif args.workload == "serve":
            out = bench_serve(args)
        elif args.workload == "residual":
            out = bench_residual(args)
        elif args.workload == "large":
            out = bench_single(args, 64, H.range_ring(64, 4), 0.01, 1e-5,
                               "N=64 range-4 ring",
                               "N=64 ring, 1/r^3 couplings r=1..4, dt 0.01, threshold 1e-5")
        else:
            out = bench_single(args, 12, H.ring(12), 0.01, 1e-8, "N=12 ring",
                               "N=12 NN ring, dt 0.01, threshold 1e-8")
        print(json.dumps(out), flush=True)
        return

    print(json.dumps(bench_headline(args)), flush=True)


def roofline_of(prof, term_bytes, s_updates, s_in_out, sample, s_in_place=0):
    """Roofline entry for the dominant kernel: the leaf kernel kind with the largest total time
    over the sampled launches, among those with an algorithmic byte model.  Algorithmic bytes
    per launch (B = 8W + 8 bytes per term; in_place: terms of the classes an appending rotation
    left in place, 2 per term in s_in_place): k_rot_write B(len_in + len_out - 2 in_place) --
    every input term of the rewritten classes read once, every output written once;
    k_rot_count B(len_in - in_place); k_small B(len_in + len_out) per rotation it runs (only
    when it ran every sampled rotation, so the counters are its own)."""
    models = {"k_rot_write": term_bytes * (s_in_out - s_in_place),
              "k_rot_count": term_bytes * (s_updates - s_in_place // 2)}
    if prof.get("k_rot_write", (0, 0))[1] == 0 and prof.get("k_small", (0, 0))[1]:
        models = {"k_small": term_bytes * s_in_out}
    leaf = {k: v for k, v in prof.items() if k not in ("rotation_pass", "partition_pass")}
    ranked = sorted(leaf, key=lambda k: leaf[k][0], reverse=True)
    dom = next((k for k in ranked if k in models and leaf[k][1]), None)
    kernels = {k: {"launches": c, "avg_us": ms / max(c, 1) * 1e3, "total_ms": ms}
               for k, (ms, c) in leaf.items() if c}
    if dom is None:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None, "kernel": None, "sample": sample,
                "kernels": kernels}
    ms, cnt = leaf[dom]
    alg = models[dom] / cnt
    avg_s = ms * 1e-3 / cnt
    achieved = alg / avg_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom, "launches": cnt,
            "avg_us": avg_s * 1e6, "algorithmic_bytes_per_launch": alg,
            "largest_kind_by_time": ranked[0] if ranked else None, "sample": sample,
            "kernels": kernels}


def _rocprof_csvs(path, suffix):
    import glob

    if os.path.isfile(path):
        return [path]
    return sorted(glob.glob(os.path.join(path, "**", f"*{suffix}"), recursive=True))


def _demangle(names):
    """Demangle Itanium C++ names (as a kernel trace may record them) with c++filt, when
    present; other names pass through."""
    import shutil
    import subprocess

    mangled = sorted({n for n in names if n.startswith("_Z")})
    tool = shutil.which("c++filt") or shutil.which("llvm-cxxfilt") or \
        (os.path.exists("/opt/rocm/llvm/bin/llvm-cxxfilt") and "/opt/rocm/llvm/bin/llvm-cxxfilt")
    if not mangled or not tool:
        return {}
    out = subprocess.run([tool], input="\n".join(mangled), capture_output=True, text=True)
    lines = out.stdout.splitlines()
    return dict(zip(mangled, lines)) if len(lines) == len(mangled) else {}


def _short_name(name):
    """'void k_rot_write<1>(BatchDev, ...)' -> 'k_rot_write<1>'."""
    name = name.strip().replace("(anonymous namespace)::", "")
    if name.endswith(".kd"):
        name = name[:-3]
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for k, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:k].strip()
    return name


def summarize_trace(path, rotations, steps, anchor="k_rot_write"):
    """Markdown table of a rocprofv3 --kernel-trace over the timed window only.  The window is
    found from the run itself: every rotation of the multi-kernel path launches the anchor kernel
    once, so the window is everything dispatched after the anchor launch that precedes the last
    `rotations` ones (the timed region is the run's last GPU work; the CPU baseline launches
    nothing)."""
    import csv

    rows = []
    for f in _rocprof_csvs(path, "kernel_trace.csv"):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    if not rows:
        raise SystemExit(f"no kernel_trace.csv under {path}")
    dm = _demangle([r["Kernel_Name"] for r in rows])
    recs = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                   _short_name(dm.get(r["Kernel_Name"], r["Kernel_Name"]))) for r in rows)
    idx = [k for k, r in enumerate(recs) if r[2].split("<")[0] == anchor]
    if len(idx) < rotations:
        names = sorted({r[2] for r in recs})[:20]
        raise SystemExit(f"only {len(idx)} {anchor} launches in the trace, window has "
                         f"{rotations}; kernel names seen: {names}")
    first = idx[len(idx) - rotations]
    prev_end = recs[idx[len(idx) - rotations - 1]][1] if len(idx) > rotations else -1
    win = [r for r in recs[:first] if r[0] > prev_end] + recs[first:]
    stats = {}
    for s, e, name in win:
        d = (e - s) * 1e-3  # us
        st = stats.setdefault(name, [0, 0.0, float("inf"), 0.0])
        st[0] += 1
        st[1] += d
        st[2] = min(st[2], d)
        st[3] = max(st[3], d)
    total = sum(v[1] for v in stats.values())
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |",
           "|---|---|---|---|---|---|---|"]
    for name, (n, tot, lo, hi) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| {name} | {n} | {tot * 1e-3:.1f} | {tot / n:.1f} | {lo:.1f} | {hi:.1f} | "
                   f"{100 * tot / total:.2f} |")
    span = (win[-1][1] - win[0][0]) * 1e-6
    out.append("")
    out.append(f"Window: {len(win)} dispatches, {rotations} rotations over {steps} steps; "
               f"device time {total * 1e-3:.1f} ms ({total * 1e-3 / steps:.2f} ms per step), "
               f"first-to-last dispatch span {span:.1f} ms ({span / steps:.2f} ms per step).")
    return "\n".join(out)


def summarize_pmc(path, rotations, copy_path=None, kernels=("k_rot_write", "k_rot_count")):
    """Markdown table of rocprofv3 --pmc runs (one counter per run): the per-dispatch average of
    each counter over the last `rotations` launches of each kernel (the timed window), and, when
    `copy_path` holds the same counters for a plain device copy, the copy's average."""
    import csv

    import functools

    @functools.lru_cache(maxsize=None)
    def dm(name):
        if name.startswith("_Z"):
            return _demangle([name]).get(name, name)
        return name

    def load(p):
        vals = {}  # (counter, kernel) -> [(dispatch, value)]
        for f in _rocprof_csvs(p, "counter_collection.csv"):
            with open(f, newline="") as fh:
                for r in csv.DictReader(fh):
                    key = (r["Counter_Name"], _short_name(dm(r["Kernel_Name"])))
                    vals.setdefault(key, []).append((int(r["Dispatch_Id"]),
                                                     float(r["Counter_Value"])))
        return vals

    out = ["| kernel | counter | dispatches | average (KiB per dispatch) |", "|---|---|---|---|"]
    for p in path:
        for (cn, kn), v in sorted(load(p).items()):
            if kn.split("<")[0] not in kernels:
                continue
            v = sorted(v)[-rotations:]
            out.append(f"| {kn} | {cn} | {len(v)} | {sum(x for _, x in v) / len(v):.1f} |")
    for p in copy_path or []:
        for (cn, kn), v in sorted(load(p).items()):
            out.append(f"| {kn} (copy) | {cn} | {len(v)} | {sum(x for _, x in v) / len(v):.1f} |")
    return "\n".join(out)


# The headline FID's operating point: Strang step 600 of the N=16 J1-J2 FID, where the map holds
# ~1.43e7 terms (the survey's windowing: the CPU path cannot run the whole 1000-step FID).
HEADLINE_START_STEP = 600
RESIDUAL_START_STEP = 150  # the bound of step 151


def bench_headline(args):
    """The N=16 J1-J2 FID at its operating point.  Untimed: steps 1..start (the map grows from 16
    to ~1.4e7 terms), then --warmup further untimed steps there.  Timed: the next --steps steps
    on the resident map.  The CPU reference path then runs from the same state (the map at the
    start of the timed region) on all of this job's cores, and on one core."""
    import numpy as np

    from ddevol_amd import engine as E
    from ddevol_amd import hamiltonians as H
    from ddevol_amd import native as N

    n, dt, thr = 16, 0.005, 1e-6
    terms = H.j1j2_n16_terms(n)
    ctx = E.Context(0)
    log(f"device: {ctx.name}")
    mp = E.Map(n)
    for s in range(n):
        mp.add_single(s, 1, 1.0)
    opts = E.options(threshold=thr, context=ctx)
    led = E.ledger(0)
    start = args.start_step if args.start_step is not None else HEADLINE_START_STEP
    grow = start + args.warmup  # untimed steps before the timed region

    t = time.perf_counter()
    if args.load_state:
        st = np.load(args.load_state)
        keys0, vals0 = st["keys"], st["values"]
        start = done = int(st["warmup"])
        mp = E.Map.from_sorted(n, keys0, vals0)
        grow = done + args.warmup
        log(f"loaded state after {done} steps: {len(mp)} terms")
    else:
        done = 0
    if done < grow:
        ev = E.Evolution(mp, terms, dt, opts)
        while done < grow:
            chunk = min(50, grow - done)
            st, fid, rad, act, acc = ev.run(chunk, led)
            if st:
                raise RuntimeError(f"warmup failed at step {done + acc}: status {st}: "
                                   f"{N.engine().tpgp_hip_last_error().decode()}")
            done += acc
            log(f"untimed step {done}: {int(act[-1])} terms, FID {fid[-1]:.9f}, "
                f"{time.perf_counter() - t:.1f} s")
        assert ev.close(mp) == 0
    keys0, vals0 = mp.arrays()
    grow_s = time.perf_counter() - t
    if args.save_state:
        np.savez(args.save_state, keys=keys0, values=vals0, warmup=done)
        log(f"saved {len(mp)} terms to {args.save_state}")
        return {"saved": args.save_state, "terms": len(mp), "warmup": done}
    start_terms = len(mp)

    # timed region: --steps steps on the resident map (slots sized for the window's growth, so
    # no re-layout falls inside it)
    ev = E.Evolution(mp, terms, dt, opts, initial_capacity=4 * len(mp) + (1 << 20))
    ev.run(0, led)  # one readout: allocation and first-touch outside the timed region
    leaf = [N.PROFILE_MERGE, N.PROFILE_SCATTER, N.PROFILE_COUNT, N.PROFILE_SCAN,
            N.PROFILE_PLAN, N.PROFILE_HISTOGRAM, N.PROFILE_SPLITS, N.PROFILE_SMALL]

    def progress(k, terms_now, secs):  # a whole FID: a line per call of 25 steps
        if args.steps > 100:
            log(f"timed step {done + k}: {terms_now} terms, {secs:.1f} s")

    w = timed_window(ev, led, args.steps, leaf, args.sample_every, progress)
    ev.close()
    elapsed = w["elapsed"]
    updates, in_out, rotations = w["updates"], w["in_out"], w["rotations"]
    term_bytes = 8 * ((n + 31) // 32) + 8
    value = updates / elapsed
    log(f"timed {args.steps} steps ({done + 1}..{done + args.steps}): {elapsed:.3f} s, "
        f"{rotations} rotations, {updates} updates, terms {start_terms} -> {w['terms_last']}")
    for name, (ms, cnt) in w["prof"].items():
        log(f"  {name:16s} {cnt:7d} launches {ms:9.1f} ms total "
            f"{ms / max(cnt, 1) * 1e3:8.1f} us avg")
    roof = roofline_of(w["prof"], term_bytes, w["s_updates"], w["s_in_out"], w["sample"],
                       w["s_in_place"])
    roof["traffic"] = (args.traffic if args.traffic is not None
                       else MEASURED_TRAFFIC.get((roof["kernel"], done, args.steps)))

    cpu = None
    if not args.no_cpu:
        threads, avail = host_threads()
        model = cpu_model()
        try:
            kflat = np.ascontiguousarray(keys0.reshape(-1))
            vflat = np.ascontiguousarray(vals0)
            cpu = cpu_baseline(kflat, vflat, n, terms, dt, thr, args.cpu_seconds, threads)
            one = cpu_baseline(kflat, vflat, n, terms, dt, thr, args.cpu_seconds, 1)
            cpu.update({
                "cpu_model": model, "cores_in_affinity": avail,
                "single_thread": {"value": one["value"], "rotations": one["rotations"],
                                  "seconds": one["seconds"]},
                "gpu_over_cpu": value / cpu["value"],
                "sample": f"{cpu['rotations']} rotations of the J1-J2 Strang sequence (from "
                          f"its first term, continuing the evolution) on the {start_terms}-term "
                          f"map at the start of the timed region, {cpu['seconds']:.1f} s, "
                          f"OpenMP branch generation on {threads} threads ({model}; all cores "
                          f"of this job's CPU share, {avail} in its affinity mask); one thread: "
                          f"{one['rotations']} rotations in {one['seconds']:.1f} s"})
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # reported, not fatal
            log(f"cpu baseline unavailable: {e}")

    return {
        "metric": "FP64 Pauli branch-updates/s (N=16 J1-J2); batched FIDs/s + p99 latency vs plain",
        "value": value,
        "unit": "branch-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded couplings; initial state sum_s X_s)",
        "config": {"workload": "N=16 J1-J2 ring FID, J2 from InputPCG(7), dt=0.005, "
                               "threshold 1e-6, binary64",
                   "operating_point": f"untimed steps 1..{start} grow the map to the window's "
                                      f"size; --warmup {args.warmup} further untimed steps "
                                      f"there",
                   "untimed_steps": done, "untimed_seconds": grow_s,
                   "timed_steps": f"{done + 1}..{done + args.steps}",
                   "terms_at_start": start_terms, "terms_at_end": w["terms_last"],
                   "rotations": rotations, "branch_updates": updates, "terms_in_out": in_out,
                   "terms_left_in_place": w["in_place"] // 2,
                   "fid_at_end": w["fid_last"],
                   # a loaded state carries no ledger of its earlier steps
                   "fid_allowance_at_end": None if args.load_state else
                   fid_allowance(E, terms, dt, led, n, done + args.steps)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }


if __name__ == "__main__":
    main()

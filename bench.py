#!/usr/bin/env python3
"""Benchmark: FP64 Pauli branch-updates/s on the N=16 J1-J2 FID, one MI355X.

Workload (the metric's configuration): N=16 periodic J1-J2 ring, J1 = 1, J2_i = 0.1 + 0.3 u_i
with u_i the first 16 uniforms of the reference's seeded InputPCG(7); secular dipolar terms
(96 per step order); observable sum_s X_s with coefficients 1; Strang steps dt = 0.005;
truncation |c| <= 1e-6; binary64 throughout.  Synthetic by construction: no dataset.

A "step" is one Strang step: 192 rotations over the resident map.  W untimed warmup steps grow
the map from 16 terms (it reaches ~6e6 terms by step 500 and ~4e7 by step 750); the timed
region is the next K steps, bracketed by device synchronisation, with the map already resident
in HBM.  One branch-update = one input term processed by one executed rotation (the
reference's ledger.rotations counts the same rotations, engine/src/tpgp_rev.c:740).

roofline: the dominant kernel is whichever of the rotation count pass (k_rot_count) and the
rotation write pass (k_rot_write) takes more time in the timed region.  Algorithmic bytes per
launch, for W key words (16 B per term for N <= 32): k_rot_write (8W + 8)(len_in + len_out) --
the input map read once and the output map written once; k_rot_count (8W + 8) len_in.  Their
durations are measured live with HIP events recorded on the batch stream around every launch
in the timed region (only those two kinds, unless --breakdown asks for all).

cpu_baseline: the reference CPU engine (engine/src/tpgp_rev.c compiled from its own sources into
oracle/_ref) -- or, where that build is absent, the oracle's C restatement -- runs the same
rotation sequence from the same map state (the state at the start of the timed region) with its
OpenMP branch backend on the host cores, for a bounded ~20 s sample.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(keys, values, n_spins, terms, dt, threshold, budget_s, threads):
    """Time the reference (or oracle) CPU path on the same state; returns a dict."""
    import numpy as np

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
    for (i, j, qi, qj, h) in seq:
        if time.perf_counter() - t0 > budget_s:
            break
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
            "sample": f"{rots} rotations (the J1-J2 Strang sequence from its first term) on the "
                      f"{len(values)}-term map at the start of the timed region, "
                      f"{el:.1f} s, OpenMP branch generation on {threads} host threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=600)
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one FID per GPU: only --gpus 1 is supported; "
                                   "multi-GPU sharding is a design, not part of this build"}))
        sys.exit(2)

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

    if args.load_state:
        st = np.load(args.load_state)
        keys0, vals0 = st["keys"], st["values"]
        args.warmup = int(st["warmup"])
        mp = E.Map.from_sorted(n, keys0, vals0)
        log(f"loaded state after {args.warmup} steps: {len(mp)} terms")
    else:
        # warmup: grow the map to the timed region's starting state (untimed)
        t = time.perf_counter()
        ev = E.Evolution(mp, terms, dt, opts)
        done = 0
        while done < args.warmup:
            chunk = min(50, args.warmup - done)
            st, fid, rad, act, acc = ev.run(chunk, led)
            if st:
                raise RuntimeError(f"warmup failed at step {done + acc}: status {st}: "
                                   f"{N.engine().tpgp_hip_last_error().decode()}")
            done += acc
            log(f"warmup step {done}: {int(act[-1])} terms, FID {fid[-1]:.9f}, "
                f"{time.perf_counter() - t:.1f} s")
        st = ev.close(mp)
        assert st == 0
        keys0, vals0 = mp.arrays()
        if args.save_state:
            np.savez(args.save_state, keys=keys0, values=vals0, warmup=args.warmup)
            log(f"saved {len(mp)} terms to {args.save_state}")
            return
    start_terms = len(mp)

    # timed region: K steps on the resident map
    ev = E.Evolution(mp, terms, dt, opts, initial_capacity=4 * len(mp) + (1 << 20))
    ev.run(0, led)  # one readout: allocation and first-touch outside the timed region
    timed_kinds = [N.PROFILE_MERGE, N.PROFILE_COUNT]
    mask = N.PROFILE_ALL if args.breakdown else sum(1 << k for k in timed_kinds)
    ev.profile(mask)
    c0 = ev.counters()
    t0 = time.perf_counter()
    st, fid, rad, act, acc = ev.run(args.steps, led)
    t1 = time.perf_counter()
    if st:
        raise RuntimeError(f"timed run failed: status {st}: "
                           f"{N.engine().tpgp_hip_last_error().decode()}")
    c1 = ev.counters()
    kinds = range(len(N.PROFILE_NAMES)) if args.breakdown else timed_kinds
    prof = {N.PROFILE_NAMES[k]: ev.profile_read(k) for k in kinds}
    ev.close()
    elapsed = t1 - t0
    updates = c1.updates - c0.updates
    in_out = c1.in_out - c0.in_out
    rotations = c1.rotations - c0.rotations
    key_bytes = 8 * ((n + 31) // 32)
    term_bytes = key_bytes + 8
    value = updates / elapsed
    # dominant kernel: the timed kind with the largest total time.  Algorithmic bytes per
    # launch: k_rot_write reads every input term once and writes every output term once;
    # k_rot_count reads every input term once.
    alg_bytes = {"k_rot_write": term_bytes * in_out, "k_rot_count": term_bytes * updates}
    dom = max(alg_bytes, key=lambda k: prof[k][0])
    dom_ms, dom_n = prof[dom]
    dom_bytes = alg_bytes[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else None
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {rotations} rotations, {updates} updates, "
        f"terms {start_terms} -> {int(act[-1])}")
    for name, (ms, cnt) in prof.items():
        log(f"  {name:16s} {cnt:7d} launches {ms:9.1f} ms total {ms / max(cnt, 1) * 1e3:8.1f} us avg")

    cpu = None
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        try:
            cpu = cpu_baseline(np.ascontiguousarray(keys0.reshape(-1)), np.ascontiguousarray(vals0),
                               n, terms, dt, thr, args.cpu_seconds, threads)
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # reported, not fatal
            log(f"cpu baseline unavailable: {e}")

    out = {
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
                   "timed_steps": f"{args.warmup + 1}..{args.warmup + args.steps}",
                   "terms_at_start": start_terms, "terms_at_end": int(act[-1]),
                   "rotations": rotations, "branch_updates": updates,
                   "fid_at_end": float(fid[-1])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": args.traffic,
                     "kernel": dom, "launches": dom_n,
                     "avg_us": dom_ms / max(dom_n, 1) * 1e3,
                     "algorithmic_bytes_per_launch": dom_bytes / max(dom_n, 1),
                     "kernels": {k: {"launches": n_, "avg_us": ms / max(n_, 1) * 1e3}
                                 for k, (ms, n_) in prof.items()}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

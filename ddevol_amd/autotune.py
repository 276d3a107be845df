"""Launch-shape autotuner: picks the partition's elements per thread (its tile size) and the
grids of the rotation's count and write passes per key width and size bucket, by timing the
candidates on the device, and writes the table that contexts load (TPGP_HIP_TUNING,
include/tpgp_hip.h).

Every candidate gives the same keys, values and ledgers (the parameters only change how a pass
is cut into workgroups), so the tuner compares time alone.  The workloads are FID trajectories
from the X-sum initial state: one-word keys on the N=16 J1-J2 ring (dt 0.005, threshold 1e-6),
two-word keys on the N=64 range-4 ring (dt 0.01, threshold 1e-5).  Each map is snapshotted as it
first reaches 2**k terms, and from each snapshot every candidate runs the same few Strang steps
on a fresh session, timed with the HIP events of the pass the parameter governs (partition
pass, count pass, write pass).  One parameter at a time, the others at the values already
chosen for that bucket; ties within 2 % keep the built-in default.

    python -m ddevol_amd.autotune --out ddevol_amd/tuning_gfx950.txt
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from . import engine as E
from . import hamiltonians as H
from . import native as N

CANDIDATES = {
    N.TUNE_PARTITION_ROUNDS: [8, 4, 12],  # the default first
    N.TUNE_COUNT_GRID: [3, 2, 6],
    N.TUNE_WRITE_GRID: [4, 3, 8],
}
KIND = {
    N.TUNE_PARTITION_ROUNDS: N.PROFILE_PARTITION,
    N.TUNE_COUNT_GRID: N.PROFILE_COUNT,
    N.TUNE_WRITE_GRID: N.PROFILE_MERGE,
}
DEFAULT = {p: c[0] for p, c in CANDIDATES.items()}
TIE = 0.02


def log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def workload(words):
    """(spins, terms, dt, threshold) of the trajectory that tunes keys of `words` words."""
    if words == 1:
        return 16, H.j1j2_n16_terms(16), 0.005, 1e-6
    return 64, H.range_ring(64, 4), 0.01, 1e-5


def snapshots(ctx, words, buckets, max_steps=2000):
    """Maps of the trajectory as it first reaches 2**k terms, k in buckets."""
    n, terms, dt, thr = workload(words)
    mp = E.Map(n)
    for s in range(n):
        mp.add_single(s, 1, 1.0)
    opts = E.options(threshold=thr, max_terms=10 ** 11, context=ctx)
    led = E.ledger(0)
    want = sorted(buckets)
    out = {}
    ev = E.Evolution(mp, terms, dt, opts, context=ctx)
    step = 0
    while want and step < max_steps:
        st, fid, rad, act, acc = ev.run(1, led)
        if st:
            raise RuntimeError(f"trajectory failed at step {step}: status {st}")
        step += 1
        if int(act[-1]) >= 1 << want[0]:
            assert ev.close(mp) == 0
            keys, vals = mp.arrays()
            while want and len(vals) >= 1 << want[0]:
                out[want.pop(0)] = (keys.copy(), vals.copy(), step)
            log(f"step {step}: {len(vals)} terms, snapshot")
            ev = E.Evolution(mp, terms, dt, opts, context=ctx)
    ev.close()
    return out


def time_trial(ctx, words, snap, steps, kind):
    """Milliseconds of pass `kind` over `steps` Strang steps from the snapshot."""
    n, terms, dt, thr = workload(words)
    keys, vals, _ = snap
    mp = E.Map.from_sorted(n, keys, vals)
    opts = E.options(threshold=thr, max_terms=10 ** 11, context=ctx)
    led = E.ledger(0)
    ev = E.Evolution(mp, terms, dt, opts, initial_capacity=4 * len(vals) + (1 << 16),
                     context=ctx)
    ev.run(0, led)  # allocation and first touch outside the timing
    ev.profile(1 << kind)
    st, fid, rad, act, acc = ev.run(steps, led)
    ms, launches = ev.profile_read(kind)
    ev.close()
    if st:
        raise RuntimeError(f"trial failed: status {st}")
    return ms, launches


def set_all(ctx, words, param, value):
    for b in range(N.TUNE_BUCKETS):
        assert ctx.set_tuning(words, param, b, value) == 0


def trial_steps(words, k):
    """Strang steps per trial: enough launches to time at 2**k terms (one J1-J2 step is 192
    rotations, one N=64 ring step 1536)."""
    if words == 1:
        return int(min(20, max(2, (1 << 21) // (1 << k))))
    return int(min(4, max(1, (1 << 19) // (1 << k))))


def tune(ctx, words, buckets=range(14, 25), reps=2):
    """{bucket: {param: value}} for keys of `words` words, from timings on this device."""
    t0 = time.perf_counter()
    snaps = snapshots(ctx, words, buckets)
    table = {}
    for k in sorted(snaps):
        steps = trial_steps(words, k)
        chosen = dict(DEFAULT)
        for param, cands in CANDIDATES.items():
            for p, v in chosen.items():
                set_all(ctx, words, p, v)
            times = {}
            for v in cands:
                set_all(ctx, words, param, v)
                best = None
                for _ in range(reps):
                    ms, launches = time_trial(ctx, words, snaps[k], steps, KIND[param])
                    best = ms if best is None else min(best, ms)
                times[v] = best
            base = times[cands[0]]
            win = min(times, key=times.get)
            if base <= 0 or times[win] > base * (1 - TIE):
                win = cands[0]
            chosen[param] = win
            log(f"W={words} bucket {k} ({len(snaps[k][1])} terms, {steps} steps) "
                f"{N.TUNE_NAMES[param]}: "
                + ", ".join(f"{v}: {t:.2f} ms" for v, t in times.items())
                + f" -> {win}  [{time.perf_counter() - t0:.0f} s]")
        table[k] = chosen
    for p in CANDIDATES:
        set_all(ctx, words, p, 0)
    return table


def write_table(path, tables, device):
    """Lines "words param bucket value" for the non-default choices ({words: {bucket: {param:
    value}}}); buckets below a width's smallest tuned one take its values, those above its
    largest take the largest one's."""
    lines = [f"# launch-shape table measured by ddevol_amd/autotune.py on {device}",
             "# workloads: 1 key word: N=16 J1-J2 FID trajectory (dt 0.005, threshold 1e-6);",
             "#            2 key words: N=64 range-4 ring FID trajectory (dt 0.01, threshold 1e-5)",
             "# words param bucket value  (params: 0 partition rounds, 1 count grid, 2 write grid;",
             "#  bucket = floor(log2(terms in the pass)); unlisted = built-in default)"]
    for words in sorted(tables):
        ks = sorted(tables[words])
        if not ks:
            continue
        for b in range(N.TUNE_BUCKETS):
            src = min(ks, key=lambda k: abs(k - b))
            for p, v in sorted(tables[words][src].items()):
                if v != DEFAULT[p]:
                    lines.append(f"{words} {p} {b} {v}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(N.HERE, "tuning_gfx950.txt"))
    ap.add_argument("--min-bucket", type=int, default=14)
    ap.add_argument("--max-bucket", type=int, default=24)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--words", type=int, nargs="+", default=[1, 2])
    args = ap.parse_args(argv)
    os.environ["TPGP_HIP_TUNING"] = ""  # measure from the built-in defaults
    ctx = E.Context(0)
    log(f"device: {ctx.name}")
    tables = {w: tune(ctx, w, range(args.min_bucket, args.max_bucket + 1), args.reps)
              for w in args.words}
    write_table(args.out, tables, ctx.name)
    log(f"wrote {args.out}")
    ctx.close()


if __name__ == "__main__":
    main()

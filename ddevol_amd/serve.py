"""FID serving with continuous batching over the fleet API (tpgp_fleet_*).

Requests (one FID each: ring size N, couplings, threshold, step count) arrive over time.  The
server keeps every admitted request's term arrays resident in HBM as one slot of a fleet and
advances all running slots of a fleet together, one Strang step per device round, so requests
that arrive while others run join the next round as soon as a slot and memory are free.  Keys
of N <= 32 spins use one 64-bit word and N > 32 two, and requests projected to grow large run
apart from the small ones, so the server runs one fleet per key width and size class, sharing
the GPU between them by virtual time (Server).

Admission is by projected term-array bytes against a fixed HBM budget.  A request's projection
is the peak term count its FID is expected to reach (`projected_peak_terms`: measured growth of
the serving rings per threshold, capped by the Pauli space the dynamics can reach), times three
term buffers, 8W + 8 bytes a term and the slots' growth headroom.  A request whose projection
alone (beside every other slot at its opening share) exceeds the budget is refused at once; the
others wait until a slot is free and the bytes the fleets hold (every slot's share, free slots
included) plus the reservations of the running requests leave room.  Free slots that grew for
an earlier request are trimmed back to the opening share (tpgp_fleet_trim) when that is what
keeps a request waiting.  Each round, a running request's reservation rises to the bytes its
slot actually holds, so growth beyond the projection is charged before anything else is
admitted, and each fleet's own ceiling (tpgp_fleet_set_budget) is what the other fleets leave
of the budget: a slot that cannot grow within it fails alone (TPGP_ENOMEM) and the others run
on.  Each request's results are exactly those of running it alone (the fleet runs each slot's
own loop).
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import engine as E
from . import hamiltonians as H


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the stream
    n: int
    terms: list
    threshold: float
    steps: int
    dt: float = 0.01
    # filled by the server
    admitted: float = math.nan
    finished: float = math.nan
    status: int = -1
    fid: np.ndarray | None = None
    active: np.ndarray | None = None

    @property
    def words(self):
        return 1 if self.n <= 32 else 2

    @property
    def latency(self):
        return self.finished - self.arrival


def poisson_stream(count, rate, seed=7, n_range=(8, 48), steps_range=(100, 2000),
                   thresholds=(1e-4, 1e-5, 1e-6), dt=0.01):
    """Seeded request stream: exponential inter-arrival times (rate per second), N uniform in
    n_range, ring NN couplings uniform in [0.5, 1.5], NNN coupling 1/2^3, threshold uniform over
    `thresholds`, steps uniform in steps_range (all draws from InputPCG(seed))."""
    g = H.InputPCG(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += -math.log(1.0 - g.uniform()) / rate
        n = n_range[0] + int(g.uniform() * (n_range[1] - n_range[0] + 1))
        nn = [0.5 + g.uniform() for _ in range(n)]
        thr = thresholds[int(g.uniform() * len(thresholds))]
        steps = steps_range[0] + int(g.uniform() * (steps_range[1] - steps_range[0] + 1))
        out.append(Request(k, t, n, H.serving_ring(n, nn), thr, steps, dt))
    return out


# Peak term counts of the serving rings (N = 48, nearest-neighbour couplings U[0.5, 1.5],
# next-nearest 1/8, dt 0.01, initial state sum_s X_s) after a given number of Strang steps, per
# truncation threshold: the engine's own trajectories on MI355X, sampled every 50 steps.  Larger
# rings hold more terms at a given step, so the N = 48 curve bounds the smaller rings.
GROWTH = {
    1e-4: [(50, 2509), (100, 8097), (150, 21011), (200, 43799), (300, 101616), (400, 145440),
           (500, 175264), (700, 213004), (1000, 248964), (1500, 276011), (2000, 280936)],
    1e-5: [(50, 11109), (100, 47087), (150, 159216), (200, 444809), (250, 1050926),
           (300, 2104761), (400, 5451302), (500, 9266370), (600, 12505092), (800, 17333826),
           (950, 20083609), (1000, 20877675)],
    1e-6: [(50, 50549), (100, 323732), (150, 1484618), (200, 5539945), (250, 17242054),
           (300, 44947444), (350, 99233726), (400, 188122001)],
}
SLOT_HEADROOM = 3.0  # a fleet slot whose map fills more than half of it grows to 3x the map


def projected_peak_terms(n, threshold, steps):
    """Term count a ring FID of n spins is expected to reach within `steps` Strang steps at this
    truncation threshold: the measured curve of the nearest threshold at or below it (smaller
    thresholds keep more terms), interpolated geometrically, and past its last sample extended
    at its last growth rate (no saturation assumed); at most 4^n / 4, the part of the Pauli space
    the dynamics reaches from sum_s X_s."""
    keys = sorted(GROWTH)
    thr = max([t for t in keys if t <= threshold * (1 + 1e-9)] or [keys[0]])
    curve = GROWTH[thr]
    if steps <= curve[0][0]:
        peak = curve[0][1] * steps / curve[0][0]
    elif steps >= curve[-1][0]:
        (s0, t0), (s1, t1) = curve[-2], curve[-1]
        peak = t1 * (t1 / t0) ** ((steps - s1) / (s1 - s0))
    else:
        k = next(i for i, (st, _) in enumerate(curve) if st >= steps)
        (s0, t0), (s1, t1) = curve[k - 1], curve[k]
        peak = t0 * (t1 / t0) ** ((steps - s0) / (s1 - s0))
    return int(min(peak, 4.0 ** n / 4)) + n


def initial_state(n):
    m = E.Map(n)
    for s in range(n):
        if m.add_single(s, 1, 1.0):
            raise RuntimeError("initial state")
    return m


LARGE_TERMS = 1 << 20  # projected peak terms from which a request runs in a large-map fleet


class Server:
    """Fleets by key width (one word: N <= 32, two: N > 32) and size class: requests projected
    to stay below LARGE_TERMS terms share a small-map fleet, the others a large-map fleet, so a
    round of small maps never waits for a step of a large one.  The fleets share the GPU by
    virtual time: the fleet stepped next is the one whose rounds have taken the least device
    time per running request so far (a fleet that becomes busy starts at the least virtual time
    of the busy ones), so every running request gets about the same share of the GPU."""

    def __init__(self, context: E.Context, slots_per_width=64, budget_bytes=64 << 30,
                 slot_terms=1 << 15, large_slots=8):
        self.ctx = context
        self.budget = budget_bytes
        self.slot_terms = slot_terms
        self.keys = [(w, c) for w in (1, 2) for c in ("small", "large")]
        self.slots = {k: slots_per_width if k[1] == "small" else large_slots for k in self.keys}
        self.fleets = {k: E.Fleet(k[0], self.slots[k], slot_terms, context=context)
                       for k in self.keys}
        for fl in self.fleets.values():
            fl.set_budget(budget_bytes)
        # what every slot holds at the least: the opening share, free or not
        self.share = {k: 3 * (8 * k[0] + 8) * slot_terms for k in self.keys}
        self.floor = sum(self.slots[k] * self.share[k] for k in self.keys)
        self.slot_req = {k: [None] * self.slots[k] for k in self.keys}
        self.reserved = {k: [0] * self.slots[k] for k in self.keys}
        self.vtime = {k: 0.0 for k in self.keys}
        self.refused = []
        self.rounds = 0
        self.peak_held = 0

    def projected_bytes(self, req: Request):
        terms = max(projected_peak_terms(req.n, req.threshold, req.steps), self.slot_terms)
        return int(3 * (8 * req.words + 8) * SLOT_HEADROOM * terms)

    def fleet_of(self, req: Request):
        big = projected_peak_terms(req.n, req.threshold, req.steps) >= LARGE_TERMS
        return req.words, "large" if big else "small"

    def running(self, k=None):
        keys = [k] if k else self.keys
        return sum(r is not None for f in keys for r in self.slot_req[f])

    def _slot_charge(self, k, s):
        """Bytes slot s of fleet k counts against the budget: what it holds, or what its running
        request is reserved, whichever is more."""
        return max(self.fleets[k].slot_bytes(s), self.reserved[k][s])

    def held_bytes(self):
        return sum(self._slot_charge(k, s) for k in self.keys for s in range(self.slots[k]))

    def fits(self, req: Request):
        """Whether `req` fits the budget at all: with every other slot at its opening share."""
        k = self.fleet_of(req)
        return self.floor - self.share[k] + max(self.projected_bytes(req), self.share[k]) \
            <= self.budget

    def _trim(self):
        """Give the memory of grown free slots back (tpgp_fleet_trim); True if any was."""
        grown = False
        for k, fl in self.fleets.items():
            if any(r is None and fl.slot_bytes(s) > self.share[k]
                   for s, r in enumerate(self.slot_req[k])):
                st = fl.trim()
                if st:
                    raise RuntimeError(f"fleet trim failed: status {st}")
                grown = True
        return grown

    def _admit(self, req: Request, now):
        """Admit `req` if a slot and the budget allow it now; True when admitted."""
        k = self.fleet_of(req)
        fl = self.fleets[k]
        need = self.projected_bytes(req)
        free = [s for s, r in enumerate(self.slot_req[k]) if r is None]
        if not free:
            return False

        def place():  # the free slot whose share covers most of the reservation, and the bytes
            j = min(free, key=lambda j: max(need, fl.slot_bytes(j)) - fl.slot_bytes(j))
            held = self.held_bytes() - self._slot_charge(k, j)
            return j, held + max(need, fl.slot_bytes(j))

        s, total = place()
        if total > self.budget and self._trim():
            s, total = place()
        if total > self.budget:
            return False
        st = fl.admit(s, initial_state(req.n), req.terms, req.dt, req.steps,
                      E.options(threshold=req.threshold, max_terms=10 ** 10))
        if st:
            raise RuntimeError(f"admission failed: status {st}")
        if not self.running(k):  # a fleet that becomes busy gets no credit for its idle time
            busy = [self.vtime[f] for f in self.keys if self.running(f)]
            self.vtime[k] = max(self.vtime[k], min(busy) if busy else 0.0)
        self.slot_req[k][s] = req
        self.reserved[k][s] = need
        req.admitted = now
        return True

    def _round(self, k, clock):
        """One step of every running slot of fleet k."""
        fl = self.fleets[k]
        # the device bytes this fleet's slots may grow to: what the others hold leaves the rest
        fl.set_budget(max(1, self.budget - sum(
            o.held_bytes() for f, o in self.fleets.items() if f != k)))
        n = self.running(k)
        t = clock()
        st, _ = fl.step()
        if st:
            raise RuntimeError(f"fleet step failed: status {st}: "
                               f"{E.N.engine().tpgp_hip_last_error().decode()}")
        self.vtime[k] += (clock() - t) / n
        self.rounds += 1
        self.peak_held = max(self.peak_held, sum(f.held_bytes() for f in self.fleets.values()))

    def run(self, requests, clock=time.perf_counter, log=None):
        """Serve the stream in real time; returns the requests with results filled in.  Refused
        requests (projection beyond the budget) keep status -1 and land in self.refused.  Ready
        requests are admitted first fit in arrival order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        ready = []
        t0 = clock()
        done = 0
        last_log = 0.0
        while done + len(self.refused) < len(requests):
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.pop(0)
                if not self.fits(r):
                    self.refused.append(r)
                    if log:
                        log(f"request {r.rid} (N={r.n}, {r.steps} steps, thr {r.threshold:g}) "
                            f"refused: projected {self.projected_bytes(r) / 2 ** 30:.1f} GiB")
                else:
                    ready.append(r)
            ready = [r for r in ready if not self._admit(r, now)]
            busy = [k for k in self.keys if self.running(k)]
            if not busy:
                if pending:  # idle until the next arrival
                    time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                    continue
                if ready:
                    raise RuntimeError(
                        f"request {ready[0].rid} fits the budget ({self.budget} bytes) but "
                        f"cannot be admitted: {self.held_bytes()} held, "
                        f"{self.projected_bytes(ready[0])} projected")
                continue
            k = min(busy, key=lambda f: self.vtime[f])
            self._round(k, clock)
            now = clock() - t0
            if log and now - last_log >= 20.0:
                last_log = now
                log(f"{now:.1f} s: round {self.rounds}, {done} done, {self.running()} running, "
                    f"{len(ready)} waiting, {self.held_bytes() / 2 ** 30:.2f} GiB charged")
            fl = self.fleets[k]
            for s, r in enumerate(self.slot_req[k]):
                if r is None:
                    continue
                state, steps_done, status, _ = fl.slot(s)
                if state in (E.N.SLOT_DONE, E.N.SLOT_FAILED):
                    r.fid, _, r.active = fl.fid(s, r.steps + 1)
                    r.status = status
                    r.finished = now
                    fl.release(s)
                    self.slot_req[k][s] = None
                    self.reserved[k][s] = 0
                    done += 1
                    if log:
                        log(f"request {r.rid} (N={r.n}, {r.steps} steps, thr {r.threshold:g}) "
                            f"done at {now:.2f} s, latency {r.latency:.2f} s, status {status}")
                else:  # growth beyond the projection is charged from now on
                    self.reserved[k][s] = max(self.reserved[k][s], fl.slot_bytes(s))
        return requests

    def close(self):
        for fl in self.fleets.values():
            fl.close()


def run_alone(context: E.Context, requests, clock=time.perf_counter, log=None, chunk=100):
    """The one-FID-at-a-time loop: each request runs by itself (a resident evolution session,
    tpgp_evolution_*) in arrival order, starting no earlier than its arrival; returns the
    requests with results filled in.  The session runs `chunk` steps per call (the samples are
    those of one call: each call's first sample is the previous call's last)."""
    t0 = clock()
    for r in sorted(requests, key=lambda q: q.arrival):
        now = clock() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        r.admitted = clock() - t0
        o = E.options(threshold=r.threshold, max_terms=10 ** 10, context=context)
        led = E.ledger(0)
        ev = E.Evolution(initial_state(r.n), r.terms, r.dt, o, context=context)
        fids, acts, done, st = [], [], 0, 0
        while done < r.steps and not st:
            st, fid, _, act, acc = ev.run(min(chunk, r.steps - done), led)
            fids.append(fid[:acc + 1] if not fids else fid[1:acc + 1])
            acts.append(act[:acc + 1] if not acts else act[1:acc + 1])
            done += acc
            if log and not st and done < r.steps:
                log(f"alone: request {r.rid}: {done} of {r.steps} steps, {int(act[acc])} terms")
        ev.close()
        r.status = st
        r.fid = np.concatenate(fids)
        r.active = np.concatenate(acts)
        r.finished = clock() - t0
        if log:
            log(f"alone: request {r.rid} (N={r.n}, {r.steps} steps, thr {r.threshold:g}) done "
                f"at {r.finished:.2f} s")
    return requests

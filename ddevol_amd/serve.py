"""FID serving with continuous batching over the fleet API (tpgp_fleet_*).

Requests (one FID each: ring size N, couplings, threshold, step count) arrive over time.  The
server keeps every admitted request's term arrays resident in HBM as one slot of a fleet and
advances all running slots together, one Strang step per device round, so requests that arrive
while others run join the next round as soon as a slot and memory are free.  Keys of N <= 32
spins use one 64-bit word and N > 32 two, so the server runs one fleet per key width.

Admission is by projected term-array bytes against a fixed HBM budget: a request needs its slot
(three term buffers of `slot_terms` terms at 8*W + 8 bytes each) before it is admitted, and the
bytes held by running slots (which grow on demand) count against the budget.  Each request's
results are exactly those of running it alone (the fleet runs each slot's own loop).
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


def initial_state(n):
    m = E.Map(n)
    for s in range(n):
        if m.add_single(s, 1, 1.0):
            raise RuntimeError("initial state")
    return m


class Server:
    def __init__(self, context: E.Context, slots_per_width=64, budget_bytes=64 << 30,
                 slot_terms=1 << 15):
        self.ctx = context
        self.budget = budget_bytes
        self.slot_terms = slot_terms
        self.fleets = {w: E.Fleet(w, slots_per_width, slot_terms, context=context)
                       for w in (1, 2)}
        self.slot_req = {w: [None] * slots_per_width for w in (1, 2)}
        self.rounds = 0

    def projected_bytes(self, req: Request):
        return 3 * self.slot_terms * (8 * req.words + 8)

    def held_bytes(self):
        total = 0
        for w, fl in self.fleets.items():
            for s, r in enumerate(self.slot_req[w]):
                if r is not None:
                    total += fl.slot_bytes(s)
        return total

    def _admit(self, req: Request, now):
        w = req.words
        slots = self.slot_req[w]
        try:
            s = slots.index(None)
        except ValueError:
            return False
        if self.held_bytes() + self.projected_bytes(req) > self.budget:
            return False
        st = self.fleets[w].admit(s, initial_state(req.n), req.terms, req.dt, req.steps,
                                  E.options(threshold=req.threshold, max_terms=10 ** 10))
        if st:
            raise RuntimeError(f"admission failed: status {st}")
        slots[s] = req
        req.admitted = now
        return True

    def run(self, requests, clock=time.perf_counter, log=None):
        """Serve the stream in real time; returns the requests with results filled in."""
        pending = sorted(requests, key=lambda r: r.arrival)
        ready = []
        t0 = clock()
        done = 0
        while done < len(requests):
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                ready.append(pending.pop(0))
            while ready and self._admit(ready[0], now):
                ready.pop(0)
            running = any(r is not None for w in (1, 2) for r in self.slot_req[w])
            if not running:
                if pending:  # idle until the next arrival
                    time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                    continue
                if ready:
                    raise RuntimeError("a request does not fit the HBM budget")
            for w, fl in self.fleets.items():
                if any(r is not None for r in self.slot_req[w]):
                    st, _ = fl.step()
                    if st:
                        raise RuntimeError(f"fleet step failed: status {st}: "
                                           f"{E.N.engine().tpgp_hip_last_error().decode()}")
            self.rounds += 1
            now = clock() - t0
            for w, fl in self.fleets.items():
                for s, r in enumerate(self.slot_req[w]):
                    if r is None:
                        continue
                    state, steps_done, status, _ = fl.slot(s)
                    if state in (E.N.SLOT_DONE, E.N.SLOT_FAILED):
                        r.fid, _, r.active = fl.fid(s, r.steps + 1)
                        r.status = status
                        r.finished = now
                        fl.release(s)
                        self.slot_req[w][s] = None
                        done += 1
                        if log:
                            log(f"request {r.rid} (N={r.n}, {r.steps} steps, thr {r.threshold:g})"
                                f" done at {now:.2f} s, latency {r.latency:.2f} s")
        return requests

    def close(self):
        for fl in self.fleets.values():
            fl.close()


def run_alone(context: E.Context, requests, clock=time.perf_counter, log=None):
    """The one-FID-at-a-time loop: each request runs by itself (tpgp_hip_evolve) in arrival
    order, starting no earlier than its arrival; returns the requests with results filled in."""
    t0 = clock()
    for r in sorted(requests, key=lambda q: q.arrival):
        now = clock() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        r.admitted = clock() - t0
        m = initial_state(r.n)
        o = E.options(threshold=r.threshold, max_terms=10 ** 10, context=context)
        led = E.ledger(0)
        st, fid, rad, act, acc = E.evolve(m, r.terms, r.dt, r.steps, o, led)
        r.status = st
        r.fid = fid[:acc + 1]
        r.active = act[:acc + 1]
        r.finished = clock() - t0
        if log:
            log(f"alone: request {r.rid} (N={r.n}, {r.steps} steps, thr {r.threshold:g}) done "
                f"at {r.finished:.2f} s")
    return requests

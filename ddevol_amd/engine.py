"""Python surface of the engine: the reference core's operations, backed by the native libraries.

Mirrors engine/include/tpgp_rev.h (maps, tpgp_rotate, tpgp_strang_step, readout) and
engine/include/tpgp_cuda_rev.h (context + branches) with the same argument meaning and statuses.
"""
from __future__ import annotations

import ctypes as C
import math
import weakref

import numpy as np

from . import native as N


class Context:
    """A device context (tpgp_hip_create / tpgp_hip_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = N.hip()
        self._children = weakref.WeakSet()  # sessions on this context, closed before it
        err = C.create_string_buffer(512)
        self.ptr = self.lib.tpgp_hip_create(device, err, 512)
        if not self.ptr:
            raise RuntimeError(f"HIP device unavailable: {err.value.decode()}")

    @property
    def name(self) -> str:
        return self.lib.tpgp_hip_device_name(self.ptr).decode()

    def set_tuning(self, words: int, param: int, bucket: int, value: int) -> int:
        """tpgp_hip_set_tuning: a launch shape for passes over about 2**bucket terms of keys of
        `words` words (0 resets)."""
        return self.lib.tpgp_hip_set_tuning(self.ptr, words, param, bucket, value)

    def get_tuning(self, words: int, param: int, bucket: int) -> int:
        """The value in force (the built-in default where none is set)."""
        v = C.c_uint()
        st = self.lib.tpgp_hip_get_tuning(self.ptr, words, param, bucket, C.byref(v))
        if st:
            raise ValueError(f"tpgp_hip_get_tuning: status {st}")
        return v.value

    def close(self):
        if self.ptr:
            for child in list(self._children):
                child.close()
            self.lib.tpgp_hip_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def branches(self, keys, values, n_words, n_spins, i, j, qi, qj, c, s, rc, rs, enclose,
                 noalloc=False, fill=0):
        """tpgp_hip_branches (or, with noalloc, tpgp_hip_branches_noalloc) on host arrays;
        returns (status, out_keys, out_values, out_radii, msg).  `fill` is the bit pattern the
        output arrays hold before the call."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1)
        values = np.ascontiguousarray(values, dtype=np.float64)
        count = len(values)
        ok = np.full(2 * count * n_words, fill, np.uint64)
        ov = np.full(2 * count, fill, np.uint64).view(np.float64)
        orad = np.full(2 * count, fill, np.uint64).view(np.float64)
        err = C.create_string_buffer(512)
        fn = self.lib.tpgp_hip_branches_noalloc if noalloc else self.lib.tpgp_hip_branches
        st = fn(self.ptr, N.as_u64p(keys), N.as_dp(values), count, n_words, n_spins, i, j, qi,
                qj, c, s, rc, rs, enclose, N.as_u64p(ok), N.as_dp(ov), N.as_dp(orad), err, 512)
        return st, ok.reshape(2 * count, n_words) if count else ok, ov, orad, err.value.decode()


class Map:
    """A host tpgp_map (sorted unique keys, reference layout)."""

    def __init__(self, n_spins: int):
        self.eng = N.engine()
        self.m = N.TpgpMap()
        st = self.eng.tpgp_map_init(C.byref(self.m), n_spins)
        if st:
            raise ValueError(f"tpgp_map_init: status {st}")

    def __del__(self):
        try:
            self.eng.tpgp_map_free(C.byref(self.m))
        except Exception:
            pass

    @property
    def n_spins(self):
        return self.m.n_spins

    @property
    def n_words(self):
        return self.m.n_words

    def __len__(self):
        return self.m.len

    def add_single(self, site, pauli, value):
        return self.eng.tpgp_map_add_single(C.byref(self.m), site, pauli, value)

    def add(self, key_words, value):
        k = np.ascontiguousarray(key_words, dtype=np.uint64)
        return self.eng.tpgp_map_add(C.byref(self.m), N.as_u64p(k), value)

    @classmethod
    def from_arrays(cls, n_spins, keys, values):
        mp = cls(n_spins)
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(len(values), -1)
        for k in range(len(values)):
            st = mp.add(keys[k], float(values[k]))
            if st:
                raise ValueError(f"tpgp_map_add: status {st}")
        return mp

    @classmethod
    def from_sorted(cls, n_spins, keys, values):
        """Install arrays that are already canonical (sorted unique, nonzero) without re-adding."""
        mp = cls(n_spins)
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1)
        values = np.ascontiguousarray(values, dtype=np.float64)
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]
        kp = libc.malloc(max(keys.nbytes, 8))
        vp = libc.malloc(max(values.nbytes, 8))
        if not kp or not vp:
            raise MemoryError("host map allocation failed")
        C.memmove(kp, keys.ctypes.data, keys.nbytes)
        C.memmove(vp, values.ctypes.data, values.nbytes)
        mp.m.keys = C.cast(kp, C.POINTER(C.c_uint64))
        mp.m.values = C.cast(vp, C.POINTER(C.c_double))
        mp.m.len = mp.m.capacity = len(values)
        return mp

    def arrays(self):
        n, w = self.m.len, self.m.n_words
        if n == 0:
            return np.zeros((0, w), np.uint64), np.zeros(0)
        keys = np.ctypeslib.as_array(self.m.keys, shape=(n * w,)).reshape(n, w).copy()
        vals = np.ctypeslib.as_array(self.m.values, shape=(n,)).copy()
        return keys, vals

    def single_sum(self, pauli=1):
        return self.eng.tpgp_map_single_sum(C.byref(self.m), pauli)

    def readout(self, pauli, normalizer):
        mid, rad = C.c_double(), C.c_double()
        st = self.eng.tpgp_readout(C.byref(self.m), pauli, normalizer, C.byref(mid),
                                   C.byref(rad))
        return st, mid.value, rad.value


def options(threshold=0.0, max_terms=0, validated=0, max_drop_l2=0.0, context=None,
            backend=N.TPGP_HIP, threads=1):
    o = N.TpgpOptions()
    o.threshold, o.max_terms, o.threads, o.backend = threshold, max_terms, threads, backend
    o.validated, o.max_drop_l2 = validated, max_drop_l2
    o.hip_context = context.ptr if context is not None else None
    return o


def ledger(validated=0):
    led = N.TpgpLedger()
    N.engine().tpgp_ledger_init(C.byref(led), validated)
    return led


def rotate(mp: Map, i, j, qi, qj, theta, opts, led):
    return N.engine().tpgp_rotate(C.byref(mp.m), i, j, qi, qj, theta, C.byref(opts),
                                  C.byref(led))


def strang_step(mp: Map, terms_arr, count, dt, opts, led):
    return N.engine().tpgp_strang_step(C.byref(mp.m), terms_arr, count, dt, C.byref(opts),
                                       C.byref(led))


def strang_bound(terms, dt):
    """Per-step Strang operator bound of the term table (tpgp_strang_bound, tpgp_rev.c:887-961)."""
    return N.engine().tpgp_strang_bound(N.terms_array(terms), len(terms), dt)


def fid_allowance(pf, led, n_spins):
    """FID allowance of the global initial state sum_s X_s after the accumulated Strang bound
    `pf`: pf + (dropped_l2 + arithmetic_l2) / sqrt(N), summed upward (the no-MPFR row of
    cli_rev.c:584-602 with input_terms = N, as the run driver prints it)."""
    e = N.engine()
    loss = (led.dropped_l2 + led.arithmetic_l2) / math.sqrt(n_spins)
    return e.tpgp_bound_add_up(pf, math.nextafter(loss, math.inf))


def residual_bound(a: Map, b: Map, terms, dt, depth, reference=False):
    """Independent endpoint bound of one step from a to b (tpgp_residual_bound, residual_rev.c:
    181-194); returns (status, upper)."""
    e = N.engine()
    fn = e.tpgp_residual_bound_reference if reference else e.tpgp_residual_bound
    up = C.c_double(float("nan"))
    st = fn(C.byref(a.m), C.byref(b.m), N.terms_array(terms), len(terms), dt, depth,
            C.byref(up))
    return st, up.value


def liouvillian(mp: Map, terms):
    """-i[H, A] on the GPU (tpgp_liouvillian); returns (status, Map)."""
    out = Map(mp.n_spins)
    ta = N.terms_array(terms)
    st = N.engine().tpgp_liouvillian(C.byref(out.m), C.byref(mp.m), ta, len(terms))
    return st, out


def evolve(mp: Map, terms, dt, steps, opts, led, exact_ledger=False, pauli=1, normalizer=None,
           initial_capacity=0):
    """Resident evolution (tpgp_hip_evolve); returns (status, fid, fid_radius, active, accepted)."""
    ta = N.terms_array(terms)
    eo = N.EvolveOptions(1 if exact_ledger else 0, pauli,
                         normalizer if normalizer else mp.n_spins, initial_capacity)
    fid = np.full(steps + 1, np.nan)
    rad = np.full(steps + 1, np.nan)
    act = np.zeros(steps + 1, np.uint64)
    acc = C.c_size_t()
    st = N.engine().tpgp_hip_evolve(C.byref(mp.m), ta, len(terms), dt, steps, C.byref(opts),
                                    C.byref(eo), C.byref(led), N.as_dp(fid), N.as_dp(rad),
                                    act.ctypes.data_as(C.POINTER(C.c_size_t)), C.byref(acc))
    return st, fid, rad, act, acc.value


class Evolution:
    """Resident evolution session (tpgp_evolution_open/run/close): the map stays in HBM."""

    def __init__(self, mp: Map, terms, dt, opts, exact_ledger=False, pauli=1, normalizer=None,
                 initial_capacity=0, context: Context | None = None):
        self.eng = N.engine()
        self.n_spins = mp.n_spins
        self._terms = N.terms_array(terms)
        self._opts = opts
        self._ctx = context  # keeps the caller's Context (if given) alive as long as the session
        if context is not None:
            context._children.add(self)
        self.eo = N.EvolveOptions(1 if exact_ledger else 0, pauli,
                                  normalizer if normalizer else mp.n_spins, initial_capacity)
        self.ptr = C.c_void_p()
        st = self.eng.tpgp_evolution_open(C.byref(self.ptr), C.byref(mp.m), self._terms,
                                          len(terms), dt, C.byref(opts), C.byref(self.eo))
        if st:
            raise RuntimeError(f"tpgp_evolution_open: status {st}: "
                               f"{self.eng.tpgp_hip_last_error().decode()}")

    def run(self, steps, led):
        fid = np.full(steps + 1, np.nan)
        rad = np.full(steps + 1, np.nan)
        act = np.zeros(steps + 1, np.uint64)
        acc = C.c_size_t()
        st = self.eng.tpgp_evolution_run(self.ptr, steps, C.byref(led), N.as_dp(fid),
                                         N.as_dp(rad), act.ctypes.data_as(C.POINTER(C.c_size_t)),
                                         C.byref(acc))
        return st, fid, rad, act, acc.value

    def counters(self):
        c = N.EvolutionCounters()
        self.eng.tpgp_evolution_counters_get(self.ptr, C.byref(c))
        return c

    def profile(self, mask=N.PROFILE_ALL):
        """Enable HIP-event timing for the kinds in `mask` (bit k = kind k; 0 disables)."""
        return self.eng.tpgp_evolution_profile(self.ptr, int(mask))

    def graph_counts(self):
        """(steps recorded as HIP graphs, steps replayed from them)."""
        rec, rep = C.c_uint64(), C.c_uint64()
        self.eng.tpgp_evolution_graph_counts(self.ptr, C.byref(rec), C.byref(rep))
        return rec.value, rep.value

    def profile_read(self, kind):
        ms, n = C.c_double(), C.c_uint64()
        self.eng.tpgp_evolution_profile_read(self.ptr, kind, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def snapshot(self) -> Map:
        """The last accepted map, copied out without closing the session."""
        mp = Map(self.n_spins)
        st = self.eng.tpgp_evolution_snapshot(self.ptr, C.byref(mp.m))
        if st:
            raise RuntimeError(f"tpgp_evolution_snapshot: status {st}: "
                               f"{self.eng.tpgp_hip_last_error().decode()}")
        return mp

    def l2(self) -> float:
        """l2 norm of the last accepted map (device reduction)."""
        out = C.c_double()
        st = self.eng.tpgp_evolution_l2(self.ptr, C.byref(out))
        if st:
            raise RuntimeError(f"tpgp_evolution_l2: status {st}: "
                               f"{self.eng.tpgp_hip_last_error().decode()}")
        return out.value

    def close(self, mp: Map | None = None):
        if self.ptr:
            st = self.eng.tpgp_evolution_close(self.ptr, C.byref(mp.m) if mp is not None else None)
            self.ptr = None
            return st
        return 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Checkpoint:
    """A TPGPRVCP version-2 restart record (tpgp_checkpoint_*; checkpoint_rev.h:7-40).

    ``load`` returns a record that owns its arrays (freed with the object); ``of`` builds one
    that borrows the caller's map and terms, for saving."""

    def __init__(self):
        self.eng = N.engine()
        self.c = N.Checkpoint()
        self._owned = False
        self._keep = None

    @classmethod
    def load(cls, path):
        ck = cls()
        err = C.create_string_buffer(512)
        st = ck.eng.tpgp_checkpoint_load(str(path).encode(), C.byref(ck.c), err, 512)
        if st:
            raise ValueError(f"checkpoint rejected: {err.value.decode()} (status {st})")
        ck._owned = True
        return ck

    @classmethod
    def of(cls, mp: Map, terms, opts, led, dt, step, total_steps, model="ring", j2=0.0,
           linear_size=0, initial_global=True, pf_bound=0.0, residual_depth=-1,
           residual_bound=float("nan"), max_fid_bound=0.0, resumed_from_step=0,
           prior_total_steps=0, prior_max_terms=0, stop_reason=0, last_status=0):
        ck = cls()
        c = ck.c
        ta = N.terms_array(terms) if not isinstance(terms, C.Array) else terms
        ck._keep = (mp, ta)
        c.map = mp.m
        c.terms = C.cast(ta, C.POINTER(N.TpgpTerm))
        c.term_count = len(terms)
        c.options = opts
        c.options.hip_context = None
        c.ledger = led
        c.dt, c.pf_bound, c.residual_bound, c.max_fid_bound = dt, pf_bound, residual_bound, \
            max_fid_bound
        c.step, c.total_steps = step, total_steps
        c.residual_depth, c.initial_global = residual_depth, 1 if initial_global else 0
        c.model = model.encode()
        c.j2, c.linear_size = j2, linear_size
        c.resumed_from_step, c.prior_total_steps = resumed_from_step, prior_total_steps
        c.prior_max_terms, c.stop_reason, c.last_status = prior_max_terms, stop_reason, last_status
        return ck

    def save(self, path):
        err = C.create_string_buffer(512)
        st = self.eng.tpgp_checkpoint_save(str(path).encode(), C.byref(self.c), err, 512)
        if st:
            raise ValueError(f"checkpoint save failed: {err.value.decode()} (status {st})")

    def same_configuration(self, other: "Checkpoint"):
        return bool(self.eng.tpgp_checkpoint_same_configuration(C.byref(self.c),
                                                                C.byref(other.c)))

    def map(self) -> Map:
        """A copy of the saved map."""
        mp = Map(self.c.map.n_spins)
        st = self.eng.tpgp_map_clone(C.byref(mp.m), C.byref(self.c.map))
        if st:
            raise MemoryError(f"tpgp_map_clone: status {st}")
        return mp

    def terms(self):
        """The saved Hamiltonian as (i, j, qi, qj, h) tuples, in order."""
        t = self.c.terms
        return [(t[k].i, t[k].j, t[k].qi, t[k].qj, t[k].h) for k in range(self.c.term_count)]

    def __getattr__(self, name):
        if name in ("eng", "c", "_owned", "_keep"):
            raise AttributeError(name)
        return getattr(self.c, name)

    def __del__(self):
        try:
            if self._owned:
                self.eng.tpgp_checkpoint_free(C.byref(self.c))
                self._owned = False
        except Exception:
            pass


class Fleet:
    """Many independent FIDs in one resident batch (tpgp_fleet_*): batched powder orientations,
    coupling draws, or served requests.  Each slot runs exactly the loop of ``evolve``."""

    def __init__(self, n_words, n_slots, slot_capacity=1 << 16, context: Context | None = None):
        self.eng = N.engine()
        self.n_slots = n_slots
        self.ptr = C.c_void_p()
        self._ctx = context  # the context must outlive the fleet
        if context is not None:
            context._children.add(self)
        st = self.eng.tpgp_fleet_open(C.byref(self.ptr), context.ptr if context else None,
                                      n_words, n_slots, slot_capacity)
        if st:
            raise RuntimeError(f"tpgp_fleet_open: status {st}: "
                               f"{self.eng.tpgp_hip_last_error().decode()}")

    def admit(self, slot, initial: Map, terms, dt, steps, opts, exact_ledger=False, pauli=1,
              normalizer=None):
        ta = N.terms_array(terms)
        job = N.FidJob()
        job.initial = C.pointer(initial.m)
        job.terms = C.cast(ta, C.POINTER(N.TpgpTerm))
        job.count = len(terms)
        job.dt = dt
        job.steps = steps
        job.options = opts
        job.evolve = N.EvolveOptions(1 if exact_ledger else 0, pauli,
                                     normalizer if normalizer else initial.n_spins, 0)
        return self.eng.tpgp_fleet_admit(self.ptr, slot, C.byref(job))

    def step(self):
        running = C.c_size_t()
        st = self.eng.tpgp_fleet_step(self.ptr, C.byref(running))
        return st, running.value

    def slot(self, slot):
        state, status = C.c_int(), C.c_int()
        done = C.c_size_t()
        led = N.TpgpLedger()
        self.eng.tpgp_fleet_slot(self.ptr, slot, C.byref(state), C.byref(done), C.byref(status),
                                 C.byref(led))
        return state.value, done.value, status.value, led

    def fid(self, slot, capacity):
        fid = np.full(capacity, np.nan)
        rad = np.full(capacity, np.nan)
        act = np.zeros(capacity, np.uint64)
        n = C.c_size_t()
        self.eng.tpgp_fleet_fid(self.ptr, slot, N.as_dp(fid), N.as_dp(rad),
                                act.ctypes.data_as(C.POINTER(C.c_size_t)), capacity, C.byref(n))
        return fid[:n.value], rad[:n.value], act[:n.value]

    def release(self, slot, final: Map | None = None):
        return self.eng.tpgp_fleet_release(self.ptr, slot,
                                           C.byref(final.m) if final is not None else None)

    def counters(self):
        c = N.EvolutionCounters()
        self.eng.tpgp_fleet_counters_get(self.ptr, C.byref(c))
        return c

    def slot_bytes(self, slot):
        return self.eng.tpgp_fleet_slot_bytes(self.ptr, slot)

    def held_bytes(self):
        """Device bytes all slots hold (free ones included: the batch keeps their share)."""
        return sum(self.slot_bytes(s) for s in range(self.n_slots))

    def set_budget(self, nbytes):
        """Ceiling on the bytes all slots hold together (tpgp_fleet_set_budget); 0: none."""
        return self.eng.tpgp_fleet_set_budget(self.ptr, int(nbytes))

    def trim(self):
        """Return grown FREE slots to the opening slot capacity (tpgp_fleet_trim)."""
        return self.eng.tpgp_fleet_trim(self.ptr)

    def graph_counts(self):
        """(rounds recorded as HIP graphs, rounds replayed from them)."""
        rec, rep = C.c_uint64(), C.c_uint64()
        self.eng.tpgp_fleet_graph_counts(self.ptr, C.byref(rec), C.byref(rep))
        return rec.value, rep.value

    def close(self):
        if self.ptr:
            self.eng.tpgp_fleet_close(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

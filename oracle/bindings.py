"""ctypes bindings for the CPU oracle and for the reference engine built from its own sources.

TEST INFRASTRUCTURE ONLY -- used as the checker by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.

Two shared libraries share one calling surface here:

* ``oracle/_build/libtpgp_oracle.so`` -- our C17 restatement (``orc_*`` symbols).
* ``oracle/_ref/libtpgp_ref.so`` -- the reference CPU engine (engine/src/tpgp_rev.c and friends)
  compiled by ``make -C oracle ref`` (``tpgp_*`` symbols).  Present only where it was built.

Both use the reference's struct layouts: map {n_spins, n_words, len, capacity, keys, values}
(engine/include/tpgp_rev.h:25-29), term {i, j, qi, qj, h, layer} (tpgp_rev.h:30-35) and ledger
(tpgp_rev.h:46-51).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libtpgp_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtpgp_ref.so")

OK, EINVAL, ENOMEM, ECAP, ENUMERIC, EBACKEND, EBUDGET = range(7)
SERIAL, PTHREAD, OPENMP = 0, 1, 2


class Map(C.Structure):
    _fields_ = [("n_spins", C.c_size_t), ("n_words", C.c_size_t), ("len", C.c_size_t),
                ("capacity", C.c_size_t), ("keys", C.POINTER(C.c_uint64)),
                ("values", C.POINTER(C.c_double))]


class Term(C.Structure):
    _fields_ = [("i", C.c_size_t), ("j", C.c_size_t), ("qi", C.c_uint), ("qj", C.c_uint),
                ("h", C.c_double), ("layer", C.c_size_t)]


class Ledger(C.Structure):
    _fields_ = [("dropped_l1", C.c_double), ("dropped_l2", C.c_double),
                ("dropped_l2_squared", C.c_double), ("arithmetic_l2", C.c_double),
                ("rotations", C.c_size_t), ("removal_events", C.c_size_t),
                ("peak_terms", C.c_size_t), ("arithmetic_enclosed", C.c_int)]

    def as_tuple(self):
        return (self.dropped_l1, self.dropped_l2, self.dropped_l2_squared, self.arithmetic_l2,
                self.rotations, self.removal_events, self.peak_terms, self.arithmetic_enclosed)

    def bits(self):
        """Bit patterns of the floating fields plus the integer fields (for bitwise equality)."""
        f = np.array([self.dropped_l1, self.dropped_l2, self.dropped_l2_squared,
                      self.arithmetic_l2], dtype=np.float64).view(np.uint64).tolist()
        return tuple(f) + (self.rotations, self.removal_events, self.peak_terms,
                           self.arithmetic_enclosed)


class OracleOptions(C.Structure):
    _fields_ = [("threshold", C.c_double), ("max_terms", C.c_size_t), ("threads", C.c_uint),
                ("backend", C.c_int), ("validated", C.c_int), ("max_drop_l2", C.c_double)]


class RefOptions(C.Structure):  # engine/include/tpgp_rev.h:36-45
    _fields_ = [("threshold", C.c_double), ("max_terms", C.c_size_t), ("threads", C.c_uint),
                ("backend", C.c_int), ("validated", C.c_int), ("max_drop_l2", C.c_double),
                ("metal_context", C.c_void_p), ("cuda_context", C.c_void_p)]


def build(ref: bool = False, force: bool = False) -> None:
    """Compile the oracle (and the reference when its sources are present)."""
    extra = ["-B"] if force else []
    subprocess.run(["make", "-s", *extra, "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", *extra, "-C", HERE, "ref"], check=True)


def terms_array(terms):
    """Term tuples (i, j, qi, qj, h) -> ctypes array in the reference layout."""
    arr = (Term * max(len(terms), 1))()
    for k, t in enumerate(terms):
        arr[k] = Term(t[0], t[1], t[2], t[3], t[4], k)
    return arr


class CpuEngine:
    """The oracle or the reference engine behind one Python surface."""

    def __init__(self, kind: str = "oracle"):
        self.kind = kind
        path = ORACLE_SO if kind == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built")
        self.lib = L = C.CDLL(path)
        p = "orc_" if kind == "oracle" else "tpgp_"
        self.p = p
        self.Options = OracleOptions if kind == "oracle" else RefOptions
        mp = C.POINTER(Map)
        self._f = {}

        def fn(name, res, *args):
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = list(args)
            self._f[name] = f

        fn("has_mpfr", C.c_int)
        fn("has_openmp", C.c_int)
        fn("ledger_init", None, C.POINTER(Ledger), C.c_int)
        fn("map_add", C.c_int, mp, C.POINTER(C.c_uint64), C.c_double)
        fn("map_add_single", C.c_int, mp, C.c_size_t, C.c_uint, C.c_double)
        fn("map_l2", C.c_double, mp)
        fn("map_single_sum", C.c_double, mp, C.c_uint)
        fn("readout", C.c_int, mp, C.c_uint, C.c_size_t, C.POINTER(C.c_double),
           C.POINTER(C.c_double))
        fn("pauli_phase", C.c_uint, C.c_uint, C.c_uint)
        opt = C.POINTER(self.Options)
        fn("rotate", C.c_int, mp, C.c_size_t, C.c_size_t, C.c_uint, C.c_uint, C.c_double, opt,
           C.POINTER(Ledger))
        fn("strang_step", C.c_int, mp, C.POINTER(Term), C.c_size_t, C.c_double, opt,
           C.POINTER(Ledger))
        fn("strang_bound_taylor", C.c_double, C.POINTER(Term), C.c_size_t, C.c_double)
        fn("strang_bound", C.c_double, C.POINTER(Term), C.c_size_t, C.c_double)
        fn("bound_add_up", C.c_double, C.c_double, C.c_double)
        fn("bound_fid_up", C.c_double, C.c_double, C.c_size_t)
        fn("liouvillian", C.c_int, mp, mp, C.POINTER(Term), C.c_size_t)
        if kind == "oracle":
            fn("map_new", mp, C.c_size_t)
            fn("map_delete", None, mp)
            fn("map_assign", C.c_int, mp, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
               C.c_size_t)
            fn("rotate_product", C.c_int, mp, C.c_size_t, C.c_size_t, C.c_uint, C.c_uint,
               C.c_double, C.c_double, opt, C.POINTER(Ledger))
            fn("trig", C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double))
            fn("branches", None, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_size_t,
               C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint, C.c_uint, C.c_double, C.c_double,
               C.c_double, C.c_double, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
               C.POINTER(C.c_double))
            for nm in ("upadd", "upmul"):
                fn(nm, C.c_double, C.c_double, C.c_double)
            fn("round_radius", C.c_double, C.c_double)
            fn("upper_l2", C.c_double, C.POINTER(C.c_double), C.c_size_t)
        else:
            fn("map_init", C.c_int, mp, C.c_size_t)
            fn("map_free", None, mp)

    def __getattr__(self, name):
        f = self.__dict__.get("_f", {}).get(name)
        if f is None:
            raise AttributeError(name)
        return f

    # ---- maps ----------------------------------------------------------------------------
    def new_map(self, n_spins: int) -> "CpuMap":
        return CpuMap(self, n_spins)

    def options(self, threshold=0.0, max_terms=0, threads=1, backend=SERIAL, validated=0,
                max_drop_l2=0.0):
        o = self.Options()
        o.threshold, o.max_terms, o.threads = threshold, max_terms, threads
        o.backend, o.validated, o.max_drop_l2 = backend, validated, max_drop_l2
        return o

    def ledger(self, validated=0) -> Ledger:
        led = Ledger()
        self._f["ledger_init"](C.byref(led), validated)
        return led


class CpuMap:
    def __init__(self, eng: CpuEngine, n_spins: int):
        self.eng = eng
        self.n_spins = n_spins
        if eng.kind == "oracle":
            self.ptr = eng.map_new(n_spins)
        else:
            self._m = Map()
            if eng.map_init(C.byref(self._m), n_spins) != OK:
                raise ValueError("map_init failed")
            self.ptr = C.pointer(self._m)

    def __del__(self):
        try:
            if self.eng.kind == "oracle":
                self.eng.map_delete(self.ptr)
            else:
                self.eng.map_free(self.ptr)
        except Exception:
            pass

    @property
    def n_words(self):
        return self.ptr.contents.n_words

    def __len__(self):
        return self.ptr.contents.len

    def arrays(self):
        m = self.ptr.contents
        n, w = m.len, m.n_words
        if n == 0:
            return np.zeros((0, w), np.uint64), np.zeros(0, np.float64)
        keys = np.ctypeslib.as_array(m.keys, shape=(n * w,)).reshape(n, w).copy()
        vals = np.ctypeslib.as_array(m.values, shape=(n,)).copy()
        return keys, vals

    def assign(self, keys: np.ndarray, values: np.ndarray) -> int:
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1)
        values = np.ascontiguousarray(values, dtype=np.float64)
        if self.eng.kind == "oracle":
            return self.eng.map_assign(self.ptr, keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                                       values.ctypes.data_as(C.POINTER(C.c_double)), len(values))
        w = self.n_words
        for k in range(len(values)):
            kk = np.ascontiguousarray(keys[k * w:(k + 1) * w])
            s = self.eng.map_add(self.ptr, kk.ctypes.data_as(C.POINTER(C.c_uint64)), values[k])
            if s:
                return s
        return OK

    def add_single(self, site, pauli, value):
        return self.eng.map_add_single(self.ptr, site, pauli, value)

    def single_sum(self, pauli=1):
        return self.eng.map_single_sum(self.ptr, pauli)

    def readout(self, pauli, normalizer):
        mid, rad = C.c_double(), C.c_double()
        s = self.eng.readout(self.ptr, pauli, normalizer, C.byref(mid), C.byref(rad))
        return s, mid.value, rad.value

    def rotate(self, i, j, qi, qj, theta, opts, ledger):
        return self.eng.rotate(self.ptr, i, j, qi, qj, theta, C.byref(opts), C.byref(ledger))

    def rotate_product(self, i, j, qi, qj, h, dur, opts, ledger):
        return self.eng.rotate_product(self.ptr, i, j, qi, qj, h, dur, C.byref(opts),
                                       C.byref(ledger))

    def strang_step(self, terms_arr, count, dt, opts, ledger):
        return self.eng.strang_step(self.ptr, terms_arr, count, dt, C.byref(opts),
                                    C.byref(ledger))

    def l2(self):
        return self.eng.map_l2(self.ptr)

    def liouvillian(self, terms_arr, count) -> tuple:
        """-i[H, A]: (status, result map)."""
        out = self.eng.new_map(self.n_spins)
        st = self.eng.liouvillian(out.ptr, self.ptr, terms_arr, count)
        return st, out

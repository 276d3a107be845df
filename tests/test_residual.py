"""The independent endpoint (residual) bound, tpgp_residual_bound: CPU part.

The bound's interval vectors are built on the GPU (tests/test_residual_gpu.py holds the parity
tests: bit for bit against the reference library and against the exact step).  Here, without a
device: the same refusals as the reference (EINVAL for malformed input, depth > 10, dt <= 0),
decided before any device work, and the helpers the GPU tests share.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB

OK, EINVAL = 0, 1


def has_mpfr():
    return N.engine().tpgp_has_mpfr() == 1


pytestmark = pytest.mark.skipif(not has_mpfr(), reason="built without MPFR")


def trajectory(oracle, n, terms, dt, steps, thr):
    m = oracle.new_map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    o = oracle.options(threshold=thr, max_terms=10 ** 8)
    led = oracle.ledger(0)
    ta = OB.terms_array(terms)
    maps = []
    for k in range(steps + 1):
        keys, vals = m.arrays()
        maps.append(E.Map.from_sorted(n, keys, vals))
        if k < steps:
            assert m.strang_step(ta, len(terms), dt, o, led) == 0
    return maps


@pytest.fixture(scope="module")
def ref_lib():
    if not os.path.exists(OB.REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    lib = C.CDLL(OB.REF_SO)
    for name in ("tpgp_residual_bound", "tpgp_residual_bound_reference"):
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double, C.c_uint,
                       C.POINTER(C.c_double)]
    return lib


def ref_bound(lib, a, b, terms, dt, depth, reference):
    fn = lib.tpgp_residual_bound_reference if reference else lib.tpgp_residual_bound
    up = C.c_double(float("nan"))
    ta = N.terms_array(terms)
    st = fn(C.addressof(a.m), C.addressof(b.m), C.cast(ta, C.c_void_p), len(terms), dt, depth,
            C.byref(up))
    return st, up.value


def bits(x):
    return np.float64(x).view(np.uint64)


def pauli_index_generator(n, terms):
    """Dense L over the 4^n Pauli strings (index = key bits): L[P ^ q, P] += +-2h for every term
    anticommuting with P, the sign from the phase of the product (I0 X1 Y2 Z3)."""
    dim = 4 ** n
    L = np.zeros((dim, dim))

    def phase(q, p):
        if not q or not p or q == p:
            return 0
        return 1 if (p + 3 - q) % 3 == 1 else 3

    for key in range(dim):
        for (i, j, qi, qj, h) in terms:
            a, b = (key >> (2 * i)) & 3, (key >> (2 * j)) & 3
            p = (phase(qi, a) + phase(qj, b)) & 3
            if p & 1 and h != 0:
                L[key ^ (qi << (2 * i)) ^ (qj << (2 * j)), key] += 2 * h if p == 1 else -2 * h
    return L


def test_refusals_match_reference(ref_lib):
    n = 4
    terms = H.ring(n)
    a = E.Map(n)
    for s in range(n):
        a.add_single(s, 1, 1.0)
    b = E.Map(n)
    b.add_single(0, 1, 0.5)
    other = E.Map(6)
    other.add_single(0, 1, 1.0)
    cases = [(a, b, terms, 0.01, 11), (a, b, terms, 0.0, 1), (a, b, terms, -0.01, 1),
             (a, b, terms, float("nan"), 1), (a, other, terms, 0.01, 1),
             (a, b, [(0, 0, 3, 3, 1.0)], 0.01, 1), (a, b, [(0, 5, 3, 3, 1.0)], 0.01, 1),
             (a, b, [(0, 1, 4, 3, 1.0)], 0.01, 1), (a, b, [(0, 1, 3, 3, float("inf"))], 0.01, 1)]
    for (x, y, t, dt, depth) in cases:
        st, up = E.residual_bound(x, y, t, dt, depth)
        rst, _ = ref_bound(ref_lib, x, y, t, dt, depth, False)
        assert st == rst == EINVAL, (dt, depth, st, rst)
        assert np.isnan(up)  # output untouched on failure


def test_no_cpu_fallback_without_a_device():
    """The vectors are built on the GPU only: with no device a valid call is EBACKEND (the
    reference's status for a missing backend), never a host computation."""
    err = C.create_string_buffer(256)
    ctx = N.hip().tpgp_hip_create(0, err, 256)
    if ctx:
        N.hip().tpgp_hip_destroy(ctx)
        pytest.skip("a device is present")
    a = E.Map(4)
    for s in range(4):
        a.add_single(s, 1, 1.0)
    st, up = E.residual_bound(a, a, H.ring(4), 0.01, 1)
    assert st == N.EBACKEND and np.isnan(up)

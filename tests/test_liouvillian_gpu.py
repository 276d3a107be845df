"""tpgp_liouvillian on the GPU: bit for bit the reference's -i[H, A] (tpgp_rev.c:810-857).

Against the golden cases written from the reference library (every model, one and two key words,
cancellations, both failure statuses), against the oracle (pinned to the reference on the CPU)
at sizes of 1e4-1e5 input keys, and the C ABI's capacity protocol.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import liouv_golden as LG
from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB

pytestmark = pytest.mark.gpu


def bits(v):
    return np.asarray(v, np.float64).view(np.uint64)


def canonical(n, count, seed):
    rng = np.random.default_rng(seed)
    w = (n + 31) // 32
    digits = rng.integers(0, 4, size=(count, n), dtype=np.uint64)
    keys = np.zeros((count, w), np.uint64)
    for s in range(n):
        keys[:, s // 32] |= digits[:, s] << np.uint64(2 * (s % 32))
    keys = np.unique(keys, axis=0)
    keys = keys[np.lexsort(tuple(keys[:, k] for k in range(w)))]
    return keys, rng.standard_normal(len(keys))


def test_golden_cases(hip_ctx):
    for case in LG.cases():
        print(f"liouvillian golden {case['name']}: N={case['n']}, {len(case['vals'])} keys, "
              f"{len(case['terms'])} terms", flush=True)
        mp = E.Map.from_sorted(case["n"], case["keys"], case["vals"])
        st, out = E.liouvillian(mp, case["terms"])
        k, v = out.arrays() if st == 0 else (np.zeros(0, np.uint64), np.zeros(0))
        assert LG.matches(case, st, k, v), (case["name"], st)


@pytest.mark.parametrize("n,count,kind", [(16, 20000, "j1j2"), (12, 60000, "ring"),
                                          (40, 8000, "random"), (64, 5000, "random")])
def test_matches_oracle_large(hip_ctx, oracle, n, count, kind):
    keys, vals = canonical(n, count, n)
    if kind == "j1j2":
        terms = H.j1j2_n16_terms(n)
    elif kind == "ring":
        terms = H.ring(n)
    else:
        rng = np.random.default_rng(n + 1)
        terms = []
        for _ in range(60):
            i, j = rng.choice(n, 2, replace=False)
            terms.append((int(i), int(j), int(rng.integers(1, 4)), int(rng.integers(1, 4)),
                          float(rng.standard_normal())))
    print(f"liouvillian N={n}: {len(vals)} keys, {len(terms)} terms", flush=True)
    st, out = E.liouvillian(E.Map.from_sorted(n, keys, vals), terms)
    om = oracle.new_map(n)
    assert om.assign(keys, vals) == 0
    st_o, g = om.liouvillian(OB.terms_array(terms), len(terms))
    assert st == st_o == 0
    ka, va = out.arrays()
    kb, vb = g.arrays()
    assert np.array_equal(ka, kb) and np.array_equal(bits(va), bits(vb))


def test_capacity_protocol(hip_ctx):
    """A result longer than the buffers: ECAPACITY with the length, nothing written; then the
    full result with buffers of that length."""
    n = 10
    keys, vals = canonical(n, 500, 3)
    terms = H.ring(n)
    ta = N.terms_array(terms)
    lib = N.hip()
    flat = np.ascontiguousarray(keys.reshape(-1))
    got = C.c_size_t()
    err = C.create_string_buffer(256)
    small_k = np.full(4, 7, np.uint64)
    small_v = np.full(4, 7.0)
    st = lib.tpgp_hip_liouvillian(hip_ctx.ptr, N.as_u64p(flat), N.as_dp(vals), len(vals), 1, n,
                                  C.cast(ta, C.POINTER(N.TpgpTerm)), len(terms),
                                  N.as_u64p(small_k), N.as_dp(small_v), 4, C.byref(got), err,
                                  256)
    assert st == N.TPGP_HIP_ECAPACITY if hasattr(N, "TPGP_HIP_ECAPACITY") else st == 16
    assert got.value > 4 and np.all(small_k == 7)
    ok = np.zeros(got.value, np.uint64)
    ov = np.zeros(got.value)
    st = lib.tpgp_hip_liouvillian(hip_ctx.ptr, N.as_u64p(flat), N.as_dp(vals), len(vals), 1, n,
                                  C.cast(ta, C.POINTER(N.TpgpTerm)), len(terms), N.as_u64p(ok),
                                  N.as_dp(ov), got.value, C.byref(got), err, 256)
    assert st == 0
    _, ref = E.liouvillian(E.Map.from_sorted(n, keys, vals), terms)
    rk, rv = ref.arrays()
    assert np.array_equal(ok, rk.reshape(-1)) and np.array_equal(bits(ov), bits(rv))

"""-i[H, A] (tpgp_liouvillian, tpgp_rev.c:810-857): the oracle's restatement against the reference
compiled from its own sources, and against what the algebra fixes.

Pins, on the CPU:
* the reference's own known answer (engine/tests/test_core_rev.c:163-171);
* bit-identical maps and statuses against the reference library on random maps and Hamiltonians
  (N = 3..64, one and two key words; ring, J1-J2, cubic and random bonds; cancelling
  contributions; invalid terms and overflow, alone and in either order);
* the algebra: for Pauli strings P, Q the commutator of single terms is 2i or 0 times the product,
  so L applied to each single key equals the sum over anticommuting terms of a signed doubled
  coupling -- checked against 4x4 Pauli matrices for all 256 two-site cases; and antisymmetry
  <B, L A> = -<L B, A> (the Liouvillian of a Hermitian H is antisymmetric in the Hilbert-Schmidt
  product).
"""
from __future__ import annotations

import itertools
import math
import os

import numpy as np
import pytest

from ddevol_amd import hamiltonians as H
from oracle import bindings as OB

PAULI = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]),
         np.array([[1, 0], [0, -1]])]


def arrays_bits(m):
    k, v = m.arrays()
    return k, np.asarray(v, np.float64).view(np.uint64)


def random_map(eng, n, count, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = (n + 31) // 32
    digits = rng.integers(0, 4, size=(count, n), dtype=np.uint64)
    keys = np.zeros((count, w), np.uint64)
    for s in range(n):
        keys[:, s // 32] |= digits[:, s] << np.uint64(2 * (s % 32))
    keys = np.unique(keys, axis=0)  # canonical: unique, ordered with the last word first
    keys = keys[np.lexsort(tuple(keys[:, k] for k in range(w)))]
    vals = rng.standard_normal(len(keys)) * scale
    m = eng.new_map(n)
    assert m.assign(keys, vals) == 0
    return m


def random_terms(n, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        i, j = rng.choice(n, 2, replace=False)
        out.append((int(i), int(j), int(rng.integers(1, 4)), int(rng.integers(1, 4)),
                    float(rng.standard_normal())))
    return out


def test_oracle_known_answer(oracle):
    m = oracle.new_map(3)
    assert m.add_single(0, 1, 1.0) == 0
    terms = [(0, 1, 3, 3, .5), (1, 2, 1, 1, -.25), (0, 2, 2, 2, .3)]
    st, g = m.liouvillian(OB.terms_array(terms), 3)
    assert st == 0
    keys, vals = g.arrays()
    got = {int(k[0]): float(v) for k, v in zip(keys, vals)}
    yz, zy = 2 | (3 << 2), 3 | (2 << 4)
    assert got == {yz: 1.0, zy: -0.6}


def test_oracle_single_strings_match_matrices(oracle):
    """All 16 x 16 two-site strings P, terms Q on (0, 1): L(P) = -i[hQ, P] as 4x4 matrices."""
    h = 0.37
    for p0, p1, q0, q1 in itertools.product(range(4), repeat=4):
        if not (q0 and q1):
            continue
        m = oracle.new_map(2)
        key = np.array([[p0 | (p1 << 2)]], np.uint64)
        assert m.assign(key, np.array([1.0])) == 0
        st, g = m.liouvillian(OB.terms_array([(0, 1, q0, q1, h)]), 1)
        assert st == 0
        Pm = np.kron(PAULI[p1], PAULI[p0])
        Qm = np.kron(PAULI[q1], PAULI[q0])
        want = -1j * h * (Qm @ Pm - Pm @ Qm)
        got = np.zeros((4, 4), complex)
        keys, vals = g.arrays()
        for k, v in zip(keys, vals):
            a, b = int(k[0]) & 3, (int(k[0]) >> 2) & 3
            got += v * np.kron(PAULI[b], PAULI[a])
        assert np.allclose(got, want, atol=1e-15), (p0, p1, q0, q1)


def hs_dot(a, b):
    ka, va = a.arrays()
    kb, vb = b.arrays()
    da = {tuple(k): v for k, v in zip(ka.tolist(), va)}
    return math.fsum(v * da.get(tuple(k), 0.0) for k, v in zip(kb.tolist(), vb))


def test_oracle_antisymmetry(oracle):
    n = 10
    terms = H.j1j2(n, 0.3)
    ta = OB.terms_array(terms)
    a = random_map(oracle, n, 400, 1)
    b = random_map(oracle, n, 400, 2)
    st, la = a.liouvillian(ta, len(terms))
    st2, lb = b.liouvillian(ta, len(terms))
    assert st == st2 == 0
    x, y = hs_dot(b, la), hs_dot(lb, a)
    assert abs(x + y) <= 1e-12 * (abs(x) + abs(y) + 1)


CASES = [
    (3, 5, "random", 4),
    (12, 300, "ring", 0),
    (16, 2000, "j1j2", 0),
    (27, 500, "cubic", 0),
    (33, 400, "random", 40),
    (48, 800, "random", 60),
    (64, 1000, "random", 80),
]


def case_terms(n, kind, count, seed):
    if kind == "ring":
        return H.ring(n)
    if kind == "j1j2":
        return H.j1j2_n16_terms(n)
    if kind == "cubic":
        return H.cubic(3)
    return random_terms(n, count, seed)


@pytest.mark.parametrize("n,count,kind,nterms", CASES)
def test_oracle_equals_reference(oracle, reference, n, count, kind, nterms):
    terms = case_terms(n, kind, nterms, n)
    ta = OB.terms_array(terms)
    for seed in range(3):
        a = random_map(oracle, n, count, 100 * n + seed)
        r = reference.new_map(n)
        k, v = a.arrays()
        assert r.assign(k, v) == 0
        st_o, go = a.liouvillian(ta, len(terms))
        st_r, gr = r.liouvillian(ta, len(terms))
        assert st_o == st_r == 0
        ko, vo = arrays_bits(go)
        kr, vr = arrays_bits(gr)
        assert np.array_equal(ko, kr) and np.array_equal(vo, vr), (n, kind, seed)


def test_oracle_cancellations_equal_reference(oracle, reference):
    """Maps built so that many contributions cancel exactly (the same coefficient reached from
    two keys with opposite signs), plus repeated terms."""
    n = 6
    terms = [(0, 1, 3, 3, 0.5), (0, 1, 3, 3, -0.5), (1, 2, 1, 1, 0.25), (2, 3, 2, 2, 0.25),
             (3, 4, 3, 3, 1.0)]
    ta = OB.terms_array(terms)
    keys = np.array([[k] for k in range(0, 4 ** n, 7)], np.uint64)
    vals = np.array([1.0 if (k % 3) else -1.0 for k in range(len(keys))])
    a, r = oracle.new_map(n), reference.new_map(n)
    assert a.assign(keys, vals) == 0 and r.assign(keys, vals) == 0
    st_o, go = a.liouvillian(ta, len(terms))
    st_r, gr = r.liouvillian(ta, len(terms))
    assert st_o == st_r == 0
    ko, vo = arrays_bits(go)
    kr, vr = arrays_bits(gr)
    assert np.array_equal(ko, kr) and np.array_equal(vo, vr)


@pytest.mark.parametrize("bad_at,overflow", [(1, False), (0, True), (2, True), (None, True)])
def test_oracle_statuses_equal_reference(oracle, reference, bad_at, overflow):
    """EINVAL at the first invalid term, ENUMERIC at the first overflow, in term order."""
    n = 8
    terms = random_terms(n, 4, 7)
    if overflow:
        terms[1] = (terms[1][0], terms[1][1], terms[1][2], terms[1][3], 1e300)
    if bad_at is not None:
        terms[bad_at] = (3, 3, 1, 1, 1.0)  # i == j
    ta = OB.terms_array(terms)
    a = random_map(oracle, n, 200, 3, scale=1e10 if overflow else 1.0)
    r = reference.new_map(n)
    k, v = a.arrays()
    assert r.assign(k, v) == 0
    st_o, _ = a.liouvillian(ta, len(terms))
    st_r, _ = r.liouvillian(ta, len(terms))
    assert st_o == st_r != 0, (st_o, st_r)


def test_oracle_reproduces_golden_cases(oracle):
    """The golden cases written from the reference library (tests/golden/liouv.npz)."""
    import liouv_golden as LG

    seen = 0
    for case in LG.cases():
        m = oracle.new_map(case["n"])
        assert m.assign(case["keys"], case["vals"]) == 0
        st, g = m.liouvillian(OB.terms_array(case["terms"]), len(case["terms"]))
        k, v = g.arrays() if st == 0 else (np.zeros(0, np.uint64), np.zeros(0))
        assert LG.matches(case, st, k, v), case["name"]
        seen += 1
    assert seen == 10

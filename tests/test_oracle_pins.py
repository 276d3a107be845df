"""CPU pins of the oracle's harness functions and of its rotation's composition.

Each function is pinned twice: bit for bit against the reference engine compiled from its own
sources (oracle/_ref: engine/src/tpgp_rev.c, residual_rev.c), and against what exact arithmetic
fixes, independently of both (Python fractions: the smallest double above a real number, the
nested-commutator constant of the Strang bound computed from the Pauli algebra by brute force).

* orc_strang_bound, orc_strang_bound_taylor   (tpgp_rev.c:887-961)
* orc_bound_add_up, orc_bound_fid_up          (residual_rev.c:195-201)
* orc_map_l2                                  (tpgp_rev.c:266-271)
* orc_rotate on single-term maps over the reference CUDA test's 16 edge coefficients x 16
  two-site keys x 16 generators x {plain, validated} (engine/tests/test_cuda_rev.c:111-122):
  the composition of orc_branches with the merge, truncation and ledger.
"""
import ctypes as C
import math
import random
import struct
import sys
from fractions import Fraction

import numpy as np
import pytest

from ddevol_amd import hamiltonians as H
from oracle import bindings as OB

DBL_MAX = sys.float_info.max
DBL_MIN = sys.float_info.min
TRUE_MIN = 5e-324
# engine/tests/test_cuda_rev.c:111-113
EDGE = [0.0, -0.0, TRUE_MIN, -TRUE_MIN, DBL_MIN, -DBL_MIN, 1.0, -1.0, DBL_MAX * .25,
        -DBL_MAX * .25, float.fromhex("0x1.fffffffffffffp-1"),
        -float.fromhex("0x1.fffffffffffffp-1"), 2.0 ** -1000, -2.0 ** -1000, 2.0 ** 1000,
        -2.0 ** 1000]


def bits(x):
    return struct.pack("<d", x)


def same(a, b):
    return bits(a) == bits(b) or (math.isnan(a) and math.isnan(b))


def five_configs():
    """Term tables and time steps of the five workloads (plain, J1-J2, powder, serving, large)."""
    g = H.InputPCG(7)
    return [
        ("ring12", H.ring(12), 0.01),
        ("j1j2_n16", H.j1j2_n16_terms(16), 0.005),
        ("cluster27", H.powder_terms(H.powder_orientations(3)[2]), 0.01),
        ("serving48", H.serving_ring(48, [0.5 + g.uniform() for _ in range(48)]), 0.01),
        ("range64", H.range_ring(64, 4), 0.01),
    ]


def random_terms(rng, n_sites, count, zero_frac=0.1):
    out = []
    for _ in range(count):
        i, j = rng.sample(range(n_sites), 2)
        h = 0.0 if rng.random() < zero_frac else rng.choice([-1, 1]) * rng.uniform(0.01, 2.0)
        out.append((i, j, rng.randrange(4), rng.randrange(4), h))
    return out


# ---- Strang bounds -----------------------------------------------------------------------
def test_strang_bound_bitwise_vs_reference(oracle, reference):
    cases = [(name, t, dt) for name, t, dt in five_configs()]
    rng = random.Random(20260711)
    for k in range(60):
        n = rng.randrange(2, 12)
        cases.append((f"random{k}", random_terms(rng, n, rng.randrange(0, 40)),
                      rng.choice([0.01, -0.02, 0.3, 1.0, 1e-6, 0.0])))
    # invalid tables give NaN in both
    cases += [("same-site", [(1, 1, 3, 3, 0.5)], 0.1), ("bad-symbol", [(0, 1, 4, 3, 0.5)], 0.1),
              ("nan-h", [(0, 1, 3, 3, math.nan)], 0.1), ("inf-dt", H.ring(4), math.inf)]
    for name, terms, dt in cases:
        ta = OB.terms_array(terms)
        for fn in ("strang_bound", "strang_bound_taylor"):
            a = getattr(oracle, fn)(ta, len(terms), dt)
            b = getattr(reference, fn)(ta, len(terms), dt)
            assert same(a, b), (name, fn, a, b)


def anticommute(a, b):
    """Two two-site Pauli terms anticommute iff an odd number of shared sites carry different,
    non-identity symbols (from the single-site algebra: X, Y, Z pairwise anticommute)."""
    sa = {a[0]: a[2], a[1]: a[3]}
    sb = {b[0]: b[2], b[1]: b[3]}
    odd = 0
    for s, p in sa.items():
        q = sb.get(s, 0)
        odd ^= int(p != 0 and q != 0 and p != q)
    return odd == 1


def nested_constant(terms):
    """C = sum_a h_a^2 sum_{b>a, b~a} |h_b| / 6 + sum_{a<b, a~b} |h_a h_b| sum_{c>a, c~PaPb}
    |h_c| / 3, exactly, where ~ is anticommutation and c~PaPb iff c anticommutes with exactly
    one of Pa, Pb: the double-commutator sums of the second-order Trotter remainder."""
    h = [Fraction(abs(t[4])) for t in terms]
    n = len(terms)
    adj = [[b for b in range(n) if b != a and terms[a][4] and terms[b][4]
            and anticommute(terms[a], terms[b])] for a in range(n)]
    C = Fraction(0)
    for a in range(n):
        C += h[a] * h[a] * sum((h[b] for b in adj[a] if b > a), Fraction(0)) / 6
        for b in adj[a]:
            if b <= a:
                continue
            one = set(adj[a]) ^ set(adj[b])
            C += h[a] * h[b] * sum((h[c] for c in one if c > a), Fraction(0)) / 3
    return C


@pytest.mark.parametrize("k", range(12))
def test_strang_bound_exact_constants(oracle, k):
    """The bound is the smallest of the Taylor form (2/3)(|dt| sum|h|)^3 and the nested form
    2 C |dt|^3, each rounded upward through chains of upward additions (one per anticommuting
    pair and triple, ~1e4 at N=27, each adding at most an ulp): so it is at or above the exact
    minimum, and within a relative 1e-9 of it.  A wrong 1/6, 1/3 or 2/3 (or a dropped
    anticommuting pair) moves it by percents."""
    rng = random.Random(1000 + k)
    if k < 5:
        _, terms, dt = five_configs()[k]
    else:
        terms = random_terms(rng, rng.randrange(3, 9), rng.randrange(2, 30))
        dt = rng.choice([0.01, 0.05, -0.2])
    t = abs(Fraction(dt))
    lam = sum((Fraction(abs(x[4])) for x in terms), Fraction(0))
    taylor = Fraction(2, 3) * (lam * t) ** 3
    nested = 2 * nested_constant(terms) * t ** 3
    exact = min(taylor, nested)
    ta = OB.terms_array(terms)
    got = Fraction(oracle.strang_bound(ta, len(terms), dt))
    assert got >= exact
    assert got <= exact * (1 + Fraction(1, 10 ** 9)) + Fraction(1, 10 ** 300)
    gt = Fraction(oracle.strang_bound_taylor(ta, len(terms), dt))
    assert taylor <= gt <= taylor * (1 + Fraction(1, 10 ** 12))


# ---- bound_add_up / bound_fid_up ---------------------------------------------------------
def special_values(rng, count):
    vals = [0.0, TRUE_MIN, DBL_MIN, 1.0, 0.1, 3.0, 1e-300, 1e300, DBL_MAX, math.inf]
    vals += [rng.uniform(0, 1) * 10.0 ** rng.randrange(-30, 30) for _ in range(count)]
    return vals


def test_bound_add_up_bitwise_vs_reference(oracle, reference):
    rng = random.Random(7)
    vals = special_values(rng, 40) + [-1.0, math.nan]
    for a in vals:
        for b in vals:
            assert same(oracle.bound_add_up(a, b), reference.bound_add_up(a, b)), (a, b)


def test_bound_fid_up_bitwise_vs_reference(oracle, reference):
    rng = random.Random(8)
    vals = special_values(rng, 60) + [-1.0, math.nan]
    for a in vals:
        for n in (0, 1, 2, 3, 12, 16, 27, 48, 64, 1000, 2 ** 40 + 1):
            assert same(oracle.bound_fid_up(a, n), reference.bound_fid_up(a, n)), (a, n)


def next_down(x):
    return math.nextafter(x, -math.inf)


def test_bound_add_up_is_the_smallest_double_above(oracle):
    """With MPFR the sum is taken at 256 bits upward and rounded upward to binary64: for two
    doubles whose exact sum is representable at 256 bits, that is the smallest double >= a + b.
    Without MPFR the reference falls back to next_up(RN(a + b)), which is >= a + b."""
    rng = random.Random(9)
    # finite values below DBL_MAX (the sum stays finite)
    vals = [x for x in special_values(rng, 40) if x < DBL_MAX]
    for a in vals:
        for b in vals:
            z = oracle.bound_add_up(a, b)
            exact = Fraction(a) + Fraction(b)
            assert Fraction(z) >= exact
            if oracle.has_mpfr():
                assert Fraction(next_down(z)) < exact, (a, b, z)


def test_bound_fid_up_is_the_smallest_double_above(oracle):
    """a / sqrt(n) with sqrt rounded down and the quotient rounded up at 256 bits, then upward to
    binary64: the smallest double z with z^2 n >= a^2 (exact integer comparison), except where
    a / sqrt(n) lies within 2^-250 relative of a double, which these inputs avoid."""
    if not oracle.has_mpfr():
        pytest.skip("bound_fid_up needs MPFR")
    rng = random.Random(10)
    for a in [x for x in special_values(rng, 60) if x < math.inf]:
        for n in (1, 2, 3, 12, 16, 27, 48, 64, 1000):
            z = oracle.bound_fid_up(a, n)
            A = Fraction(a)
            assert Fraction(z) ** 2 * n >= A ** 2
            if a > 0 and math.isqrt(n) ** 2 != n:
                assert Fraction(next_down(z)) ** 2 * n < A ** 2, (a, n, z)
            if a > 0 and math.isqrt(n) ** 2 == n:  # exact quotient: the smallest double above
                assert Fraction(next_down(z)) < A / math.isqrt(n)


# ---- map_l2 ------------------------------------------------------------------------------
def random_maps(rng, count):
    for k in range(count):
        n = rng.choice([6, 16, 33, 64])
        w = (n + 31) // 32
        size = rng.randrange(1, 400)
        keys = set()
        while len(keys) < size:
            key = tuple(rng.getrandbits(min(64, 2 * (n - 32 * q))) for q in range(w))
            keys.add(key)
        keys = sorted(keys, key=lambda t: t[::-1])
        scale = 10.0 ** rng.randrange(-150, 150)
        vals = [rng.uniform(-1, 1) * scale for _ in keys]
        yield n, np.array(keys, np.uint64).reshape(size, w), np.array(vals)


def test_map_l2_bitwise_vs_reference_and_exact(oracle, reference):
    rng = random.Random(11)
    for n, keys, vals in random_maps(rng, 40):
        a = oracle.new_map(n)
        b = reference.new_map(n)
        assert a.assign(keys, vals) == 0 and b.assign(keys, vals) == 0
        la, lb = a.l2(), b.l2()
        assert same(la, lb), (n, la, lb)
        # long-double accumulation, then a rounded square root: within 2 ulp of the exact norm
        exact2 = sum((Fraction(v) ** 2 for v in vals), Fraction(0))
        up, dn = math.nextafter(la, math.inf), next_down(la)
        assert Fraction(math.nextafter(dn, -math.inf)) ** 2 <= exact2 <= \
            Fraction(math.nextafter(up, math.inf)) ** 2


def test_map_l2_on_trajectories(oracle, reference):
    for name, terms, dt in five_configs()[:2]:
        n = 12 if name == "ring12" else 16
        a, b = oracle.new_map(n), reference.new_map(n)
        for s in range(n):
            a.add_single(s, 1, 1.0)
            b.add_single(s, 1, 1.0)
        ta = OB.terms_array(terms)
        oa, ob = oracle.options(threshold=1e-8), reference.options(threshold=1e-8)
        la, lb = oracle.ledger(0), reference.ledger(0)
        for _ in range(8):
            assert a.strang_step(ta, len(terms), dt, oa, la) == 0
            assert b.strang_step(ta, len(terms), dt, ob, lb) == 0
            assert same(a.l2(), b.l2())


# ---- rotation composition over the edge-value matrix -------------------------------------
@pytest.mark.parametrize("validated", [0, 1])
def test_edge_matrix_single_term_rotations(oracle, reference, validated):
    """Every edge coefficient on every two-site Pauli key, rotated by every generator, as a
    single-term map: status, keys, coefficients and ledger equal the reference's bit for bit.
    A dropped branch, a wrong phase sign, a swapped slot or a missing radius fails here."""
    thetas = (0.3, -0.8, 1e-3)
    n_ok = 0
    for theta in thetas:
        for p in range(16):
            for v in EDGE:
                for qi in range(4):
                    for qj in range(4):
                        res = []
                        for eng in (oracle, reference):
                            m = eng.new_map(2)
                            key = (C.c_uint64 * 1)(p)
                            st = eng.map_add(m.ptr, key, v)  # adding a zero leaves no term
                            o = eng.options(threshold=1e-300, validated=validated,
                                            max_terms=16)
                            led = eng.ledger(validated)
                            if st == 0:
                                st = m.rotate(0, 1, qi, qj, theta, o, led)
                            k, vals = m.arrays()
                            res.append((st, k.tolist(), vals.view(np.uint64).tolist(),
                                        led.bits()))
                        assert res[0] == res[1], (theta, p, v, qi, qj)
                        n_ok += res[0][0] == 0
    assert n_ok > 0

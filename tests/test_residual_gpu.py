"""The independent endpoint (residual) bound with its interval vectors built on the MI355X.

tpgp_residual_bound builds D1 and D2 (residual_rev.c:181-187: four generator applications and
two combinations, 256-bit intervals with directed rounding) with tpgp_hip_residual_vectors and
finishes the serial Gram/norm sums on the host.  Checked here:
1. Bit for bit against the reference library compiled from its own sources (oracle/_ref), both
   variants (incident terms with the Gram enclosure; every term scanned with direct norms),
   depths 0..4, one- and two-word keys, thresholds 0 and positive, on consecutive maps of
   Strang trajectories from the oracle; then on maps of 1e4..1e5 terms (N=16 J1-J2 and N=40),
   where each generator application has 1e5..1e7 contributions.
2. Against the mathematics: for N = 4 the generator L is a 256 x 256 matrix over Pauli strings;
   the exact one-step image exp(dt L) a lies within the bound of the Strang step's result b.
3. Edge cases: no terms (the bound is the endpoint difference alone), terms with h = 0 (the
   scan variant creates entries for them, the incident-terms variant does not), an empty map,
   and coefficients of very different magnitudes (1e-300 .. 1e300) so that sums round.
4. The vectors themselves: D1 of a one-term map against the closed form 6b - 6a - ... built
   from the dense generator, and the C ABI's argument refusals.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB
from test_residual import OK, bits, pauli_index_generator, ref_bound, trajectory

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(N.engine().tpgp_has_mpfr() != 1, reason="built without MPFR")]


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


def same_as_reference(ref_lib, a, b, terms, dt, depth, reference, where):
    st, up = E.residual_bound(a, b, terms, dt, depth, reference)
    rst, rup = ref_bound(ref_lib, a, b, terms, dt, depth, reference)
    assert st == rst == OK, (where, st, rst, N.engine().tpgp_hip_last_error())
    assert bits(up) == bits(rup), (where, up, rup)
    return up


@pytest.mark.parametrize("n,model,dt,thr", [(6, "j1j2", 0.05, 1e-6), (8, "ring", 0.02, 0.0),
                                            (12, "ring", 0.01, 1e-8), (40, "serving", 0.01, 1e-4),
                                            (64, "ring", 0.01, 1e-5)])
def test_bit_identical_to_reference(oracle, ref_lib, n, model, dt, thr):
    terms = (H.j1j2(n, 0.3) if model == "j1j2" else H.ring(n) if model == "ring"
             else H.serving_ring(n, [0.5 + 0.01 * i for i in range(n)]))
    maps = trajectory(oracle, n, terms, dt, 2, thr)
    for k in range(2):
        for depth in (0, 1, 4):
            for reference in (False, True):
                print(f"N={n} {model} step {k} depth {depth} scan={reference}", flush=True)
                same_as_reference(ref_lib, maps[k], maps[k + 1], terms, dt, depth, reference,
                                  (n, k, depth, reference))


@pytest.mark.parametrize("n,model,dt,thr,steps", [(16, "j1j2", 0.005, 1e-6, 100),
                                                  (40, "serving", 0.01, 1e-6, 12)])
def test_large_maps_bit_identical(oracle, ref_lib, n, model, dt, thr, steps):
    """Maps of 1e4..1e5 terms: each generator application has 1e5..1e7 contributions, so the
    device's sort, run placement and chains are exercised at scale (the reference's CPU takes
    seconds here)."""
    terms = (H.j1j2_n16_terms(n) if model == "j1j2"
             else H.serving_ring(n, [0.5 + 0.02 * (i % 7) for i in range(n)]))
    maps = trajectory(oracle, n, terms, dt, steps, thr)
    a, b = maps[-2], maps[-1]
    print(f"N={n} {model}: {len(a)} -> {len(b)} terms", flush=True)
    t0 = time.time()
    st, up = E.residual_bound(a, b, terms, dt, 1)
    t1 = time.time()
    rst, rup = ref_bound(ref_lib, a, b, terms, dt, 1, False)
    t2 = time.time()
    assert st == rst == OK and bits(up) == bits(rup), (up, rup)
    print(f"  bound {up:.6e}: device vectors + host sums {t1 - t0:.2f} s, reference "
          f"{t2 - t1:.2f} s", flush=True)


@pytest.mark.parametrize("dt", [0.02, 0.1, 0.3])
def test_bound_encloses_exact_step(oracle, dt):
    from scipy.linalg import expm

    n = 4
    terms = H.j1j2(n, 0.4)
    maps = trajectory(oracle, n, terms, dt, 3, 0.0)
    L = pauli_index_generator(n, terms)
    U = expm(dt * L)
    for k in range(3):
        a = np.zeros(4 ** n)
        b = np.zeros(4 ** n)
        for vec, mp in ((a, maps[k]), (b, maps[k + 1])):
            keys, vals = mp.arrays()
            vec[keys[:, 0].astype(np.int64)] = vals
        err = np.linalg.norm(b - U @ a)
        for depth in (0, 2):
            st, up = E.residual_bound(maps[k], maps[k + 1], terms, dt, depth)
            assert st == OK
            assert err <= up * (1 + 1e-12), (k, depth, err, up)
        # and it is a real statement: within a modest factor of the true error
        assert up < 1e3 * max(err, 1e-300)


def test_edge_cases(ref_lib):
    n = 6
    terms = H.ring(n)
    a = E.Map(n)
    for s in range(n):
        a.add_single(s, 1, 1.0)
    b = E.Map(n)
    b.add_single(0, 1, 0.5)
    b.add_single(3, 2, -0.25)
    empty = E.Map(n)
    # no terms: the generator vanishes, the bound is the pure endpoint difference
    for depth in (0, 3):
        up = same_as_reference(ref_lib, a, b, [], 0.01, depth, False, ("no terms", depth))
        assert up > 0
    # terms with h = 0 (and -0): entries in the scan variant only
    zero_h = terms[:3] + [(0, 1, 3, 3, 0.0), (1, 2, 1, 1, -0.0)] + terms[3:]
    for reference in (False, True):
        for depth in (0, 2):
            same_as_reference(ref_lib, a, b, zero_h, 0.01, depth, reference, ("h=0", reference))
    # an empty map on either side
    for x, y in ((empty, b), (a, empty), (empty, empty)):
        for reference in (False, True):
            same_as_reference(ref_lib, x, y, terms, 0.01, 1, reference, ("empty", reference))
    # magnitudes far apart: products and sums of the chains round at 256 bits
    wide = E.Map(n)
    rng = np.random.default_rng(5)
    for k in range(300):
        key = int(rng.integers(1, 4 ** n))
        wide.add(np.array([key], dtype=np.uint64), float(10.0 ** rng.uniform(-300, 250)) *
                 (1 if k % 2 else -1))
    wide2 = E.Map(n)
    keys, vals = wide.arrays()
    for k in range(len(vals)):
        wide2.add(keys[k], vals[k] * (1 + 2.0 ** -30))
    for reference in (False, True):
        for depth in (0, 1, 3):
            same_as_reference(ref_lib, wide, wide2, terms, 1e-3, depth, reference,
                              ("wide", reference, depth))


def device_vectors(a, b, terms, dt, incident):
    err = C.create_string_buffer(512)
    ctx = N.hip().tpgp_hip_create(0, err, 512)
    assert ctx, err.value
    out = (N.HipIntervalVector * 2)()
    ta = N.terms_array(terms)
    try:
        st = N.hip().tpgp_hip_residual_vectors(
            ctx, a.m.keys, a.m.values, a.m.len, b.m.keys, b.m.values, b.m.len, a.m.n_words,
            a.m.n_spins, ta, len(terms), dt, incident, out, err, 512)
        assert st == 0, err.value
        res = []
        for d in range(2):
            v = out[d]
            W = a.m.n_words
            keys = np.ctypeslib.as_array(v.keys, (v.len * W,)).copy() if v.len else \
                np.zeros(0, np.uint64)

            def value(x):
                if not x.m[3]:
                    return -0.0 if x.neg else 0.0
                m = (x.m[3] << 192) | (x.m[2] << 128) | (x.m[1] << 64) | x.m[0]
                from fractions import Fraction
                f = Fraction(m, 1 << 256) * (Fraction(2) ** x.e)
                return -f if x.neg else f

            lo = [value(v.lower[i]) for i in range(v.len)]
            hi = [value(v.upper[i]) for i in range(v.len)]
            extra = {"partner": [v.partner[i] for i in range(v.len)],
                     "square": [(value(v.square[0][i]), value(v.square[1][i]))
                                for i in range(v.len)]}
            if d == 0:
                extra["product"] = [(value(v.product[0][i]), value(v.product[1][i]))
                                    for i in range(v.len)]
            res.append((keys.reshape(-1, W), lo, hi, extra))
            N.hip().tpgp_hip_interval_vector_free(C.byref(out[d]))
        return res
    finally:
        N.hip().tpgp_hip_destroy(ctx)


def test_vectors_against_dense_generator():
    """D1 and D2 of small exact maps against 6b - 6a - 4dt La - 2dt Lb - dt^2 LLa (and D2's
    combination) computed exactly in rationals from the dense generator: every interval must
    contain the exact value, and where the chain is exact (dyadic inputs, few terms) the
    interval is the point itself."""
    from fractions import Fraction

    n = 4
    terms = [(0, 1, 3, 3, 0.5), (0, 1, 1, 1, -0.25), (1, 2, 2, 2, -0.25), (2, 3, 3, 3, 0.75)]
    a = E.Map(n)
    b = E.Map(n)
    for s in range(n):
        a.add_single(s, 1, 1.0)
        b.add_single(s, 1, 0.5 + 0.125 * s)
    b.add_single(1, 2, 0.0625)
    dt = 0.125
    L = pauli_index_generator(n, terms)  # entries +-2h: exact in binary64 and in Fraction
    dim = 4 ** n

    def vec(mp):
        v = [Fraction(0)] * dim
        keys, vals = mp.arrays()
        for k, x in zip(keys[:, 0], vals):
            v[int(k)] = Fraction(x)
        return v

    def apply(v):
        out = [Fraction(0)] * dim
        nz = np.nonzero(L)
        for r, c in zip(*nz):
            out[r] += Fraction(L[r, c]) * v[c]
        return out

    va, vb = vec(a), vec(b)
    La, Lb = apply(va), apply(vb)
    LLa, LLb = apply(La), apply(Lb)
    fdt = Fraction(dt)
    d1 = [6 * vb[i] - 6 * va[i] - 4 * fdt * La[i] - 2 * fdt * Lb[i] - fdt * fdt * LLa[i]
          for i in range(dim)]
    d2 = [6 * vb[i] - 6 * va[i] - 2 * fdt * La[i] - 4 * fdt * Lb[i] + fdt * fdt * LLb[i]
          for i in range(dim)]
    for incident in (1, 0):
        got = device_vectors(a, b, terms, dt, incident)
        # the key join and the Gram terms: on point intervals every term is the exact square or
        # product of the entries
        (k1, lo1, _, x1), (k2, lo2, _, x2) = got
        pos2 = {int(k): j for j, k in enumerate(k2[:, 0])}
        pos1 = {int(k): i for i, k in enumerate(k1[:, 0])}
        for i, k in enumerate(k1[:, 0]):
            j = pos2.get(int(k), 0xffffffff)
            assert x1["partner"][i] == j
            assert x1["square"][i] == (lo1[i] * lo1[i], lo1[i] * lo1[i])
            if j != 0xffffffff:
                assert x1["product"][i] == (lo1[i] * lo2[j], lo1[i] * lo2[j])
        for j, k in enumerate(k2[:, 0]):
            assert x2["partner"][j] == pos1.get(int(k), 0xffffffff)
            assert x2["square"][j] == (lo2[j] * lo2[j], lo2[j] * lo2[j])
        for (keys, lo, hi, _), exact in zip(got, (d1, d2)):
            assert len(set(int(k) for k in keys[:, 0])) == len(keys)  # unique entries
            seen = set()
            for k, l, h in zip(keys[:, 0], lo, hi):
                x = exact[int(k)]
                assert l <= x <= h, (int(k), l, x, h)
                assert l == h == x  # dyadic data, short chains: every operation is exact
                seen.add(int(k))
            # every nonzero exact entry is present
            assert all(exact[i] == 0 for i in range(dim) if i not in seen)


def test_vector_refusals():
    err = C.create_string_buffer(512)
    ctx = N.hip().tpgp_hip_create(0, err, 512)
    assert ctx, err.value
    try:
        a = E.Map(4)
        a.add_single(0, 1, 1.0)
        ta = N.terms_array(H.ring(4))
        out = (N.HipIntervalVector * 2)()
        f = N.hip().tpgp_hip_residual_vectors
        cases = [dict(dt=0.0), dict(dt=float("nan")), dict(words=2), dict(spins=65, words=3),
                 dict(incident=2), dict(terms=N.terms_array([(0, 0, 3, 3, 1.0)]), count=1),
                 dict(terms=N.terms_array([(0, 9, 3, 3, 1.0)]), count=1)]
        for c in cases:
            st = f(ctx, a.m.keys, a.m.values, a.m.len, a.m.keys, a.m.values, a.m.len,
                   c.get("words", 1), c.get("spins", 4), c.get("terms", ta), c.get("count", 12),
                   c.get("dt", 0.01), c.get("incident", 1), out, err, 512)
            assert st == 1, (c, st)  # TPGP_HIP_EINVAL
            assert out[0].len == 0 and not out[0].keys
    finally:
        N.hip().tpgp_hip_destroy(ctx)

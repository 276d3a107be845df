"""CPU tests pinning the oracle (oracle/tpgp_oracle.c).

1. Against golden vectors recorded from the reference CPU engine (tests/golden/golden.json,
   written by tests/golden/make_golden.py from engine/src/tpgp_rev.c built from its sources).
2. Against the compiled reference itself, live, where oracle/_ref is built here.
3. Against what the mathematics fixes, independent of both: dense density-matrix FIDs, the
   short-time moment series of the NN ring, the Ising closed form, 4x4 conjugations.
"""
import ctypes as C
import json
import math
import os
import struct

import numpy as np
import pytest

from ddevol_amd import hamiltonians as H
from oracle import bindings as OB

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.json")


def fromhex(h):
    return struct.unpack(">d", bytes.fromhex(h))[0]


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def run_traj(eng, n, terms, dt, steps, threshold, validated=0):
    m = eng.new_map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    o = eng.options(threshold=threshold, validated=validated, max_terms=10 ** 9)
    led = eng.ledger(validated)
    ta = OB.terms_array(terms)
    fids, rads, act, leds = [], [], [], []
    for k in range(steps + 1):
        st, mid, rad = m.readout(1, n)
        fids.append(mid)
        rads.append(rad)
        act.append(len(m))
        if k == steps:
            break
        assert m.strang_step(ta, len(terms), dt, o, led) == 0
        leds.append(led.bits())
    return m, fids, rads, act, leds


def ledger_from_hex(entry):
    f = [np.float64(fromhex(h)).view(np.uint64).item() for h in entry[:4]]
    return tuple(f) + tuple(entry[4:])


def check_against(rec, m, fids, rads, act, leds):
    from tests.golden_util import digest

    assert [struct.pack(">d", x).hex() for x in fids] == rec["fid"]
    assert [struct.pack(">d", x).hex() for x in rads] == rec["radius"]
    assert act == rec["active"]
    assert leds == [ledger_from_hex(e) for e in rec["ledger"]]
    keys, vals = m.arrays()
    assert digest(keys, vals) == rec["final_digest"]


def test_golden_j1j2_n6(oracle, golden):
    for rec in golden["j1j2_n6"]:
        thr = fromhex(rec["threshold"])
        m, fids, rads, act, leds = run_traj(oracle, 6, H.j1j2(6, 0.3), 1.0 / 64, 3, thr,
                                            rec["validated"])
        check_against(rec, m, fids, rads, act, leds)
        keys, vals = m.arrays()
        assert [[int(w) for w in k] for k in keys] == rec["final_keys"]
        assert [struct.pack(">d", v).hex() for v in vals] == rec["final_values"]


@pytest.mark.parametrize("name", ["ring12", "j1j2_n16", "cluster27", "ring48", "range64"])
def test_golden_configs(oracle, golden, name):
    rec = golden[name]
    terms = config_terms(name)
    m, fids, rads, act, leds = run_traj(oracle, rec["n"], terms, fromhex(rec["dt"]), rec["steps"],
                                        fromhex(rec["threshold"]))
    check_against(rec, m, fids, rads, act, leds)


def config_terms(name):
    if name == "ring12":
        return H.ring(12)
    if name == "j1j2_n16":
        return H.j1j2_n16_terms(16)
    if name == "cluster27":
        return H.powder_terms(H.powder_orientations(1)[0])
    if name == "ring48":
        g = H.InputPCG(7)
        return H.serving_ring(48, [0.5 + g.uniform() for _ in range(48)])
    if name == "range64":
        return H.range_ring(64, 4)
    raise KeyError(name)


def test_golden_trig_matches_this_libm(golden):
    """The host libm must reproduce the recorded cos/sin bits (the GPU engine uses host trig)."""
    for name, dt in (("ring12", 0.01), ("j1j2_n16", 0.005)):
        half = dt * 0.5
        for (hh, ch, sh) in golden["trig"][name]:
            h = fromhex(hh)
            x = 2 * h * half
            assert struct.pack(">d", math.cos(x)).hex() == ch
            assert struct.pack(">d", math.sin(x)).hex() == sh


def test_oracle_equals_reference_live(oracle, reference):
    """Bitwise equality with the compiled reference on a validated and a thresholded run."""
    for (thr, val) in ((1e-4, 1), (0.0, 0)):
        a = run_traj(oracle, 8, H.j1j2(8, 0.3), 0.01, 4, thr, val)
        b = run_traj(reference, 8, H.j1j2(8, 0.3), 0.01, 4, thr, val)
        ka, va = a[0].arrays()
        kb, vb = b[0].arrays()
        assert np.array_equal(ka, kb) and np.array_equal(va.view(np.uint64), vb.view(np.uint64))
        assert a[1:] == b[1:]


# ---- mathematics -------------------------------------------------------------------------
PAULI = [np.eye(2), np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]),
         np.array([[1, 0], [0, -1]], complex)]


def dense_op(n, sites_syms):
    op = np.array([[1.0 + 0j]])
    d = dict(sites_syms)
    for s in reversed(range(n)):  # site 0 is the least significant tensor factor
        op = np.kron(op, PAULI[d.get(s, 0)])
    return op


def dense_fid(n, terms, times):
    """Exact infinite-temperature FID F(t) = Tr(X(t) X) / (N 2^N), X = sum_s X_s."""
    Hm = sum(h * dense_op(n, [(i, qi), (j, qj)]) for (i, j, qi, qj, h) in terms)
    X = sum(dense_op(n, [(s, 1)]) for s in range(n))
    w, V = np.linalg.eigh(Hm)
    Xe = V.conj().T @ X @ V
    out = []
    for t in times:
        ph = np.exp(1j * w * t)
        Xt = (ph[:, None] * Xe) * ph.conj()[None, :]  # e^{iHt} X e^{-iHt} in the eigenbasis
        out.append(np.real(np.trace(Xt @ Xe)) / (n * 2 ** n))
    return np.array(out)


@pytest.mark.parametrize("n,model", [(8, "ring"), (6, "j1j2")])
def test_oracle_fid_vs_dense_within_strang_bound(oracle, n, model):
    terms = H.ring(n) if model == "ring" else H.j1j2(n, 0.3)
    dt, steps = 0.02, 15
    _, fids, _, _, _ = run_traj(oracle, n, terms, dt, steps, 0.0)
    exact = dense_fid(n, terms, [k * dt for k in range(steps + 1)])
    per_step = oracle.strang_bound(OB.terms_array(terms), len(terms), dt)
    for k in range(steps + 1):
        # |dF| <= k * per-step operator bound (global initial state: dual norm sqrt(N)/N)
        assert abs(fids[k] - exact[k]) <= k * per_step + 1e-13, (k, fids[k], exact[k])
    # the allowance must be meaningful, not vacuous
    assert steps * per_step < 1e-2


def test_dense_reference_matches_ring_moments():
    """Short-time expansion with the exact NN-ring moments M2=9/2, M4=135/4, M6=1161/4
    (engine/src/ring_moments.cpp:27-28; engine/tests/check_moments_rev.c:63-65)."""
    n = 8
    t = np.array([0.01, 0.02, 0.05])
    exact = dense_fid(n, H.ring(n), t)
    series = 1 - 4.5 * t ** 2 / 2 + (135 / 4) * t ** 4 / 24 - (1161 / 4) * t ** 6 / 720
    assert np.all(np.abs(exact - series) < 2e-9 * (t / 0.05) ** 8 + 1e-15)


def test_j1j2_second_moment():
    """M2 = (9/(2N)) sum_bonds d^2 = 981/200 at J2 = .3 (check_moments_rev.c:66)."""
    n = 8
    t = 1e-3
    exact = dense_fid(n, H.j1j2(n, 0.3), [t])[0]
    m2 = 2 * (1 - exact) / t ** 2
    assert abs(m2 - 981 / 200) < 1e-4


def test_ising_closed_form(oracle):
    """Ising ring N=1000, one Strang step to t=.125: len 4000, F = cos^2(t/2)
    (engine/tests/test_core_rev.c:285-341)."""
    n, t = 1000, 0.125
    m = oracle.new_map(n)
    for s in range(n):
        m.add_single(s, 1, 1.0)
    terms = [(k, (k + 1) % n, 3, 3, 0.25) for k in range(n)]
    o = oracle.options(max_terms=4000, validated=1)
    led = oracle.ledger(1)
    ta = OB.terms_array(terms)
    assert oracle.strang_bound(ta, n, t) == 0
    assert m.strang_step(ta, n, t, o, led) == 0
    assert len(m) == 4000
    assert abs(m.single_sum(1) / n - math.cos(t / 2) ** 2) < 2e-14
    assert led.dropped_l1 == 0 and led.dropped_l2 == 0


def test_two_site_conjugations(oracle):
    """All 256 two-site conjugations against 4x4 complex matrices (test_core_rev.c:36-79)."""
    theta = .371
    o = oracle.options(validated=1)
    for qa in range(4):
        for qb in range(4):
            Q = dense_op(2, [(0, qa), (1, qb)])
            U = math.cos(theta) * np.eye(4) - 1j * math.sin(theta) * Q
            for pa in range(4):
                for pb in range(4):
                    R = U @ dense_op(2, [(0, pa), (1, pb)]) @ U.conj().T
                    m = oracle.new_map(2)
                    m.assign(np.array([[pa | (pb << 2)]], np.uint64), np.array([1.0]))
                    assert m.rotate(0, 1, qa, qb, theta, o, oracle.ledger(1)) == 0
                    keys, vals = m.arrays()
                    got = {int(k[0]): v for k, v in zip(keys, vals)}
                    for a in range(4):
                        for b in range(4):
                            c = np.trace(dense_op(2, [(0, a), (1, b)]).conj().T @ R) / 4
                            assert abs(got.get(a | (b << 2), 0.0) - c.real) < 8e-16


def test_directed_rounding_helpers(oracle):
    up = lambda x: math.nextafter(x, math.inf)  # noqa: E731
    for a, b in [(1.0, 1e-17), (0.1, 0.2), (0.0, 3.0), (5.0, 0.0), (-1.0, 1.0)]:
        want = b if a == 0 else a if b == 0 else up(a + b)
        assert oracle.upadd(a, b) == want
        wantm = 0.0 if a == 0 or b == 0 else up(a * b)
        assert oracle.upmul(a, b) == wantm
    for x in [1.0, 0.0, 5e-324, -2.5, 1e300]:
        assert oracle.round_radius(x) == max(abs(up(x) - x), abs(x - math.nextafter(x, -math.inf)))
    assert oracle.round_radius(math.inf) == math.inf


def test_strang_bound_taylor_closed_form(oracle):
    """Taylor fallback = (2/3)(|dt| sum|h|)^3 = 1/768 (test_core_rev.c:274-284)."""
    terms = [(0, 1, 3, 3, .5), (1, 2, 1, 1, .25), (0, 2, 2, 2, .25)]
    ta = OB.terms_array(terms)
    f = oracle.strang_bound_taylor(ta, 3, .125)
    assert 1 / 768 <= f < (1 / 768) * (1 + 3e-15)
    assert oracle.strang_bound(ta, 3, .125) <= f
    assert oracle.strang_bound_taylor(ta, 3, -.125) == f


def test_pauli_phase_table(oracle):
    """phase[q][p] of tpgp_rev.c:414-418, from the Pauli algebra QP = i^phase (QxP)."""
    for q in range(4):
        for p in range(4):
            prod = PAULI[q] @ PAULI[p]
            r = q ^ p
            for ph in range(4):
                if np.allclose(prod, (1j ** ph) * PAULI[r]):
                    break
            assert oracle.pauli_phase(q, p) == ph


def test_readout_single_term(oracle):
    """One X term over normalizer 3: midpoint 1/3 with a tiny positive radius
    (test_core_rev.c:263-269)."""
    if not oracle.has_mpfr():
        pytest.skip("MPFR readout unavailable")
    m = oracle.new_map(2)
    m.add_single(0, 1, 1.0)
    st, mid, rad = m.readout(1, 3)
    assert st == 0 and mid == 1.0 / 3.0 and 0 < rad < 2e-17

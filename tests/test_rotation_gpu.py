"""Rotation and Strang-step parity of the GPU engine against the oracle, bitwise.

The GPU path is include/tpgp_engine.h (C17 host) over include/tpgp_hip.h (HIP kernels): maps,
merged coefficients and the (exact-mode) ledger must equal the oracle's restatement of
engine/src/tpgp_rev.c:535-809 bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB

pytestmark = pytest.mark.gpu


def random_map(rng, n, count, max_weight=5, tiny_fraction=0.2):
    """Sorted unique random Pauli strings with low weight (so partners collide) plus identity."""
    words = (n + 31) // 32
    if words == 1 and count > 100000:
        return random_map_one_word(rng, n, count, max_weight, tiny_fraction)
    keys = set()
    keys.add((0,) * words)
    while len(keys) < count:
        w = int(rng.integers(1, max_weight + 1))
        sites = [int(x) for x in rng.choice(n, size=min(w, n), replace=False)]
        k = [0] * words
        for s in sites:
            k[s // 32] |= int(rng.integers(1, 4)) << (2 * (s % 32))
        keys.add(tuple(k))
    keys = sorted(keys, key=lambda k: tuple(reversed(k)))
    vals = rng.standard_normal(len(keys))
    tiny = rng.random(len(keys)) < tiny_fraction
    vals[tiny] *= 1e-4
    return np.array(keys, np.uint64).reshape(len(keys), words), vals


def random_map_one_word(rng, n, count, max_weight, tiny_fraction):
    """random_map for large one-word maps, drawn in vectorised batches: each candidate keeps
    each site with probability max_weight / (2n), a random non-identity Pauli on kept sites."""
    keys = np.zeros(1, np.uint64)
    while len(keys) < count:
        live = rng.random((count, n)) < max_weight / (2 * n)
        ops = rng.integers(1, 4, size=(count, n), dtype=np.uint64) * live
        cand = np.bitwise_or.reduce(ops << (np.uint64(2) * np.arange(n, dtype=np.uint64)), axis=1)
        keys = np.unique(np.concatenate([keys, cand]))
    keys = np.sort(rng.choice(keys[1:], size=count - 1, replace=False))
    keys = np.concatenate([np.zeros(1, np.uint64), keys])
    vals = rng.standard_normal(count)
    tiny = rng.random(count) < tiny_fraction
    vals[tiny] *= 1e-4
    return keys.reshape(count, 1), vals


def both_maps(oracle, n, keys, vals):
    om = oracle.new_map(n)
    assert om.assign(keys, vals) == 0
    gm = E.Map.from_arrays(n, keys, vals)
    return om, gm


def assert_same_map(om, gm):
    ok, ov = om.arrays()
    gk, gv = gm.arrays()
    assert ok.shape == gk.shape, (ok.shape, gk.shape)
    assert np.array_equal(ok, gk)
    assert np.array_equal(ov.view(np.uint64), gv.view(np.uint64))


@pytest.mark.parametrize("n", [2, 5, 16, 31, 32, 33, 48, 64])
def test_rotate_random_maps(hip_ctx, oracle, n):
    rng = np.random.default_rng(1000 + n)
    for trial in range(6):
        keys, vals = random_map(rng, n, int(rng.integers(1, 400)) if n > 2 else 10)
        i, j = (int(x) for x in rng.choice(n, size=2, replace=False))
        qi, qj = int(rng.integers(0, 4)), int(rng.integers(0, 4))
        thr = [0.0, 1e-5, 3e-4][trial % 3]
        validated = trial % 2
        theta = float(rng.uniform(-1, 1))
        om, gm = both_maps(oracle, n, keys, vals)
        oo = oracle.options(threshold=thr, validated=validated)
        go = E.options(threshold=thr, validated=validated, context=hip_ctx)
        ol, gl = oracle.ledger(validated), E.ledger(validated)
        so = om.rotate(i, j, qi, qj, theta, oo, ol)
        sg = E.rotate(gm, i, j, qi, qj, theta, go, gl)
        assert so == sg == 0, (so, sg)
        assert_same_map(om, gm)
        assert ol.bits() == gl.bits(), (n, trial, ol.as_tuple(), gl.bits())


@pytest.mark.parametrize("threshold,validated", [(0.0, 0), (1e-5, 0), (0.0, 1), (1e-5, 1)])
def test_strang_j1j2_n6(hip_ctx, oracle, threshold, validated):
    """engine/tests/test_cuda_trajectory_rev.c:27-89: J1-J2 N=6, dt 1/64, 3 steps."""
    n = 6
    terms = H.j1j2(n, 0.3)
    assert len(terms) == 36
    om = oracle.new_map(n)
    gm = E.Map(n)
    for s in range(n):
        assert om.add_single(s, 1, 1.0) == 0
        assert gm.add_single(s, 1, 1.0) == 0
    oo = oracle.options(threshold=threshold, validated=validated, max_terms=100000)
    go = E.options(threshold=threshold, validated=validated, max_terms=100000, context=hip_ctx)
    ol, gl = oracle.ledger(validated), E.ledger(validated)
    ot, gt = OB.terms_array(terms), N.terms_array(terms)
    for step in range(3):
        assert om.strang_step(ot, len(terms), 1.0 / 64, oo, ol) == 0
        assert E.strang_step(gm, gt, len(terms), 1.0 / 64, go, gl) == 0
        assert_same_map(om, gm)
        assert ol.bits() == gl.bits()
        so, fo, ro = om.readout(1, n)
        sg, fg, rg = gm.readout(1, n)
        assert so == sg
        assert np.float64(fo).view(np.uint64) == np.float64(fg).view(np.uint64)
        assert np.float64(ro).view(np.uint64) == np.float64(rg).view(np.uint64)


def test_rollback_and_ledger(hip_ctx, oracle):
    """ECAP and EBUDGET are transactional (engine/tests/test_core_rev.c:80-112)."""
    validated = 1
    gm = E.Map(2)
    gm.add_single(0, 1, 1.0)
    k0, v0 = gm.arrays()
    go = E.options(validated=validated, max_terms=1, context=hip_ctx)
    gl = E.ledger(validated)
    before = gl.bits()
    assert E.rotate(gm, 0, 1, 3, 3, .2, go, gl) == N.ECAP
    k1, v1 = gm.arrays()
    assert np.array_equal(k0, k1) and np.array_equal(v0, v1) and gl.bits() == before
    go = E.options(validated=validated, threshold=.03, context=hip_ctx)
    assert E.rotate(gm, 0, 1, 3, 3, .01, go, gl) == 0
    assert E.rotate(gm, 0, 1, 3, 3, .01, go, gl) == 0
    assert len(gm) == 1
    assert gl.removal_events == 2
    assert gl.dropped_l2 >= .03998
    assert gl.dropped_l2 > np.sqrt(gl.dropped_l2_squared) * 1.4
    assert gl.dropped_l2 >= gl.dropped_l1 and gl.dropped_l2 - gl.dropped_l1 < 1e-15
    k0, v0 = gm.arrays()
    before = gl.bits()
    go = E.options(validated=validated, threshold=.03, max_drop_l2=gl.dropped_l2,
                   context=hip_ctx)
    assert E.rotate(gm, 0, 1, 3, 3, .01, go, gl) == N.EBUDGET
    k1, v1 = gm.arrays()
    assert np.array_equal(k0, k1) and np.array_equal(v0, v1) and gl.bits() == before


def test_local_dense_conjugations(hip_ctx):
    """All 256 two-site conjugations vs independent 4x4 complex matrices (test_core_rev.c:36-79)."""
    P = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]),
         np.array([[1, 0], [0, -1]])]

    def pauli(a, b):  # site 0 is the low bit of the basis index (test_core_rev.c:21-27)
        return np.kron(P[b], P[a])

    theta = .371
    go = E.options(validated=1, context=hip_ctx)
    for qa in range(4):
        for qb in range(4):
            Q = pauli(qa, qb)
            U = np.cos(theta) * np.eye(4) - 1j * np.sin(theta) * Q
            for pa in range(4):
                for pb in range(4):
                    R = U @ pauli(pa, pb) @ U.conj().T
                    gm = E.Map(2)
                    assert gm.add(np.array([pa | (pb << 2)], np.uint64), 1.0) == 0
                    gl = E.ledger(1)
                    assert E.rotate(gm, 0, 1, qa, qb, theta, go, gl) == 0
                    keys, vals = gm.arrays()
                    got = {int(k[0]): v for k, v in zip(keys, vals)}
                    for a in range(4):
                        for b in range(4):
                            coef = np.trace(pauli(a, b).conj().T @ R) / 4
                            assert abs(got.get(a | (b << 2), 0.0) - coef.real) < 8e-16
                            assert abs(coef.imag) < 1e-15
                    assert abs(np.sqrt(np.sum(vals ** 2)) - 1) < 5e-16


@pytest.mark.parametrize("exact", [True, False])
def test_evolve_ring12(hip_ctx, oracle, exact):
    """Resident evolution: N=12 ring, dt .01, 20 steps, threshold 1e-8 vs oracle steps."""
    n, steps, dt, thr = 12, 20, 0.01, 1e-8
    terms = H.ring(n)
    om = oracle.new_map(n)
    gm = E.Map(n)
    for s in range(n):
        om.add_single(s, 1, 1.0)
        gm.add_single(s, 1, 1.0)
    oo = oracle.options(threshold=thr, max_terms=10 ** 8)
    ol = oracle.ledger(0)
    ot = OB.terms_array(terms)
    fid_o = []
    for k in range(steps + 1):
        st, mid, rad = om.readout(1, n)
        fid_o.append(mid)
        if k < steps:
            assert om.strang_step(ot, len(terms), dt, oo, ol) == 0
    go = E.options(threshold=thr, max_terms=10 ** 8, context=hip_ctx)
    gl = E.ledger(0)
    st, fid, rad, act, acc = E.evolve(gm, terms, dt, steps, go, gl, exact_ledger=exact)
    assert st == 0 and acc == steps
    assert np.array_equal(np.array(fid_o).view(np.uint64), fid.view(np.uint64))
    assert_same_map(om, gm)
    if exact:
        assert ol.bits() == gl.bits()
    else:
        # device ledger: same counts; sums are rigorous upward-rounded bounds within 1e-12 rel.
        assert (ol.rotations, ol.removal_events, ol.peak_terms) == \
            (gl.rotations, gl.removal_events, gl.peak_terms)
        for a, b in [(ol.dropped_l1, gl.dropped_l1), (ol.dropped_l2, gl.dropped_l2),
                     (ol.dropped_l2_squared, gl.dropped_l2_squared)]:
            assert b >= a * (1 - 1e-12) and abs(b - a) <= 1e-9 * max(a, 1e-300)


@pytest.mark.parametrize("n,count", [(16, 60000), (40, 30000), (16, 600000)])
def test_rotate_large_multi_tile(hip_ctx, oracle, n, count):
    """Classes spanning many merge tiles and partition tiles (regression: tile splits).  At
    600000 terms a map has several hundred tiles: the one-launch scan walks several chunks."""
    rng = np.random.default_rng(77 + n)
    keys, vals = random_map(rng, n, count, max_weight=7)
    for (i, j, qi, qj) in [(0, 1, 1, 1), (n - 1, 3, 2, 2), (7, n // 2, 3, 3), (2, n - 2, 1, 0)]:
        om, gm = both_maps(oracle, n, keys, vals)
        oo = oracle.options(threshold=1e-5)
        go = E.options(threshold=1e-5, context=hip_ctx)
        ol, gl = oracle.ledger(0), E.ledger(0)
        assert om.rotate(i, j, qi, qj, .377, oo, ol) == 0
        assert E.rotate(gm, i, j, qi, qj, .377, go, gl) == 0
        assert_same_map(om, gm)
        assert ol.bits() == gl.bits()


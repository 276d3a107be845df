"""The small-map kernel with the map resident in LDS, bit for bit against the oracle.

A map of at most 4096 terms (one key word) or 2816 (two words) runs a whole call's partitions
and rotations in one workgroup's LDS; a map that starts larger runs them on global memory, and
one that outgrows LDS during the call writes the input of the pass that would overflow back to
its global buffer and finishes there.  Each case below forces the small-map path
(TPGP_HIP_SMALL_LIMIT=1e8) and runs whole Strang steps (tpgp_strang_step: one call, exact
ledger), validated and not, from random low-weight maps sized to land before, across and past
the LDS capacity.  Maps, coefficients and ledgers must equal the oracle's restatement of
engine/src/tpgp_rev.c:535-809 bit for bit, and the resident loop's FID samples too.
"""
import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB

pytestmark = pytest.mark.gpu


def low_weight_map(rng, n, count, max_weight=4):
    words = (n + 31) // 32
    keys = {(0,) * words}
    while len(keys) < count:
        w = int(rng.integers(1, max_weight + 1))
        k = [0] * words
        for s in rng.choice(n, size=w, replace=False):
            k[int(s) // 32] |= int(rng.integers(1, 4)) << (2 * (int(s) % 32))
        keys.add(tuple(k))
    keys = sorted(keys, key=lambda k: tuple(reversed(k)))
    vals = rng.standard_normal(len(keys))
    vals[rng.random(len(keys)) < 0.2] *= 1e-3  # some terms sit near the threshold
    return np.array(keys, np.uint64).reshape(len(keys), words), vals


def same_map(om, gm):
    ok, ov = om.arrays()
    gk, gv = gm.arrays()
    return ok.shape == gk.shape and np.array_equal(ok, gk) and \
        np.array_equal(ov.view(np.uint64), gv.view(np.uint64))


@pytest.mark.parametrize("n,count", [(16, 1), (16, 40), (16, 1500), (16, 3600), (16, 4096),
                                     (16, 5000), (40, 1), (40, 1200), (40, 2500), (40, 3000)])
@pytest.mark.parametrize("validated", [0, 1])
def test_strang_steps_small_path(hip_ctx, oracle, monkeypatch, n, count, validated):
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", "100000000")
    rng = np.random.default_rng(n * 100003 + count * 7 + validated)
    keys, vals = low_weight_map(rng, n, count)
    terms = H.j1j2(n, 0.3)
    dt, thr = 0.002, 1e-3  # a map of 3600 terms doubles in the first step, then holds
    om = oracle.new_map(n)
    assert om.assign(keys, vals) == 0
    gm = E.Map.from_sorted(n, keys, vals)
    oo = oracle.options(threshold=thr, validated=validated, max_terms=10 ** 8)
    go = E.options(threshold=thr, validated=validated, max_terms=10 ** 8, context=hip_ctx)
    ol, gl = oracle.ledger(validated), E.ledger(validated)
    ot, gt = OB.terms_array(terms), N.terms_array(terms)
    for step in range(2):
        print(f"N={n} {count} terms validated={validated}: step {step + 1}, {len(om)} terms",
              flush=True)
        assert om.strang_step(ot, len(terms), dt, oo, ol) == 0
        st = E.strang_step(gm, gt, len(terms), dt, go, gl)
        assert st == 0, N.engine().tpgp_hip_last_error().decode()
        assert same_map(om, gm), (step, len(om), len(gm))
        assert ol.bits() == gl.bits()


@pytest.mark.parametrize("n,terms_fn,dt,thr,steps", [
    (12, lambda: H.ring(12), 0.01, 1e-8, 30),
    (16, lambda: H.j1j2_n16_terms(16), 0.005, 1e-6, 48),
    (40, lambda: H.j1j2(40, 0.3), 0.01, 1e-5, 14),
])
def test_resident_evolution_small_path(hip_ctx, oracle, monkeypatch, n, terms_fn, dt, thr,
                                       steps):
    """The resident loop from sum_s X_s on the small-map path: FID samples, lengths, final map
    and exact ledger equal the oracle's step by step (each map grows through the LDS capacity
    during its run: 4096 terms at one key word, 2816 at two)."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", "100000000")
    terms = terms_fn()
    om = oracle.new_map(n)
    for s in range(n):
        om.add_single(s, 1, 1.0)
    oo, ol = oracle.options(threshold=thr, max_terms=10 ** 8), oracle.ledger(0)
    ot = OB.terms_array(terms)
    fid_o, act_o = [], []
    for k in range(steps + 1):
        fid_o.append(om.readout(1, n)[1])
        act_o.append(len(om))
        if k < steps:
            assert om.strang_step(ot, len(terms), dt, oo, ol) == 0
    gm = E.Map(n)
    for s in range(n):
        gm.add_single(s, 1, 1.0)
    gl = E.ledger(0)
    st, fid, rad, act, acc = E.evolve(gm, terms, dt, steps,
                                      E.options(threshold=thr, max_terms=10 ** 8,
                                                context=hip_ctx), gl, exact_ledger=True)
    print(f"N={n}: {steps} steps, {act_o[-1]} terms at the end", flush=True)
    assert st == 0 and acc == steps
    assert np.array_equal(np.array(fid_o).view(np.uint64), fid.view(np.uint64))
    assert [int(a) for a in act] == act_o
    assert same_map(om, gm)
    assert ol.bits() == gl.bits()

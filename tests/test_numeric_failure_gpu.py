"""The resident path's ENUMERIC branch: a merged coefficient (or its radius) that is not finite
fails the rotation, and nothing is committed.

The reference checks every raw branch value and radius (engine/src/tpgp_rev.c:682-688) and every
merged sum as it is formed, and replaces the map and ledger only when the whole rotation
succeeded (tpgp_rev.c:754-757); a Strang step is transactional as a whole (tpgp_rev.c:776-809).
On the device a rotation's merge sets a nonfinite flag in its record, and the host fold turns
the first flagged rotation into TPGP_ENUMERIC and rolls the map back.

An overflow needs two addends meeting at one key: an anticommuting key K receives its own
a-slot v_K c and the d-slot +-v_{K xor Q} s of its partner.  With |v| near DBL_MAX and a
rotation angle where c and s are both large, the sum leaves the binary64 range.  Each case is
run through the oracle as well; the statuses must agree, and on failure the GPU map and ledger
must be bit for bit what they were before the call.  Both device paths run every case: the
multi-kernel chain (TPGP_HIP_SMALL_LIMIT=0) and the single-workgroup small-map kernel
(TPGP_HIP_SMALL_LIMIT=1e8).
"""
import math
import sys

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB

pytestmark = pytest.mark.gpu

BIG = 0.75 * sys.float_info.max
PATHS = ["0", "100000000"]


def key_of(symbols):
    """Reference key (one word, N <= 32) of {site: Pauli symbol}."""
    k = 0
    for s, p in symbols.items():
        k |= p << (2 * s)
    return k


def canon(keys, vals):
    order = np.argsort(keys, kind="stable")
    return np.asarray(keys, np.uint64)[order].reshape(-1, 1), np.asarray(vals, float)[order]


def same_map(gm, keys, vals):
    gk, gv = gm.arrays()
    return (gk.shape == keys.shape and np.array_equal(gk, keys)
            and np.array_equal(gv.view(np.uint64), np.asarray(vals).view(np.uint64)))


def overflow_pair_map(n, extra, va, vb, rng):
    """A map on n sites holding K = X_0 at va and K xor ZZ_{0,1} = Y_0 Z_1 at vb, plus `extra`
    random low-weight terms with ordinary coefficients (so the failing pair sits inside a map of
    many tiles when extra is large).  The rotation generator is antisymmetric on the pair, so
    with |va| = |vb| = BIG one of the two merged sums is BIG (c + s) in magnitude, whatever the
    signs, and overflows; with vb = 1 both stay finite."""
    keys = {key_of({0: 1}): va, key_of({0: 2, 1: 3}): vb}
    while len(keys) < extra + 2:
        w = int(rng.integers(1, 5))
        sites = rng.choice(np.arange(2, n), size=w, replace=False)
        k = key_of({int(s): int(rng.integers(1, 4)) for s in sites})
        keys.setdefault(k, float(rng.standard_normal()))
    return canon(list(keys), list(keys.values()))


def run_both(oracle, hip_ctx, n, keys, vals, call, validated):
    om = oracle.new_map(n)
    assert om.assign(keys, vals) == 0
    gm = E.Map.from_sorted(n, keys, vals)
    oo = oracle.options(validated=validated, max_terms=10 ** 8)
    go = E.options(validated=validated, max_terms=10 ** 8, context=hip_ctx)
    ol, gl = oracle.ledger(validated), E.ledger(validated)
    so = call("o", om, oo, ol)
    before = gl.bits()
    sg = call("g", gm, go, gl)
    return so, sg, om, ol, gm, gl, before


@pytest.mark.parametrize("limit", PATHS)
@pytest.mark.parametrize("validated", [0, 1])
@pytest.mark.parametrize("extra", [0, 6000])
def test_rotation_overflow_rolls_back(hip_ctx, oracle, monkeypatch, limit, validated, extra):
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", limit)
    n = 16
    theta = math.pi / 8  # c = cos(pi/4), s = sin(pi/4): both addends ~0.53 DBL_MAX
    seen = set()
    for va, vb in ((BIG, BIG), (BIG, -BIG), (-BIG, BIG), (-BIG, -BIG), (BIG, 1.0)):
        rng = np.random.default_rng(5)
        keys, vals = overflow_pair_map(n, extra, va, vb, rng)

        def call(who, m, o, led):
            if who == "o":
                return m.rotate(0, 1, 3, 3, theta, o, led)
            return E.rotate(m, 0, 1, 3, 3, theta, o, led)

        so, sg, om, ol, gm, gl, before = run_both(oracle, hip_ctx, n, keys, vals, call,
                                                  validated)
        print(f"pair ({va:.3g}, {vb:.3g}): oracle {so}, gpu {sg}", flush=True)
        assert so == sg
        seen.add(so)
        if sg == N.ENUMERIC:
            assert same_map(gm, keys, vals), "map changed by a failed rotation"
            assert gl.bits() == before, "ledger changed by a failed rotation"
        else:
            ok, ov = om.arrays()
            assert same_map(gm, ok, ov)
            assert gl.bits() == ol.bits()
    assert seen == {0, N.ENUMERIC}


def huge_maps(n, count, seed):
    """Random low-weight maps mixing ordinary coefficients with a few near DBL_MAX."""
    rng = np.random.default_rng(seed)
    keys = {}
    while len(keys) < count:
        w = int(rng.integers(1, 4))
        sites = rng.choice(n, size=w, replace=False)
        k = key_of({int(s): int(rng.integers(1, 4)) for s in sites})
        huge = rng.random() < 0.3
        keys.setdefault(k, (rng.choice([-1.0, 1.0]) * 0.9999 * sys.float_info.max if huge
                            else float(rng.standard_normal())))
    return canon(list(keys), list(keys.values()))


@pytest.mark.parametrize("limit", PATHS)
def test_strang_step_overflow_is_transactional(hip_ctx, oracle, monkeypatch, limit):
    """A Strang step whose k-th rotation overflows fails as a whole: the map and ledger are those
    from before the step, whichever rotation failed.  Steps that stay finite match the oracle
    bit for bit."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", limit)
    n, dt = 6, 0.3
    terms = H.j1j2(n, 0.3)
    ot, gt = OB.terms_array(terms), N.terms_array(terms)
    outcomes = []
    for seed in range(24):
        keys, vals = huge_maps(n, 40, seed)

        def call(who, m, o, led):
            if who == "o":
                return m.strang_step(ot, len(terms), dt, o, led)
            return E.strang_step(m, gt, len(terms), dt, o, led)

        so, sg, om, ol, gm, gl, before = run_both(oracle, hip_ctx, n, keys, vals, call, seed % 2)
        print(f"seed {seed}: oracle {so}, gpu {sg}", flush=True)
        assert so == sg, seed
        outcomes.append(sg)
        if sg == N.ENUMERIC:
            assert same_map(gm, keys, vals) and gl.bits() == before, seed
        else:
            ok, ov = om.arrays()
            assert same_map(gm, ok, ov) and gl.bits() == ol.bits(), seed
    assert N.ENUMERIC in outcomes and 0 in outcomes


def late_overflow_case(sign):
    """An FID that overflows at its third Strang step.  Sites 2..7 carry a J1-J2 ring from the
    state sum_s X_s; sites 0, 1 carry only H_01 = Z_0 Z_1 (h = 1) acting on the pair K = X_0,
    K' = Y_0 Z_1 with coefficients BIG and sign*BIG.  Each step (dt = 0.1) turns the pair's
    coefficient vector, of norm 1.06 DBL_MAX, by 0.2 rad from the diagonal towards an axis; at
    the third step one coefficient passes DBL_MAX."""
    n = 8
    terms = [(0, 1, 3, 3, 1.0)] + [(i + 2, j + 2, qi, qj, h)
                                   for (i, j, qi, qj, h) in H.j1j2(6, 0.3)]
    d = {key_of({0: 1}): BIG, key_of({0: 2, 1: 3}): sign * BIG}
    for s in range(2, n):
        d[key_of({s: 1})] = 1.0
    keys, vals = canon(list(d), list(d.values()))
    return n, terms, keys, vals


@pytest.mark.parametrize("limit", PATHS)
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_resident_evolution_stops_at_the_failing_step(hip_ctx, oracle, monkeypatch, limit, sign):
    """The resident loop accepts the steps before the one that overflows and returns their map,
    FID samples and exact ledger, equal to the oracle's after the same number of steps."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", limit)
    n, terms, keys, vals = late_overflow_case(sign)
    dt, steps = 0.1, 8
    ot = OB.terms_array(terms)
    om = oracle.new_map(n)
    assert om.assign(keys, vals) == 0
    oo, ol = oracle.options(threshold=1e-8, max_terms=10 ** 8), oracle.ledger(0)
    fid_o, fail_at = [om.readout(1, n)[1]], None
    for k in range(steps):
        st = om.strang_step(ot, len(terms), dt, oo, ol)
        if st:
            assert st == OB.ENUMERIC
            fail_at = k
            break
        fid_o.append(om.readout(1, n)[1])
    assert fail_at == 2
    gm = E.Map.from_sorted(n, keys, vals)
    go, gl = E.options(threshold=1e-8, max_terms=10 ** 8, context=hip_ctx), E.ledger(0)
    st, fid, rad, act, acc = E.evolve(gm, terms, dt, steps, go, gl, exact_ledger=True)
    print(f"gpu status {st}, accepted {acc} steps", flush=True)
    assert st == N.ENUMERIC and acc == fail_at
    assert np.array_equal(np.array(fid_o).view(np.uint64), fid[:acc + 1].view(np.uint64))
    ok, ov = om.arrays()
    assert same_map(gm, ok, ov)
    assert gl.bits() == ol.bits()


def sum_x(n):
    m = E.Map(n)
    for s in range(n):
        assert m.add_single(s, 1, 1.0) == 0
    return m


@pytest.mark.parametrize("limit", PATHS)
def test_fleet_overflow_fails_one_slot(hip_ctx, monkeypatch, limit):
    """In a fleet, the slot whose FID overflows fails alone with ENUMERIC after its accepted
    steps; its neighbours run to the end, bit for bit equal to their alone runs."""
    monkeypatch.setenv("TPGP_HIP_SMALL_LIMIT", limit)
    n, terms, keys, vals = late_overflow_case(1.0)
    dt, steps = 0.1, 6
    fl = E.Fleet(1, 3, 256, context=hip_ctx)
    opts = E.options(threshold=1e-8, max_terms=10 ** 8)
    normal = []
    for slot in (0, 2):
        assert fl.admit(slot, sum_x(n), terms, dt, steps, opts, exact_ledger=True) == 0
        ref, led = sum_x(n), E.ledger(0)
        st, fid, rad, act, acc = E.evolve(ref, terms, dt, steps,
                                          E.options(threshold=1e-8, max_terms=10 ** 8,
                                                    context=hip_ctx), led, exact_ledger=True)
        assert st == 0 and acc == steps
        normal.append((slot, fid, ref, led))
    bad = E.Map.from_sorted(n, keys, vals)
    led_bad = E.ledger(0)
    st, fid_bad, _, _, acc_bad = E.evolve(E.Map.from_sorted(n, keys, vals), terms, dt, steps,
                                          E.options(threshold=1e-8, max_terms=10 ** 8,
                                                    context=hip_ctx), led_bad,
                                          exact_ledger=True)
    assert st == N.ENUMERIC and acc_bad == 2
    assert fl.admit(1, bad, terms, dt, steps, opts, exact_ledger=True) == 0
    for _ in range(1000):
        st, running = fl.step()
        assert st == 0, N.engine().tpgp_hip_last_error().decode()
        if running == 0:
            break
    state, done, status, fled = fl.slot(1)
    assert state == N.SLOT_FAILED and status == N.ENUMERIC and done == acc_bad
    f, _, _ = fl.fid(1, steps + 4)
    assert np.array_equal(f.view(np.uint64), fid_bad[:acc_bad + 1].view(np.uint64))
    assert fled.bits() == led_bad.bits()
    for slot, fid, ref, led in normal:
        state, done, status, fled = fl.slot(slot)
        assert state == N.SLOT_DONE and status == 0
        f, _, _ = fl.fid(slot, steps + 4)
        assert np.array_equal(f.view(np.uint64), fid.view(np.uint64))
        out = E.Map(n)
        assert fl.release(slot, out) == 0
        ak, av = ref.arrays()
        assert same_map(out, ak, av)
        assert fled.bits() == led.bits()
    fl.close()

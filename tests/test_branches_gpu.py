"""tpgp_hip_branches parity: the case set of engine/tests/test_cuda_rev.c:98-182, bitwise.

The expected slots come from the oracle's scalar branch spec (tpgp_rev.c:432-467, identical to
the independent spec of test_cuda_rev.c:27-50).
"""
import ctypes as C
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DBL_TRUE_MIN = 5e-324
DBL_MIN = 2.2250738585072014e-308
DBL_MAX = 1.7976931348623157e308
EDGE = [0.0, -0.0, DBL_TRUE_MIN, -DBL_TRUE_MIN, DBL_MIN, -DBL_MIN, 1.0, -1.0, DBL_MAX * .25,
        -DBL_MAX * .25, float.fromhex("0x1.fffffffffffffp-1"),
        -float.fromhex("0x1.fffffffffffffp-1"), 2.0 ** -1000, -2.0 ** -1000, 2.0 ** 1000,
        -2.0 ** 1000]


def xorshift(state):
    s = state[0]
    s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
    s ^= s >> 7
    s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
    state[0] = s
    return s


def expected(oracle, keys, values, words, i, j, qi, qj, c, s, rc, rs, enclose):
    count = len(values)
    ok = np.zeros(2 * count * words, np.uint64)
    ov = np.zeros(2 * count)
    orad = np.zeros(2 * count)
    k = np.ascontiguousarray(keys, np.uint64).reshape(-1)
    v = np.ascontiguousarray(values, np.float64)
    oracle.branches(k.ctypes.data_as(C.POINTER(C.c_uint64)), v.ctypes.data_as(C.POINTER(C.c_double)),
                    count, words, i, j, qi, qj, c, s, rc, rs, enclose,
                    ok.ctypes.data_as(C.POINTER(C.c_uint64)), ov.ctypes.data_as(C.POINTER(C.c_double)),
                    orad.ctypes.data_as(C.POINTER(C.c_double)))
    return ok.reshape(2 * count, words), ov, orad


def compare(ctx, oracle, keys, values, words, spins, i, j, qi, qj, c, s, rc, rs, enclose):
    st, gk, gv, gr, msg = ctx.branches(keys, values, words, spins, i, j, qi, qj, c, s, rc, rs,
                                       enclose)
    assert st == 0, msg
    ek, ev, er = expected(oracle, keys, values, words, i, j, qi, qj, c, s, rc, rs, enclose)
    assert np.array_equal(gk.reshape(-1), ek.reshape(-1)), (i, j, qi, qj)
    assert np.array_equal(gv.view(np.uint64), ev.view(np.uint64)), (i, j, qi, qj, enclose)
    assert np.array_equal(gr.view(np.uint64), er.view(np.uint64)), (i, j, qi, qj, enclose)


def test_two_site_edge_values(hip_ctx, oracle):
    """8,192 cases: 16 keys x 16 edge values x 16 generators x enclose (test_cuda_rev.c:110-122)."""
    keys = np.array([p for p in range(16) for _ in range(16)], np.uint64)
    values = np.array([EDGE[v] for _ in range(16) for v in range(16)])
    for qi in range(4):
        for qj in range(4):
            for enclose in (0, 1):
                compare(hip_ctx, oracle, keys, values, 1, 2, 0, 1, qi, qj, .6, -.8, 2.0 ** -54,
                        2.0 ** -1050, enclose)


def test_n1000_multiword(hip_ctx, oracle):
    """257 random 32-word keys at boundary site pairs plus exact angles (test_cuda_rev.c:124-140)."""
    COUNT, WORDS = 257, 32
    st = [0x9E3779B97F4A7C15]
    large = np.zeros((COUNT, WORDS), np.uint64)
    coeff = np.zeros(COUNT)
    for k in range(COUNT):
        for w in range(WORDS):
            large[k, w] = xorshift(st)
        large[k, WORDS - 1] &= 0xFFFF
        if k < 16:
            coeff[k] = EDGE[k]
        else:
            r = xorshift(st)
            coeff[k] = np.ldexp((r >> 11) * 2.0 ** -53, (k % 1800) - 900)
    for (i, j) in [(0, 15), (15, 16), (31, 32), (511, 999), (999, 0)]:
        for qi in range(4):
            for qj in range(4):
                compare(hip_ctx, oracle, large, coeff, WORDS, 1000, i, j, qi, qj, -.8, .6,
                        2.0 ** -54, 2.0 ** -54, 1)
    compare(hip_ctx, oracle, large, coeff, WORDS, 1000, 0, 999, 3, 3, 1.0, -0.0, 0, 0, 1)
    compare(hip_ctx, oracle, large, coeff, WORDS, 1000, 0, 999, 3, 3, 0.0, -1.0, 0, 0, 1)


def test_invalid_inputs(hip_ctx):
    """Statuses of test_cuda_rev.c:142-172 (EINVAL=1, ENUMERIC=4)."""
    from ddevol_amd import native as N

    L = N.hip()
    err = C.create_string_buffer(512)
    key = np.array([1], np.uint64)
    val = np.array([1.0])
    ok = np.zeros(2, np.uint64)
    ov = np.zeros(2)
    orad = np.zeros(2)

    def call(ctx=hip_ctx.ptr, k=key, v=val, n=1, w=1, sp=2, i=0, j=1, qi=3, qj=3, c=.6, s=.8,
             rc=0.0, rs=0.0, en=1, outk=None, outv=None, outr=None):
        kp = N.as_u64p(k) if k is not None else None
        vp = N.as_dp(v) if v is not None else None
        okp = N.as_u64p(outk if outk is not None else ok)
        ovp = N.as_dp(outv if outv is not None else ov)
        orp = N.as_dp(outr if outr is not None else orad)
        return L.tpgp_hip_branches(ctx, kp, vp, n, w, sp, i, j, qi, qj, c, s, rc, rs, en, okp,
                                   ovp, orp, err, 512)

    assert call(ctx=None) == N.HIP_EINVAL
    assert call(k=None) == N.HIP_EINVAL
    assert call(n=(1 << 64) - 1) == N.HIP_EINVAL
    assert call(w=2) == N.HIP_EINVAL
    assert call(j=0) == N.HIP_EINVAL
    assert call(qi=4) == N.HIP_EINVAL
    assert call(c=float("nan")) == N.HIP_EINVAL
    assert call(rc=float("inf")) == N.HIP_EINVAL
    assert call(rc=-1.0) == N.HIP_EINVAL
    assert call(en=2) == N.HIP_EINVAL
    assert call(outk=key) == N.HIP_EINVAL  # output aliases the input
    assert L.tpgp_hip_branches(hip_ctx.ptr, None, None, 0, 1, 2, 0, 1, 3, 3, .6, .8, 0.0, 0.0, 1,
                               None, None, None, err, 512) == N.HIP_OK
    assert call(k=np.array([16], np.uint64)) == N.HIP_EINVAL  # padding bits above 2 sites
    assert call(v=np.array([float("nan")])) == N.HIP_ENUMERIC
    assert call(v=np.array([float("inf")])) == N.HIP_ENUMERIC
    assert call(k=np.array([0], np.uint64), v=np.array([float("nan")])) == N.HIP_ENUMERIC
    assert call(v=np.array([DBL_MAX]), c=2.0) == N.HIP_ENUMERIC
    assert call() == N.HIP_OK  # reuse after failure
    assert not L.tpgp_hip_create(-1, err, 512)
    assert not L.tpgp_hip_create(2 ** 31 - 1, err, 512)


def test_two_concurrent_callers(hip_ctx, oracle):
    """Two host threads share one context (test_cuda_rev.c:84-96, 174-180)."""
    failures = []

    def worker(seed):
        st = [seed]
        keys = np.zeros((64, 32), np.uint64)
        for k in range(64):
            for w in range(32):
                keys[k, w] = xorshift(st)
            keys[k, 31] &= 0xFFFF
        values = np.array([(k + 1) / 64 for k in range(64)])
        try:
            for q in (1, 2, 3):
                compare(hip_ctx, oracle, keys, values, 32, 1000, 511, 999, q, q, .6, -.8,
                        2.0 ** -54, 2.0 ** -54, 1)
        except AssertionError as e:
            failures.append(e)

    th = [threading.Thread(target=worker, args=(s,)) for s in (0x431217, 0x187721)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not failures

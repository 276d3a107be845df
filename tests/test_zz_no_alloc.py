"""tpgp_hip_branches_noalloc: the allocation-free branch call equals tpgp_hip_branches.

The new entry point has no counterpart in the reference; its contract is tpgp_hip_branches'
(engine/include/tpgp_cuda_rev.h:43-49) with the device copies held in a fixed per-GPU array.
Checked through the C ABI, bit for bit against the allocating entry point (itself pinned to the
oracle by test_branches_gpu.py): two calls in a row, the second smaller than the first, with the
outputs pre-filled with a poison pattern so a stale or missing write shows; two-word and 32-word
keys; validated radii; the nonfinite and padding statuses; a call past the fixed size is refused
with a message and leaves the entry point working.
"""
import numpy as np
import pytest

from ddevol_amd import native as N

pytestmark = pytest.mark.gpu

POISON = 0x7FF4DEADBEEF0001  # a signalling-NaN bit pattern no branch ever writes
LIMIT = 32 << 20  # TPGP_HIP_NOALLOC_SCRATCH_BYTES (include/tpgp_hip.h)


def random_keys(rng, count, spins):
    words = (spins + 31) // 32
    digits = rng.integers(0, 4, size=(count, spins), dtype=np.uint64)
    keys = np.zeros((count, words), np.uint64)
    for s in range(spins):
        keys[:, s // 32] |= digits[:, s] << np.uint64(2 * (s % 32))
    return keys, words


def both(ctx, keys, values, words, spins, args):
    a = ctx.branches(keys, values, words, spins, *args, fill=POISON)
    b = ctx.branches(keys, values, words, spins, *args, noalloc=True, fill=POISON)
    return a, b


def assert_same(a, b):
    sa, ka, va, ra, ma = a
    sb, kb, vb, rb, mb = b
    assert sa == sb, (sa, sb, ma, mb)
    if sa == N.HIP_OK:
        assert np.array_equal(ka, kb)
        assert np.array_equal(va.view(np.uint64), vb.view(np.uint64))
        assert np.array_equal(ra.view(np.uint64), rb.view(np.uint64))
        assert not np.any(vb.view(np.uint64) == POISON)


def test_noalloc_equals_branches_twice_shrinking(hip_ctx):
    rng = np.random.default_rng(20260711)
    # first call: large two-word map (N=40), validated radii
    keys, words = random_keys(rng, 200000, 40)
    values = rng.standard_normal(len(keys)) * 10.0 ** rng.integers(-30, 30, len(keys))
    print("noalloc call 1: 200000 terms, N=40, enclose", flush=True)
    a, b = both(hip_ctx, keys, values, words, 40, (3, 37, 1, 2, .6, -.8, 2.0 ** -54, 2.0 ** -54, 1))
    assert a[0] == N.HIP_OK, a[4]
    assert_same(a, b)
    # second call: smaller, one-word keys, no radii; nothing of the first call may show
    keys, words = random_keys(rng, 5000, 12)
    values = rng.standard_normal(len(keys))
    print("noalloc call 2: 5000 terms, N=12", flush=True)
    a, b = both(hip_ctx, keys, values, words, 12, (0, 11, 3, 3, .28, .96, 0.0, 0.0, 0))
    assert a[0] == N.HIP_OK, a[4]
    assert_same(a, b)
    # N=1000 keys of 32 words
    keys, words = random_keys(rng, 257, 1000)
    values = rng.standard_normal(len(keys))
    print("noalloc call 3: 257 terms, N=1000", flush=True)
    a, b = both(hip_ctx, keys, values, words, 1000, (999, 0, 2, 1, -.8, .6, 2.0 ** -54, 0.0, 1))
    assert_same(a, b)


def test_noalloc_statuses(hip_ctx):
    keys = np.array([[1]], np.uint64)
    print("noalloc statuses", flush=True)
    for v in (np.nan, np.inf):  # nonfinite input: ENUMERIC from both
        a, b = both(hip_ctx, keys, np.array([v]), 1, 2, (0, 1, 3, 3, .6, .8, 0.0, 0.0, 1))
        assert a[0] == b[0] == N.HIP_ENUMERIC
    a, b = both(hip_ctx, np.array([[16]], np.uint64), np.array([1.0]), 1, 2,
                (0, 1, 3, 3, .6, .8, 0.0, 0.0, 1))  # padding bits set: EINVAL from both
    assert a[0] == b[0] == N.HIP_EINVAL
    a, b = both(hip_ctx, np.zeros((0, 1), np.uint64), np.zeros(0), 1, 2,
                (0, 1, 3, 3, .6, .8, 0.0, 0.0, 1))  # empty batch
    assert a[0] == b[0] == N.HIP_OK


def test_noalloc_refuses_past_its_scratch_then_works(hip_ctx):
    per_term = 24 * 1 + 40
    count = (LIMIT - 256) // per_term + 1  # one term past what the fixed array holds
    keys = np.zeros((count, 1), np.uint64)
    values = np.ones(count)
    print(f"noalloc past its scratch: {count} terms", flush=True)
    st, _, _, _, msg = hip_ctx.branches(keys, values, 1, 2, 0, 1, 3, 3, .6, .8, 0.0, 0.0, 0,
                                        noalloc=True)
    assert st == N.HIP_EINVAL and "scratch" in msg
    count -= 1  # exactly the largest call it takes
    keys = np.ones((count, 1), np.uint64)
    values = np.full(count, 0.5)
    print(f"noalloc at its scratch: {count} terms", flush=True)
    a, b = both(hip_ctx, keys, values, 1, 2, (0, 1, 3, 3, .6, .8, 0.0, 0.0, 0))
    assert a[0] == N.HIP_OK
    assert_same(a, b)

"""TPGPRVCP version-2 checkpoints (tpgp_checkpoint_*): the record is the reference's, byte for byte.

Pins, on the CPU:
* the SHA-256 against the FIPS 180-2 example digests and hashlib over every length 0..300
  (all block-boundary cases);
* records written here and by the reference's checkpoint_rev.c (compiled from its sources into
  oracle/_ref) are byte-identical, manifests included, and each side loads the other's;
* the reference's rejections (checksum, truncation, trailing bytes, broken invariants) are ours;
* tpgp_checkpoint_same_configuration agrees with the reference on every single-field change.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import math
import os
import struct

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import hamiltonians as H
from ddevol_amd import native as N
from oracle import bindings as OB


def sha(data: bytes) -> str:
    out = (C.c_ubyte * 32)()
    buf = C.create_string_buffer(data, len(data))
    N.engine().tpgp_payload_sha256(buf, len(data), out)
    return bytes(out).hex()


def test_sha256_known_answers():
    assert sha(b"abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert sha(b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    assert sha(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") == \
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"
    assert sha(b"a" * 1000000) == \
        "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0"
    rng = np.random.default_rng(5)
    for n in range(301):
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert sha(data) == hashlib.sha256(data).hexdigest(), n


def random_map(n, count, seed):
    rng = np.random.default_rng(seed)
    w = (n + 31) // 32
    digits = rng.integers(0, 4, size=(count, n), dtype=np.uint64)
    keys = np.zeros((count, w), np.uint64)
    for s in range(n):
        keys[:, s // 32] |= digits[:, s] << np.uint64(2 * (s % 32))
    keys = np.unique(keys, axis=0)
    # canonical order: lexicographic with the highest word most significant
    order = np.lexsort(tuple(keys[:, k] for k in range(w)))
    keys = keys[order]
    values = rng.standard_normal(len(keys))
    values[values == 0] = 0.5
    return E.Map.from_sorted(n, keys, values)


def record(n=40, count=300, seed=1, validated=0, **kw):
    mp = random_map(n, count, seed)
    terms = H.j1j2(n, 0.3)
    o = E.options(threshold=1e-6, max_terms=10 ** 6, validated=validated)
    o.backend = 0
    led = E.ledger(validated)
    led.dropped_l1, led.dropped_l2, led.dropped_l2_squared = 1e-9, 3e-10, 9e-20
    led.rotations, led.removal_events, led.peak_terms = 777, 12, 5000
    args = dict(dt=0.01, step=7, total_steps=40, model="j1j2", j2=0.3, pf_bound=2.5e-7,
                initial_global=True)
    args.update(kw)
    return E.Checkpoint.of(mp, terms, o, led, **args)


@pytest.fixture(scope="module")
def ref_lib():
    """The reference's checkpoint functions, compiled from its sources into oracle/_ref."""
    if not os.path.exists(OB.REF_SO):
        pytest.skip("reference library not built (oracle/_ref)")
    L = C.CDLL(OB.REF_SO)
    cp = C.POINTER(N.Checkpoint)
    for name, res, args in [("tpgp_checkpoint_save", C.c_int, [C.c_char_p, cp, C.c_char_p,
                                                               C.c_size_t]),
                            ("tpgp_checkpoint_load", C.c_int, [C.c_char_p, cp, C.c_char_p,
                                                               C.c_size_t]),
                            ("tpgp_checkpoint_free", None, [cp]),
                            ("tpgp_checkpoint_same_configuration", C.c_int, [cp, cp])]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def ref_save(L, ck, path):
    err = C.create_string_buffer(512)
    return L.tpgp_checkpoint_save(str(path).encode(), C.byref(ck.c), err, 512), err.value


def ref_load(L, path):
    rec = N.Checkpoint()
    err = C.create_string_buffer(512)
    st = L.tpgp_checkpoint_load(str(path).encode(), C.byref(rec), err, 512)
    return st, rec


def fields(c):
    """Every serialised field of a record, floats as bit patterns."""
    def b(x):
        return np.float64(x).view(np.uint64).item()
    o, l = c.options, c.ledger
    n = c.map.len * c.map.n_words
    keys = np.ctypeslib.as_array(c.map.keys, (n,)).tolist() if n else []
    vals = [b(c.map.values[k]) for k in range(c.map.len)]
    terms = [(t.i, t.j, t.qi, t.qj, b(t.h), t.layer) for t in
             (c.terms[k] for k in range(c.term_count))]
    return (c.map.n_spins, c.map.n_words, c.map.len, keys, vals, terms, b(c.dt),
            b(c.pf_bound), b(c.residual_bound), b(c.max_fid_bound), c.step, c.total_steps,
            c.residual_depth, c.initial_global, bytes(c.model), b(c.j2), c.linear_size,
            c.resumed_from_step, c.prior_total_steps, c.prior_max_terms, c.stop_reason,
            c.last_status, b(o.threshold), o.max_terms, o.threads, o.backend, o.validated,
            b(o.max_drop_l2), b(l.dropped_l1), b(l.dropped_l2), b(l.dropped_l2_squared),
            b(l.arithmetic_l2), l.rotations, l.removal_events, l.peak_terms,
            l.arithmetic_enclosed)


@pytest.mark.parametrize("n,count,extra", [
    (40, 300, {}),                                    # two key words
    (12, 50, {"residual_depth": 2, "residual_bound": 1e-8, "stop_reason": 2}),
    (32, 0, {"step": 40, "stop_reason": 1}),          # empty map, completed run
    (7, 20, {"model": "edges", "stop_reason": 4, "last_status": 3, "max_fid_bound": 0.0}),
])
def test_round_trip_and_byte_identity_with_reference(tmp_path, ref_lib, n, count, extra):
    ck = record(n, count, seed=n, **extra)
    ours, theirs = tmp_path / "ours.ck", tmp_path / "ref.ck"
    ck.save(ours)
    st, msg = ref_save(ref_lib, ck, theirs)
    assert st == 0, msg
    assert ours.read_bytes() == theirs.read_bytes()
    assert (tmp_path / "ours.ck.manifest.json").read_text() == \
        (tmp_path / "ref.ck.manifest.json").read_text()
    payload = ours.read_bytes()
    assert payload[-32:].hex() == hashlib.sha256(payload[:-32]).hexdigest()
    # each side loads the other's record with every field intact
    back = E.Checkpoint.load(theirs)
    assert fields(back.c) == fields(ck.c)
    st, rec = ref_load(ref_lib, ours)
    assert st == 0
    assert fields(rec) == fields(ck.c)
    ref_lib.tpgp_checkpoint_free(C.byref(rec))
    assert back.same_configuration(ck)


def corrupt_cases(data: bytes):
    flip = bytearray(data)
    flip[100] ^= 1
    yield "bit flip", bytes(flip)
    yield "truncated", data[:-1]
    yield "trailing byte", data + b"\0"
    yield "too short", data[:200]
    # valid checksums over payloads that break the format
    body = bytearray(data[:-32])
    yield "version 3", _resealed(body, 8, (3).to_bytes(8, "little"))
    yield "bad magic", _resealed(body, 0, b"TPGPRVCQ")
    # header fields whose stored value is outside what the field admits, resealed
    for name, offset, value in [("depth 12", 64, 12), ("global 2", 72, 2),
                                ("stop reason 7", 152, 7), ("status 7", 160, 7),
                                ("threads 65537", 184, 65537), ("threads 0", 184, 0),
                                ("backend 5", 192, 5), ("validated 2", 200, 2),
                                ("zero spins", 16, 0), ("one more term", 24, None),
                                ("step past total", 40, 10 ** 6)]:
        if value is None:  # the header's length disagrees with the body
            value = int.from_bytes(body[24:32], "little") + 1
        yield name, _resealed(body, offset, value.to_bytes(8, "little"))
    yield "negative dt", _resealed(body, 56, struct.pack("<d", -0.01))
    yield "model with a capital", _resealed(body, 80, b"J")


def _resealed(body: bytearray, offset, patch):
    b = bytearray(body)
    b[offset:offset + len(patch)] = patch
    return bytes(b) + hashlib.sha256(bytes(b)).digest()


def test_corrupt_records_rejected_like_the_reference(tmp_path, ref_lib):
    ck = record(20, 80, seed=3)
    path = tmp_path / "a.ck"
    ck.save(path)
    data = path.read_bytes()
    # unsorted keys under a valid checksum: swap the first two map entries
    map_at = len(data) - 32 - ck.c.map.len * 16  # one key word + one value per entry
    body = bytearray(data[:-32])
    e0, e1 = body[map_at:map_at + 16], body[map_at + 16:map_at + 32]
    body[map_at:map_at + 32] = e1 + e0
    swapped = bytes(body) + hashlib.sha256(bytes(body)).digest()
    cases = list(corrupt_cases(data)) + [("unsorted keys", swapped)]
    for name, blob in cases:
        bad = tmp_path / "bad.ck"
        bad.write_bytes(blob)
        with pytest.raises(ValueError):
            E.Checkpoint.load(bad)
        st, rec = ref_load(ref_lib, bad)
        assert st != 0, name
    # a rejected load leaves the destination record as it was
    keep = E.Checkpoint.load(path)
    before = fields(keep.c)
    err = C.create_string_buffer(256)
    assert N.engine().tpgp_checkpoint_load(str(tmp_path / "bad.ck").encode(), C.byref(keep.c),
                                           err, 256) != 0
    assert fields(keep.c) == before and err.value
    # invalid records are refused at save by both
    o = E.options(threshold=-1.0)
    o.backend = 0
    bad_ck = E.Checkpoint.of(random_map(8, 10, 2), H.ring(8), o, E.ledger(0), dt=0.01, step=0,
                             total_steps=3)
    with pytest.raises(ValueError):
        bad_ck.save(tmp_path / "x.ck")
    assert ref_save(ref_lib, bad_ck, tmp_path / "y.ck")[0] != 0
    assert not (tmp_path / "x.ck").exists()


def test_same_configuration_matches_reference(ref_lib):
    base = record(16, 30, seed=9)
    variants = [dict(dt=0.02), dict(total_steps=41), dict(model="ring"), dict(j2=0.25),
                dict(linear_size=3), dict(residual_depth=1, residual_bound=0.0),
                dict(initial_global=False), dict(step=9), dict(pf_bound=1.0),
                dict(stop_reason=2)]
    for v in variants:
        other = record(16, 30, seed=9, **v)
        ours = base.same_configuration(other)
        theirs = bool(ref_lib.tpgp_checkpoint_same_configuration(C.byref(base.c),
                                                                 C.byref(other.c)))
        assert ours == theirs, v
    other = record(16, 30, seed=9)
    other.c.options.threshold = 2e-6
    assert base.same_configuration(other) is False
    other = record(16, 30, seed=9)
    other.c.terms[3].h = math.nextafter(other.c.terms[3].h, 1.0)
    assert base.same_configuration(other) is False
    assert not ref_lib.tpgp_checkpoint_same_configuration(C.byref(base.c), C.byref(other.c))

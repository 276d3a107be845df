"""Launch-shape tuning (tpgp_hip_set_tuning / tpgp_hip_get_tuning, TPGP_HIP_TUNING).

The parameters change only how a pass is cut into workgroups, so every value must give the
oracle's keys, values and ledger bit for bit: partitions with tiles of 1024, 2048 and 3072
terms, and count/write grids from far fewer workgroups than tiles to several per tile, on maps
of hundreds of tiles with one- and two-word keys.
"""
import os

import numpy as np
import pytest

from ddevol_amd import engine as E
from ddevol_amd import native as N

from test_rotation_gpu import assert_same_map, both_maps, random_map

pytestmark = pytest.mark.gpu

DEFAULTS = {N.TUNE_PARTITION_ROUNDS: 8, N.TUNE_COUNT_GRID: 3, N.TUNE_WRITE_GRID: 4}


@pytest.fixture(scope="module")
def tuned_ctx():
    """A context on the built-in defaults (the committed table left out)."""
    old = os.environ.get("TPGP_HIP_TUNING")
    os.environ["TPGP_HIP_TUNING"] = ""
    try:
        ctx = E.Context(0)
    finally:
        if old is None:
            os.environ.pop("TPGP_HIP_TUNING", None)
        else:
            os.environ["TPGP_HIP_TUNING"] = old
    yield ctx
    ctx.close()


def test_committed_table_loads():
    """With TPGP_HIP_TUNING unset, a context loads the committed table from beside the library
    (tuning_gfx950.txt in the package directory, above lib/)."""
    old = os.environ.pop("TPGP_HIP_TUNING", None)
    try:
        ctx = E.Context(0)
    finally:
        if old is not None:
            os.environ["TPGP_HIP_TUNING"] = old
    try:
        with open(N.TUNING_TABLE) as f:
            rows = [tuple(int(x) for x in ln.split("#")[0].split())
                    for ln in f if ln.split("#")[0].strip()]
        for w, p, b, v in rows:
            assert ctx.get_tuning(w, p, b) == v
    finally:
        ctx.close()


def set_all(ctx, words, param, value):
    for b in range(N.TUNE_BUCKETS):
        assert ctx.set_tuning(words, param, b, value) == N.HIP_OK


def test_set_get_and_refusals(tuned_ctx):
    ctx = tuned_ctx
    for w in (1, 2):
        for p, d in DEFAULTS.items():
            assert ctx.get_tuning(w, p, 20) == d
    assert ctx.set_tuning(1, N.TUNE_PARTITION_ROUNDS, 20, 12) == N.HIP_OK
    assert ctx.get_tuning(1, N.TUNE_PARTITION_ROUNDS, 20) == 12
    assert ctx.get_tuning(1, N.TUNE_PARTITION_ROUNDS, 19) == 8  # per bucket
    assert ctx.get_tuning(2, N.TUNE_PARTITION_ROUNDS, 20) == 8  # per key width
    assert ctx.set_tuning(1, N.TUNE_PARTITION_ROUNDS, 20, 0) == N.HIP_OK  # back to the default
    assert ctx.get_tuning(1, N.TUNE_PARTITION_ROUNDS, 20) == 8
    bad = [(0, 0, 0, 8), (3, 0, 0, 8), (1, N.TUNE_PARAMS, 0, 1), (1, 0, N.TUNE_BUCKETS, 8),
           (1, 0, 3, 5), (2, 0, 3, 16), (1, 0, 3, 2), (1, 1, 3, 25), (2, 2, 3, 255)]
    for w, p, b, v in bad:
        assert ctx.set_tuning(w, p, b, v) == N.HIP_EINVAL, (w, p, b, v)
    for args in ((1, N.TUNE_PARAMS, 0), (0, 0, 0), (3, 0, 0), (1, 0, N.TUNE_BUCKETS)):
        with pytest.raises(ValueError):
            ctx.get_tuning(*args)
    for w in (1, 2):  # nothing changed by the refusals
        for p, d in DEFAULTS.items():
            assert ctx.get_tuning(w, p, 3) == d


def test_table_from_environment(tmp_path):
    good = tmp_path / "good.txt"
    good.write_text("# comment\n1 0 18 12\n\n1 1 18 6   # trailing comment\n2 2 0 1\n")
    bad = tmp_path / "bad.txt"
    bad.write_text("1 0 18 12\n1 0 18 7\n")
    old_format = tmp_path / "old.txt"
    old_format.write_text("0 18 12\n")
    old = os.environ.get("TPGP_HIP_TUNING")
    try:
        os.environ["TPGP_HIP_TUNING"] = str(good)
        ctx = E.Context(0)
        assert ctx.get_tuning(1, N.TUNE_PARTITION_ROUNDS, 18) == 12
        assert ctx.get_tuning(2, N.TUNE_PARTITION_ROUNDS, 18) == 8
        assert ctx.get_tuning(1, N.TUNE_COUNT_GRID, 18) == 6
        assert ctx.get_tuning(2, N.TUNE_WRITE_GRID, 0) == 1
        assert ctx.get_tuning(1, N.TUNE_WRITE_GRID, 0) == 4
        ctx.close()
        os.environ["TPGP_HIP_TUNING"] = str(bad)
        with pytest.raises(RuntimeError, match="line 2"):
            E.Context(0)
        os.environ["TPGP_HIP_TUNING"] = str(old_format)  # three numbers: refused
        with pytest.raises(RuntimeError, match="line 1"):
            E.Context(0)
        os.environ["TPGP_HIP_TUNING"] = ""  # empty: built-in defaults
        ctx = E.Context(0)
        assert ctx.get_tuning(1, N.TUNE_PARTITION_ROUNDS, 18) == 8
        ctx.close()
    finally:
        if old is None:
            os.environ.pop("TPGP_HIP_TUNING", None)
        else:
            os.environ["TPGP_HIP_TUNING"] = old


SETTINGS = [(N.TUNE_PARTITION_ROUNDS, 4), (N.TUNE_PARTITION_ROUNDS, 12),
            (N.TUNE_COUNT_GRID, 1), (N.TUNE_COUNT_GRID, 24),
            (N.TUNE_WRITE_GRID, 1), (N.TUNE_WRITE_GRID, 24)]


@pytest.mark.parametrize("n,count", [(16, 300000), (40, 40000)])
def test_every_setting_bit_exact(tuned_ctx, oracle, n, count):
    words = (n + 31) // 32
    rng = np.random.default_rng(500 + n)
    keys, vals = random_map(rng, n, count, max_weight=7)
    rotations = [(0, 1, 1, 1), (n - 1, 3, 2, 2), (7, n // 2, 3, 3)]
    expected = []
    for (i, j, qi, qj) in rotations:
        om, _ = both_maps(oracle, n, keys, vals)
        ol = oracle.ledger(0)
        assert om.rotate(i, j, qi, qj, .377, oracle.options(threshold=1e-5), ol) == 0
        expected.append((om, ol.bits()))
    for param, value in SETTINGS:
        print(f"n={n} count={count} {N.TUNE_NAMES[param]}={value}", flush=True)
        set_all(tuned_ctx, words, param, value)
        try:
            for (i, j, qi, qj), (om, obits) in zip(rotations, expected):
                gm = E.Map.from_arrays(n, keys, vals)
                gl = E.ledger(0)
                go = E.options(threshold=1e-5, context=tuned_ctx)
                assert E.rotate(gm, i, j, qi, qj, .377, go, gl) == 0
                assert_same_map(om, gm)
                assert obits == gl.bits()
        finally:
            set_all(tuned_ctx, words, param, 0)

"""The autotuner's table (ddevol_amd/autotune.py) on the CPU: its format, its bucket fill, and
that the committed table holds only what tpgp_hip_set_tuning accepts (include/tpgp_hip.h), since
every context loads it and a bad line fails the create."""
import os

from ddevol_amd import autotune as A
from ddevol_amd import native as N

VALID = {N.TUNE_PARTITION_ROUNDS: {4, 8, 12}, N.TUNE_COUNT_GRID: set(range(1, 25)),
         N.TUNE_WRITE_GRID: set(range(1, 25))}


def parse(path):
    out = {}
    with open(path) as f:
        for no, line in enumerate(f, 1):
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            w, p, b, v = (int(x) for x in line.split())
            assert w in (1, 2) and 0 <= p < N.TUNE_PARAMS and 0 <= b < N.TUNE_BUCKETS, (path, no)
            assert v in VALID[p], (path, no, v)
            out[(w, p, b)] = v
    return out


def test_write_table_fills_buckets(tmp_path):
    one = {14: {0: 4, 1: 3, 2: 4}, 20: {0: 8, 1: 3, 2: 6}, 24: {0: 12, 1: 2, 2: 4}}
    two = {16: {0: 8, 1: 3, 2: 3}}
    path = tmp_path / "t.txt"
    A.write_table(str(path), {1: one, 2: two}, "test device")
    got = parse(path)
    # defaults are not written; every bucket takes the nearest tuned bucket's choice
    assert got[(1, 0, 0)] == 4 and got[(1, 0, 16)] == 4 and (1, 0, 18) not in got
    assert (1, 0, 20) not in got and got[(1, 2, 19)] == 6 and got[(1, 2, 21)] == 6
    assert got[(1, 0, 23)] == 12 and got[(1, 0, 47)] == 12 and got[(1, 1, 47)] == 2
    assert (1, 1, 14) not in got and (1, 2, 24) not in got
    # each key width has its own table
    assert got[(2, 2, 0)] == 3 and got[(2, 2, 47)] == 3 and (2, 0, 16) not in got


def test_committed_table_is_valid():
    assert os.path.exists(N.TUNING_TABLE)
    got = parse(N.TUNING_TABLE)
    assert got, "the committed table sets something"


def test_candidates_are_settable():
    for p, cands in A.CANDIDATES.items():
        assert set(cands) <= VALID[p]
        assert cands[0] == A.DEFAULT[p]

"""The run driver's host logic on the CPU: every command line it must refuse, it refuses as the
reference's driver does (exit 2, nothing run), before any device work."""
from __future__ import annotations

import os
import shutil

import pytest

import cli_golden as G

pytestmark = pytest.mark.skipif(not os.path.exists(G.OURS), reason="driver not built")


def test_help_and_backend_refusals(tmp_path):
    assert G.run(G.OURS, ["--help"], tmp_path / "h.csv")[0] == 0
    for args in (["--backend", "serial"], ["--backend", "openmp"], ["--bogus", "1"],
                 ["--n"], ["--residual-depth", "11"], ["--dt", "nan"], ["--cap", "-5"],
                 ["--threads", "0"], ["--max-fid-bound", "0"]):
        code, _, rows, err = G.run(G.OURS, args, tmp_path / "x.csv")
        assert code == 2 and not rows, (args, code, err)


CASES = [
    # walltime / accuracy stops need a checkpoint path
    (["--walltime", "5"], False),
    (["--max-fid-bound", "1e-3", "--validated"], False),
    # END must be an integer multiple of STEP
    (["--time", "0.105"], True),
    # geometry
    (["--model", "j1j2", "--n", "4"], True),
    (["--model", "cubic", "--linear-size", "2"], True),
    (["--model", "edges", "--n", "16"], True),           # no coupling file
    (["--model", "triangle"], True),
    (["--initial", "somewhere"], True),
    # an accuracy stop needs an enclosed bound (validated mode or the residual bound)
    (["--max-fid-bound", "1e-3", "--checkpoint", "{ck}"], True),
    # the residual bound is for the global initial state only
    (["--residual-depth", "1", "--initial", "local"], True),
]


@pytest.mark.parametrize("args,ref_agrees", CASES)
def test_refusals_match_reference(tmp_path, args, ref_agrees):
    a = G.argv(args, tmp_path / "c.ck")
    code, _, rows, err = G.run(G.OURS, a, tmp_path / "ours.csv")
    assert code == 2 and not rows, (args, code, err)
    if ref_agrees and os.path.exists(G.REF_CLI):
        assert G.run(G.REF_CLI, G.argv(args, tmp_path / "r.ck"), tmp_path / "ref.csv")[0] == 2


def test_restart_refusals(tmp_path):
    """A corrupted record, options that contradict the record, a coupling file that differs
    from the saved Hamiltonian, and overwriting a checkpoint that is not the restart record."""
    ck = tmp_path / "r.ck"
    shutil.copyfile(os.path.join(G.GOLDEN, "restart.0.ck"), ck)
    cases = [["--restart", str(ck), "--n", "12"],
             ["--restart", str(ck), "--threshold", "1e-6"],
             ["--restart", str(ck), "--dt", "0.02"],
             ["--restart", str(ck), "--model", "j1j2"],
             ["--restart", str(ck), "--time", "0.03"],           # shorter than the record
             ["--restart", str(ck), "--cap", "10"],              # smaller cap than the record
             ["--restart", str(ck), "--model", "edges", "--edges",
              os.path.join(G.GOLDEN, "j1j2_n16.edges")],
             ["--n", "10", "--checkpoint", str(ck)]]             # existing record, no restart
    for args in cases:
        code, _, rows, err = G.run(G.OURS, args, tmp_path / "o.csv")
        assert code == 2 and not rows, (args, code, err)
        if os.path.exists(G.REF_CLI):
            assert G.run(G.REF_CLI, args, tmp_path / "r.csv")[0] == 2, args
        for p in (tmp_path / "o.csv", tmp_path / "r.csv"):
            if p.exists():
                p.unlink()
    bad = tmp_path / "bad.ck"
    data = bytearray(ck.read_bytes())
    data[200] ^= 0x10
    bad.write_bytes(bytes(data))
    code, _, _, err = G.run(G.OURS, ["--restart", str(bad)], tmp_path / "b.csv")
    assert code == 2 and "Restart rejected" in err


def test_lock_refuses_second_run(tmp_path):
    import fcntl

    ck = tmp_path / "l.ck"
    with open(str(ck) + ".lock", "w") as f:
        fcntl.flock(f, fcntl.LOCK_EX | fcntl.LOCK_NB)
        code, _, rows, err = G.run(G.OURS, ["--n", "8", "--checkpoint", str(ck)],
                                   tmp_path / "o.csv")
    assert code == 2 and "locked" in err


def test_golden_restart_chain_equals_uninterrupted_run():
    """The reference's own invariant, on its golden runs: a run to step 6 resumed to step 12
    prints rows 6..12 of the run straight to step 12, and a walltime pause resumed to the end
    prints the rows of an uninterrupted run."""
    g = G.golden()
    r = g["restart"]
    assert r[1]["rows"] == r[2]["rows"][6:]
    c = g["cap"]
    assert c[0]["exit"] == 1 and c[1]["exit"] == 0 and c[1]["rows"][0] == c[0]["rows"][-1]

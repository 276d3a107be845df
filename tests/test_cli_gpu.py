"""The run driver on the MI355X against the reference's driver (golden runs in tests/golden/cli).

For every case and every invocation of its chain (restarts included), the driver must give the
reference's exit code, header lines (the backend field names each driver's own device) and CSV
rows: time, fid, active, drop_l1, drop_l2_sum, drop_l2_squared, arithmetic_l2, the two observable
bounds, the residual columns, readout_radius and arithmetic_enclosed character for character;
l2_norm within 2 units in the last place (the reference sums squares in long double in map
order; the device sums them exactly to ~2^-104 and rounds once).  Checkpoint records must be
the reference's byte for byte once the backend field is set to the reference's.  Records cross
over both ways: this driver resumes the reference's records, and the reference resumes ours.
"""
from __future__ import annotations

import os
import shutil

import pytest

import cli_golden as G
from ddevol_amd import engine as E

pytestmark = pytest.mark.gpu

GOLD = G.golden()


def check_rows(ours, ref, where):
    assert len(ours) == len(ref), (where, len(ours), len(ref))
    for k, (a, b) in enumerate(zip(ours, ref)):
        for c in range(14):
            if c == 3:
                assert G.ulps(a[c], b[c]) <= 2, (where, k, "l2_norm", a[c], b[c])
            else:
                assert a[c] == b[c], (where, k, c, a[c], b[c])


def record_as_reference(path, out):
    """Our record with the backend field set to the reference driver's serial backend."""
    ck = E.Checkpoint.load(path)
    ck.c.options.backend = 0
    ck.save(out)
    return out


@pytest.mark.parametrize("case", sorted(GOLD))
def test_driver_matches_reference(tmp_path, case):
    ck = tmp_path / "run.ck"
    for k, inv in enumerate(GOLD[case]):
        print(f"case {case}[{k}]: {' '.join(inv['args'])}", flush=True)
        code, header, rows, err = G.run(G.OURS, G.argv(inv["args"], ck), tmp_path / f"{k}.csv")
        assert code == inv["exit"], (case, k, code, err)
        assert G.same_header(header, inv["header"]), (case, k, header, inv["header"])
        check_rows(rows, inv["rows"], (case, k))
        if inv["checkpoint"]:
            mine = record_as_reference(ck, tmp_path / f"{k}.as_ref.ck")
            want = os.path.join(G.GOLDEN, f"{case}.{k}.ck")
            assert mine.read_bytes() == open(want, "rb").read(), (case, k)


@pytest.mark.parametrize("case,first", [("restart", 0), ("cap", 0), ("walltime", 0),
                                        ("residual_restart", 0)])
def test_driver_resumes_reference_records(tmp_path, case, first):
    """The reference's record after invocation `first`, resumed here, continues exactly as the
    reference's own resumed run."""
    ck = tmp_path / "run.ck"
    shutil.copyfile(os.path.join(G.GOLDEN, f"{case}.{first}.ck"), ck)
    inv = GOLD[case][first + 1]
    code, header, rows, err = G.run(G.OURS, G.argv(inv["args"], ck), tmp_path / "r.csv")
    assert code == inv["exit"], err
    assert G.same_header(header, inv["header"])
    check_rows(rows, inv["rows"], case)
    mine = record_as_reference(ck, tmp_path / "as_ref.ck")
    assert mine.read_bytes() == open(os.path.join(G.GOLDEN, f"{case}.{first + 1}.ck"),
                                     "rb").read()


@pytest.mark.skipif(not os.path.exists(G.REF_ON_ENGINE),
                    reason="reference driver not linked against the engine (make -C oracle ref)")
@pytest.mark.parametrize("case", sorted(GOLD))
def test_reference_driver_on_engine(tmp_path, case):
    """The drop-in of the core API: the reference's own driver (cli_rev.c, compiled from its
    sources without edits) linked against libtpgp_engine.so and libtpgp_cuda_shim.so in place of
    its core library and CUDA stub.  With `--backend cuda` (the slot the GPU occupies) every
    step runs tpgp_strang_step on the MI355X; each invocation of every golden chain must give
    the reference's exit code, header, rows and records, as this repository's driver does."""
    ck = tmp_path / "run.ck"
    for k, inv in enumerate(GOLD[case]):
        args = G.argv(inv["args"], ck)
        if "--restart" not in args:
            args += ["--backend", "cuda"]
        print(f"case {case}[{k}]: {' '.join(args)}", flush=True)
        code, header, rows, err = G.run(G.REF_ON_ENGINE, args, tmp_path / f"{k}.csv")
        assert code == inv["exit"], (case, k, code, err)
        assert G.same_header(header, inv["header"]), (case, k, header, inv["header"])
        check_rows(rows, inv["rows"], (case, k))
        if inv["checkpoint"]:
            mine = record_as_reference(ck, tmp_path / f"{k}.as_ref.ck")
            want = os.path.join(G.GOLDEN, f"{case}.{k}.ck")
            assert mine.read_bytes() == open(want, "rb").read(), (case, k)


@pytest.mark.skipif(not os.path.exists(G.REF_CLI), reason="reference driver not built")
@pytest.mark.parametrize("case", ["restart", "cap", "walltime", "residual_restart"])
def test_reference_resumes_our_records(tmp_path, case):
    """Our record after the first invocation, resumed by the reference's driver (on its CPU
    backend), continues exactly as the reference's own chain."""
    ck = tmp_path / "run.ck"
    inv0, inv1 = GOLD[case][0], GOLD[case][1]
    code = G.run(G.OURS, G.argv(inv0["args"], ck), tmp_path / "0.csv")[0]
    assert code == inv0["exit"]
    ref_ck = tmp_path / "ref.ck"
    record_as_reference(ck, ref_ck)
    code, header, rows, err = G.run(G.REF_CLI, G.argv(inv1["args"], ref_ck),
                                    tmp_path / "1.csv")
    assert code == inv1["exit"], err
    assert header == inv1["header"]
    assert rows == inv1["rows"]


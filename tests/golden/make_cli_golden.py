#!/usr/bin/env python3
"""Write tests/golden/cli/ from the reference's run driver.

The reference driver (engine/src/cli_rev.c with the core, residual and checkpoint sources) is
compiled from its own sources by ``make -C oracle ref`` into oracle/_ref/tpgp_rev_cli.  This
script runs it on small FID runs and records what ddevol_amd/lib/tpgp_rev must reproduce:

* cli.json: per case, the command-line arguments of every invocation in order, each one's exit
  code, its four header lines and its CSV rows without the wall_seconds and peak_rss_mb columns
  (which measure the machine, not the run);
* <case>.<k>.ck: the checkpoint record left after invocation k, for the cases that checkpoint
  (restart chains, failure and accuracy stops);
* j1j2_n16.edges: the coupling file of the J1-J2 N=16 case (J2_i from InputPCG(7)).

Cases: every model (ring, j1j2 validated, ising, xy, cubic L=3, edges), two-word keys (ring
N=40, validated J1-J2 N=36), the local initial state, a drop-budget failure, a restart with a
longer horizon, a cap failure resumed with a larger cap, an accuracy stop, a walltime pause
resumed to completion, and the independent endpoint bound (--residual-depth: plain, validated,
a restart chain and an accuracy stop it drives).

Run in the build container (where the reference sources exist):
    make -C oracle ref && python tests/golden/make_cli_golden.py
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "cli")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "tpgp_rev_cli")
sys.path.insert(0, ROOT)

from ddevol_amd import hamiltonians as H  # noqa: E402

EDGES = "j1j2_n16.edges"

# name -> list of invocations; "{ck}" is the case's checkpoint path, "{edges}" the edge file
CASES = {
    "ring10": [["--n", "10", "--time", "0.1", "--threshold", "1e-5"]],
    "j1j2_10_validated": [["--model", "j1j2", "--n", "10", "--time", "0.05", "--threshold",
                           "1e-6", "--validated"]],
    "ising8": [["--model", "ising", "--n", "8", "--time", "0.05", "--threshold", "1e-6"]],
    "xy8": [["--model", "xy", "--n", "8", "--time", "0.05", "--threshold", "1e-6"]],
    "cubic3": [["--model", "cubic", "--linear-size", "3", "--time", "0.02", "--threshold",
                "1e-4"]],
    "edges16": [["--model", "edges", "--n", "16", "--edges", "{edges}", "--dt", "0.005",
                 "--time", "0.05", "--threshold", "1e-5"]],
    "local8": [["--n", "8", "--initial", "local", "--time", "0.05", "--threshold", "1e-6"]],
    "ring40": [["--n", "40", "--time", "0.03", "--threshold", "1e-4"]],
    "j1j2_36_validated": [["--model", "j1j2", "--n", "36", "--time", "0.02", "--threshold",
                           "1e-4", "--validated"]],
    "budget": [["--n", "10", "--time", "0.1", "--threshold", "1e-5", "--drop-budget", "2e-3",
                "--checkpoint", "{ck}"]],
    "restart": [["--n", "10", "--time", "0.06", "--threshold", "1e-5", "--checkpoint", "{ck}",
                 "--checkpoint-every", "4"],
                ["--restart", "{ck}", "--time", "0.12"],
                ["--n", "10", "--time", "0.12", "--threshold", "1e-5"]],
    "cap": [["--n", "12", "--time", "0.1", "--threshold", "1e-7", "--cap", "600",
             "--checkpoint", "{ck}"],
            ["--restart", "{ck}", "--cap", "100000"]],
    "accuracy": [["--model", "j1j2", "--n", "8", "--time", "0.1", "--threshold", "1e-5",
                  "--validated", "--max-fid-bound", "1e-4", "--checkpoint", "{ck}"]],
    "walltime": [["--model", "j1j2", "--n", "9", "--time", "0.05", "--threshold", "1e-6",
                  "--walltime", "1e-9", "--checkpoint", "{ck}"],
                 ["--restart", "{ck}"]],
    # the independent endpoint bound (residual columns, records, accuracy stops)
    "residual8": [["--n", "8", "--time", "0.05", "--threshold", "1e-6", "--residual-depth",
                   "2"]],
    "residual_j1j2_validated": [["--model", "j1j2", "--n", "8", "--time", "0.03", "--threshold",
                                 "1e-6", "--validated", "--residual-depth", "1"]],
    "residual_restart": [["--n", "8", "--time", "0.03", "--threshold", "1e-6",
                          "--residual-depth", "1", "--checkpoint", "{ck}",
                          "--checkpoint-every", "2"],
                         ["--restart", "{ck}", "--time", "0.06"]],
    "residual_accuracy": [["--n", "8", "--time", "0.1", "--threshold", "1e-6",
                           "--residual-depth", "0", "--max-fid-bound", "1e-7",
                           "--checkpoint", "{ck}"]],
}


def parse_csv(text):
    lines = text.splitlines()
    header = lines[:4]
    assert lines[4].startswith("time,fid,active,l2_norm"), lines[4]
    rows = [line.split(",")[:14] for line in lines[5:]]
    return header, rows


def write_edges(path):
    j2 = H.j1j2_n16_couplings(16)
    with open(path, "w") as f:
        f.write("# J1-J2 ring N=16: NN coupling 1, NNN coupling J2_i = 0.1 + 0.3 u_i, "
                "u_i from InputPCG(7)\n")
        for i in range(16):
            f.write(f"{i} {(i + 1) % 16} 1\n")
            f.write(f"{i} {(i + 2) % 16} {j2[i]!r}\n")


def main():
    if not os.path.exists(REF_CLI):
        sys.exit("oracle/_ref/tpgp_rev_cli is not built: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    write_edges(os.path.join(OUT, EDGES))
    golden = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, runs in CASES.items():
            ck = os.path.join(tmp, name + ".ck")
            out = []
            for k, args in enumerate(runs):
                argv = [a.replace("{ck}", ck).replace("{edges}", os.path.join(OUT, EDGES))
                        for a in args]
                csv = os.path.join(tmp, f"{name}.{k}.csv")
                r = subprocess.run([REF_CLI, *argv, "--output", csv], capture_output=True,
                                   text=True, timeout=600)
                header, rows = parse_csv(open(csv).read())
                saved = any("{ck}" in a for a in args) and os.path.exists(ck)
                if saved:
                    shutil.copyfile(ck, os.path.join(OUT, f"{name}.{k}.ck"))
                out.append({"args": args, "exit": r.returncode, "header": header,
                            "rows": rows, "checkpoint": saved})
                print(f"{name}[{k}]: exit {r.returncode}, {len(rows)} rows, "
                      f"checkpoint {saved}", flush=True)
            golden[name] = out
    with open(os.path.join(OUT, "cli.json"), "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()

"""Shared helpers for the run-driver tests: run a driver, parse its CSV, normalise records."""
from __future__ import annotations

import json
import math
import os
import struct
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "cli")
OURS = os.path.join(ROOT, "ddevol_amd", "lib", "tpgp_rev")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "tpgp_rev_cli")
# the reference's driver, unedited, linked against libtpgp_engine.so + libtpgp_cuda_shim.so
REF_ON_ENGINE = os.path.join(ROOT, "oracle", "_ref", "tpgp_rev_cli_on_engine")


def golden():
    with open(os.path.join(GOLDEN, "cli.json")) as f:
        return json.load(f)


def argv(args, ck, edges=os.path.join(GOLDEN, "j1j2_n16.edges")):
    return [a.replace("{ck}", str(ck)).replace("{edges}", edges) for a in args]


def run(binary, args, csv, timeout=300):
    """Run a driver; returns (exit code, header lines, rows of the 14 run columns, stderr)."""
    r = subprocess.run([binary, *args, "--output", str(csv)], capture_output=True, text=True,
                       timeout=timeout)
    header, rows = [], []
    if os.path.exists(csv):
        lines = open(csv).read().splitlines()
        header = lines[:4]
        rows = [line.split(",")[:14] for line in lines[5:]]
    return r.returncode, header, rows, r.stderr


def ulps(a: str, b: str) -> int:
    """Distance in binary64 units in the last place between two printed finite doubles."""
    x, y = float(a), float(b)
    if x == y:
        return 0
    if not (math.isfinite(x) and math.isfinite(y)) or (x < 0) != (y < 0):
        return 1 << 62
    ix = struct.unpack("<q", struct.pack("<d", abs(x)))[0]
    iy = struct.unpack("<q", struct.pack("<d", abs(y)))[0]
    return abs(ix - iy)


def same_header(ours, ref):
    """Header lines equal, with the backend field of line 1 naming each driver's own device."""
    import re

    norm = [re.sub(r"backend=\d+", "backend=*", h) for h in ours]
    want = [re.sub(r"backend=\d+", "backend=*", h) for h in ref]
    return norm == want

"""The 256-bit arithmetic of the device residual vectors (ddevol_amd/csrc/tpgp_f256.h) against
MPFR at 256 bits, on the CPU.

The residual bound's entries are chains of mpfr_mul and mpfr_add at precision 256, rounded
toward -inf for lower ends and toward +inf for upper ends (residual_rev.c:55, 105-106).  The
device computes them with tpgp_f256.h, which is plain C as well; tests/f256_mpfr_check.c
compiles that header on the host and compares every operation with MPFR bit for bit (value and
sign of zero): sums across every exponent gap that matters to the 384-bit window, exact and
near cancellations, carries, signed zeros, products with scalars of 1..128 bits, binary64
conversions including subnormals, and the generator's accumulate chains.
"""
from __future__ import annotations

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def mpfr_prefix():
    for d in ("/opt/conda", "/usr/local", "/usr"):
        if os.path.exists(os.path.join(d, "include", "mpfr.h")):
            return d
    return None


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    pre = mpfr_prefix()
    if not pre:
        pytest.skip("MPFR headers not present")
    exe = str(tmp_path_factory.mktemp("f256") / "f256_check")
    subprocess.run(["gcc", "-std=c17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{pre}/include",
                    os.path.join(HERE, "f256_mpfr_check.c"), f"-L{pre}/lib",
                    f"-Wl,-rpath,{pre}/lib", "-lmpfr", "-lgmp", "-lm", "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_f256_matches_mpfr(checker, seed):
    r = subprocess.run([checker, "200000", str(seed)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout
    assert int(r.stdout.split()[1]) > 600000

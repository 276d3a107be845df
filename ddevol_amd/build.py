"""Build the native libraries in-tree.

* ``ddevol_amd/lib/libtpgp_hip.so``    -- HIP kernels + C ABI (hipcc, gfx950)
* ``ddevol_amd/lib/libtpgp_engine.so`` -- C17 host engine (gcc), linked to the above, with MPFR
  when its headers are present (as the reference's TPGP_HAVE_MPFR build).
* ``ddevol_amd/lib/libtpgp_cuda_shim.so`` -- the reference's five ``tpgp_cuda_*`` backend symbols
  forwarded to the above, for programs built against the reference's CUDA header.
* ``ddevol_amd/lib/tpgp_rev``           -- the FID run driver (command line of the reference's
  ``tpgp_rev``) on the engine above.

Run ``python -m ddevol_amd.build`` (or ``__graft_entry__.build()``).  Every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ARCH = os.environ.get("TPGP_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # binary64 exactly as written: no contraction into FMA, no fast-math
    "-ffp-contract=off", "-fno-fast-math",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wextra",
]
CC_FLAGS = ["-std=c17", "-O2", "-fPIC", "-shared", "-fno-fast-math", "-ffp-contract=off",
            "-Wall", "-Wextra"]


def mpfr_prefix():
    for d in ("/opt/conda", "/usr/local", "/usr"):
        if os.path.exists(os.path.join(d, "include", "mpfr.h")):
            return d
    return None


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise FileNotFoundError("hipcc not found")


def run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"  # the host pass does not know the device-only feature
    err = "\n".join(line for line in r.stderr.splitlines() if noise not in line)
    if r.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError(f"build failed: {cmd[0]}")
    return err


def build(verbose: bool = False) -> None:
    os.makedirs(LIB, exist_ok=True)
    hip_so = os.path.join(LIB, "libtpgp_hip.so")
    tmp = hip_so + ".tmp"
    # the device units compile in parallel, then link into one library
    units = ["tpgp_hip", "tpgp_generator"]
    objs = [os.path.join(LIB, u + ".o") for u in units]
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    procs = [(subprocess.Popen([hipcc(), *flags, "-c", "-o", o, os.path.join(CSRC, u + ".hip")],
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), u)
             for u, o in zip(units, objs)]
    for p, u in procs:
        _, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write("\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l))
            raise RuntimeError(f"build failed: hipcc {u}.hip")
    run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, hip_so)
    for o in objs:
        os.remove(o)
    eng_so = os.path.join(LIB, "libtpgp_engine.so")
    cmd = ["gcc", *CC_FLAGS]
    pre = mpfr_prefix()
    libs = []
    if pre:
        cmd += ["-DTPGP_HAVE_MPFR=1", f"-I{pre}/include"]
        libs = [f"-L{pre}/lib", f"-Wl,-rpath,{pre}/lib", "-lmpfr", "-lgmp"]
    tmp = eng_so + ".tmp"
    cmd += ["-o", tmp, os.path.join(CSRC, "tpgp_engine.c"),
            os.path.join(CSRC, "tpgp_checkpoint.c"), os.path.join(CSRC, "tpgp_residual.c"),
            f"-L{LIB}", "-ltpgp_hip",
            "-Wl,-rpath,$ORIGIN", *libs, "-lm", "-lpthread"]
    run(cmd)
    os.replace(tmp, eng_so)
    shim = os.path.join(LIB, "libtpgp_cuda_shim.so")
    tmp = shim + ".tmp"
    run(["gcc", *CC_FLAGS, "-o", tmp, os.path.join(CSRC, "tpgp_cuda_shim.c"), f"-L{LIB}",
         "-ltpgp_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, shim)
    cli = os.path.join(LIB, "tpgp_rev")
    tmp = cli + ".tmp"
    run(["gcc", *[f for f in CC_FLAGS if f not in ("-shared", "-fPIC")], "-o", tmp,
         os.path.join(CSRC, "tpgp_cli.c"), f"-L{LIB}", "-ltpgp_engine", "-ltpgp_hip",
         "-Wl,-rpath,$ORIGIN", "-Wl,--allow-shlib-undefined", *libs, "-lm"])
    os.replace(tmp, cli)


if __name__ == "__main__":
    build()

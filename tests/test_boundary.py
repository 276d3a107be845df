"""The C-ABI libraries load on the CPU host and export every function include/*.h declares.

No compute call is made (there is no GPU here); loading libtpgp_hip.so only resolves symbols.
"""
import ctypes as C
import os
import re

import pytest

from ddevol_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tpgp_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("tpgp_hip.h", N.HIP_SO), ("tpgp_engine.h", N.ENGINE_SO)])
def test_exports(header, lib):
    if not os.path.exists(lib):
        pytest.fail(f"{lib} not built (run python -m ddevol_amd.build)")
    so = C.CDLL(lib)
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_cuda_shim_exports():
    """libtpgp_cuda_shim.so carries the reference's five backend symbols
    (tpgp_cuda_rev.h:20-49), forwarding to tpgp_hip_*; without a device it reports the build
    capability and a create failure, as the reference's backend does."""
    path = os.path.join(N.LIB, "libtpgp_cuda_shim.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run python -m ddevol_amd.build)")
    so = C.CDLL(path)
    for name in ("tpgp_cuda_compiled", "tpgp_cuda_create", "tpgp_cuda_destroy",
                 "tpgp_cuda_device_name", "tpgp_cuda_branches"):
        assert hasattr(so, name), name
    assert so.tpgp_cuda_compiled() == 1
    so.tpgp_cuda_device_name.restype = C.c_char_p
    so.tpgp_cuda_device_name.argtypes = [C.c_void_p]
    assert so.tpgp_cuda_device_name(None) == b"unavailable"


def test_binding_lists_cover_headers():
    assert set(declared("tpgp_hip.h")) <= set(N.HIP_SYMBOLS)
    assert set(declared("tpgp_engine.h")) <= set(N.ENGINE_SYMBOLS)


def test_compiled_flag_without_device():
    assert N.hip().tpgp_hip_compiled() == 1


def test_no_cpu_backend():
    """The engine exposes only the GPU backend slot: CPU backends are rejected (no fallback)."""
    from ddevol_amd import engine as E

    eng = N.engine()
    assert eng.tpgp_has_openmp() == 0
    mp = E.Map(4)
    mp.add_single(0, 1, 1.0)
    for backend in (0, 1, 2, 3):  # serial, pthread, OpenMP, Metal slots of the reference
        o = E.options(backend=backend)
        assert E.rotate(mp, 0, 1, 3, 3, .1, o, E.ledger(0)) == N.EBACKEND


def test_host_map_container():
    """tpgp_map_* host operations keep the reference's sorted-unique semantics."""
    from ddevol_amd import engine as E

    mp = E.Map(40)
    assert mp.n_words == 2
    assert mp.add_single(39, 3, 0.5) == 0
    assert mp.add_single(0, 1, 1.0) == 0
    assert mp.add_single(0, 1, -1.0) == 0  # exact cancellation removes the key
    assert len(mp) == 1
    keys, vals = mp.arrays()
    assert int(keys[0][1]) == 3 << 14 and vals[0] == 0.5
    assert mp.add(np.array([0, 1 << 16], dtype=np.uint64), 1.0) == N.EINVAL  # padding bits


import numpy as np  # noqa: E402

"""Loading the Liouvillian golden cases (tests/golden/liouv.npz, written by make_liouv_golden.py)."""
from __future__ import annotations

import hashlib
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "liouv.npz")


def cases():
    z = np.load(PATH)  # plain arrays only (no pickles)
    names = sorted(k[:-len("_meta")] for k in z.files if k.endswith("_meta"))
    for name in names:
        n, status, out_len = (int(x) for x in z[f"{name}_meta"])
        ij, q, h = z[f"{name}_terms_ij"], z[f"{name}_terms_q"], z[f"{name}_terms_h"]
        terms = [(int(a), int(b), int(c), int(d), float(e))
                 for (a, b), (c, d), e in zip(ij.tolist(), q.tolist(), h.tolist())]
        case = {"name": name, "n": n, "status": status, "out_len": out_len, "terms": terms,
                "keys": z[f"{name}_in_keys"], "vals": z[f"{name}_in_vals"]}
        if f"{name}_out_keys" in z.files:
            case["out_keys"], case["out_vals"] = z[f"{name}_out_keys"], z[f"{name}_out_vals"]
        else:
            case["digest"] = bytes(z[f"{name}_digest"])
        yield case


def digest(keys, vals):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(keys, "<u8").tobytes())
    h.update(np.ascontiguousarray(vals, "<f8").tobytes())
    return h.digest()


def matches(case, status, keys, vals):
    """The status, and the output whole or by digest."""
    if status != case["status"]:
        return False
    if status:
        return True
    w = (case["n"] + 31) // 32
    keys = np.asarray(keys, np.uint64).reshape(-1, w)
    vals = np.asarray(vals, np.float64)
    if len(vals) != case["out_len"]:
        return False
    if "digest" in case:
        return digest(keys, vals) == case["digest"]
    return (np.array_equal(keys, case["out_keys"]) and
            np.array_equal(vals.view(np.uint64), case["out_vals"].view(np.uint64)))

"""ctypes bindings to the native libraries (include/tpgp_hip.h, include/tpgp_engine.h).

Loading fails loudly when a library is missing: the engine has no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib")
HIP_SO = os.path.join(LIB, "libtpgp_hip.so")
ENGINE_SO = os.path.join(LIB, "libtpgp_engine.so")

# statuses of include/tpgp_hip.h and include/tpgp_engine.h
HIP_OK, HIP_EINVAL, HIP_ENOMEM, HIP_EDEVICE, HIP_ENUMERIC = range(5)
HIP_ECAPACITY = 16
OK, EINVAL, ENOMEM, ECAP, ENUMERIC, EBACKEND, EBUDGET = range(7)
TPGP_HIP = 4

# launch-shape tuning (tpgp_hip_set_tuning): per key width (1 or 2 words), parameter and size
# bucket (floor(log2 terms))
TUNE_PARTITION_ROUNDS, TUNE_COUNT_GRID, TUNE_WRITE_GRID = range(3)
TUNE_PARAMS, TUNE_BUCKETS = 3, 48
TUNE_NAMES = ["partition_rounds", "count_grid", "write_grid"]
# the table the autotuner measured on MI355X (ddevol_amd/autotune.py); tpgp_hip_create loads it
# from beside the library unless TPGP_HIP_TUNING names another file (empty: built-in defaults)
TUNING_TABLE = os.path.join(HERE, "tuning_gfx950.txt")

_hip = None
_eng = None


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"native library {path} is not built; run python -m ddevol_amd.build "
                           "(there is no CPU fallback)")
    return C.CDLL(path)


class TpgpMap(C.Structure):  # tpgp_rev.h:25-29 layout
    _fields_ = [("n_spins", C.c_size_t), ("n_words", C.c_size_t), ("len", C.c_size_t),
                ("capacity", C.c_size_t), ("keys", C.POINTER(C.c_uint64)),
                ("values", C.POINTER(C.c_double))]


class TpgpTerm(C.Structure):  # tpgp_rev.h:30-35
    _fields_ = [("i", C.c_size_t), ("j", C.c_size_t), ("qi", C.c_uint), ("qj", C.c_uint),
                ("h", C.c_double), ("layer", C.c_size_t)]


class TpgpOptions(C.Structure):  # tpgp_rev.h:36-45 layout, GPU slot
    _fields_ = [("threshold", C.c_double), ("max_terms", C.c_size_t), ("threads", C.c_uint),
                ("backend", C.c_int), ("validated", C.c_int), ("max_drop_l2", C.c_double),
                ("metal_context", C.c_void_p), ("hip_context", C.c_void_p)]


class TpgpLedger(C.Structure):  # tpgp_rev.h:46-51
    _fields_ = [("dropped_l1", C.c_double), ("dropped_l2", C.c_double),
                ("dropped_l2_squared", C.c_double), ("arithmetic_l2", C.c_double),
                ("rotations", C.c_size_t), ("removal_events", C.c_size_t),
                ("peak_terms", C.c_size_t), ("arithmetic_enclosed", C.c_int)]

    def bits(self):
        f = np.array([self.dropped_l1, self.dropped_l2, self.dropped_l2_squared,
                      self.arithmetic_l2], dtype=np.float64).view(np.uint64).tolist()
        return tuple(f) + (self.rotations, self.removal_events, self.peak_terms,
                           self.arithmetic_enclosed)


class EvolveOptions(C.Structure):
    _fields_ = [("exact_ledger", C.c_int), ("pauli", C.c_uint), ("normalizer", C.c_size_t),
                ("initial_capacity", C.c_size_t)]


class FidJob(C.Structure):  # tpgp_fid_job
    _fields_ = [("initial", C.POINTER(TpgpMap)), ("terms", C.POINTER(TpgpTerm)),
                ("count", C.c_size_t), ("dt", C.c_double), ("steps", C.c_size_t),
                ("options", TpgpOptions), ("evolve", EvolveOptions)]


SLOT_FREE, SLOT_RUNNING, SLOT_DONE, SLOT_FAILED = range(4)


class EvolutionCounters(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("rotations", C.c_uint64), ("updates", C.c_uint64),
                ("in_out", C.c_uint64), ("in_place", C.c_uint64)]


(PROFILE_MERGE, PROFILE_SCATTER, PROFILE_ROTATION, PROFILE_PARTITION, PROFILE_COUNT,
 PROFILE_SCAN, PROFILE_PLAN, PROFILE_HISTOGRAM, PROFILE_SPLITS, PROFILE_SMALL) = range(10)
PROFILE_ALL = 0x3FF
PROFILE_NAMES = ["k_rot_write", "k_part_scatter", "rotation_pass", "partition_pass",
                 "k_rot_count", "k_rot_scan", "k_rot_plan", "k_part_count", "k_rot_splits",
                 "k_small"]


class HipRotation(C.Structure):
    _fields_ = [("active", C.c_int), ("qi", C.c_uint), ("qj", C.c_uint), ("validated", C.c_int),
                ("log", C.c_int), ("i", C.c_size_t), ("j", C.c_size_t), ("c", C.c_double),
                ("s", C.c_double), ("rc", C.c_double), ("rs", C.c_double),
                ("threshold", C.c_double)]


class HipRecord(C.Structure):
    _fields_ = [("len_in", C.c_uint64), ("len_out", C.c_uint64), ("dropped", C.c_uint64),
                ("l1_up", C.c_double), ("sq_up", C.c_double), ("rad_sq_up", C.c_double),
                ("radii", C.c_uint64), ("executed", C.c_uint32), ("nonfinite", C.c_uint32),
                ("overflow", C.c_uint32), ("rotation", C.c_uint32), ("in_place", C.c_uint64)]


class HipLogEntry(C.Structure):
    _fields_ = [("key", C.c_uint64 * 2), ("value", C.c_double), ("aux", C.c_double),
                ("map", C.c_uint32), ("rotation", C.c_uint32), ("kind", C.c_uint32),
                ("pad", C.c_uint32)]


class HipF256(C.Structure):  # tpgp_hip_f256: (-1)^neg (m / 2^256) 2^e, m[3] most significant
    _fields_ = [("m", C.c_uint64 * 4), ("e", C.c_int32), ("neg", C.c_uint32)]


class HipIntervalVector(C.Structure):  # tpgp_hip_interval_vector
    _fields_ = [("len", C.c_size_t), ("keys", C.POINTER(C.c_uint64)),
                ("lower", C.POINTER(HipF256)), ("upper", C.POINTER(HipF256)),
                ("partner", C.POINTER(C.c_uint32)), ("square", C.POINTER(HipF256) * 2),
                ("product", C.POINTER(HipF256) * 2)]


class Checkpoint(C.Structure):  # checkpoint_rev.h:7-23 layout
    _fields_ = [("map", TpgpMap), ("terms", C.POINTER(TpgpTerm)), ("term_count", C.c_size_t),
                ("options", TpgpOptions), ("ledger", TpgpLedger), ("dt", C.c_double),
                ("pf_bound", C.c_double), ("residual_bound", C.c_double),
                ("max_fid_bound", C.c_double), ("step", C.c_size_t),
                ("total_steps", C.c_size_t), ("residual_depth", C.c_int),
                ("initial_global", C.c_int), ("model", C.c_char * 32), ("j2", C.c_double),
                ("linear_size", C.c_size_t), ("resumed_from_step", C.c_size_t),
                ("prior_total_steps", C.c_size_t), ("prior_max_terms", C.c_size_t),
                ("stop_reason", C.c_int), ("last_status", C.c_int)]


HIP_SYMBOLS = ["tpgp_hip_compiled", "tpgp_hip_create", "tpgp_hip_destroy",
               "tpgp_hip_device_name", "tpgp_hip_branches", "tpgp_hip_branches_noalloc",
               "tpgp_hip_batch_create",
               "tpgp_hip_batch_destroy", "tpgp_hip_batch_upload", "tpgp_hip_batch_download",
               "tpgp_hip_batch_rotate", "tpgp_hip_batch_rotate_many", "tpgp_hip_batch_readout", "tpgp_hip_batch_sync",
               "tpgp_hip_batch_commit", "tpgp_hip_batch_rollback", "tpgp_hip_batch_reserve",
               "tpgp_hip_batch_info", "tpgp_hip_batch_stream", "tpgp_hip_batch_profile",
               "tpgp_hip_batch_profile_read", "tpgp_hip_batch_assign",
               "tpgp_hip_batch_reserve_log", "tpgp_hip_batch_reserve_many",
               "tpgp_hip_batch_resize", "tpgp_hip_batch_graph_replay",
               "tpgp_hip_batch_graph_begin", "tpgp_hip_batch_graph_end",
               "tpgp_hip_batch_sum_squares", "tpgp_hip_liouvillian", "tpgp_hip_set_tuning",
               "tpgp_hip_get_tuning", "tpgp_hip_residual_vectors",
               "tpgp_hip_interval_vector_free"]
ENGINE_SYMBOLS = ["tpgp_status_string", "tpgp_has_mpfr", "tpgp_has_openmp",
                  "tpgp_default_options", "tpgp_ledger_init", "tpgp_map_init", "tpgp_map_free",
                  "tpgp_map_clone", "tpgp_key_get", "tpgp_key_set", "tpgp_map_add",
                  "tpgp_map_add_single", "tpgp_map_get", "tpgp_map_l2", "tpgp_map_single_sum",
                  "tpgp_readout", "tpgp_pauli_phase", "tpgp_rotate", "tpgp_strang_step",
                  "tpgp_strang_bound", "tpgp_strang_bound_taylor", "tpgp_bound_add_up",
                  "tpgp_bound_fid_up", "tpgp_residual_bound", "tpgp_residual_bound_reference",
                  "tpgp_hip_evolve", "tpgp_hip_last_error",
                  "tpgp_evolution_open", "tpgp_evolution_run", "tpgp_evolution_close",
                  "tpgp_evolution_counters_get", "tpgp_evolution_profile",
                  "tpgp_evolution_profile_read", "tpgp_evolution_key_words",
                  "tpgp_evolution_graph_counts",
                  "tpgp_fleet_open", "tpgp_fleet_close", "tpgp_fleet_admit", "tpgp_fleet_step",
                  "tpgp_fleet_slot", "tpgp_fleet_fid", "tpgp_fleet_release",
                  "tpgp_fleet_counters_get", "tpgp_fleet_slot_bytes", "tpgp_fleet_set_budget",
                  "tpgp_fleet_trim", "tpgp_fleet_graph_counts",
                  "tpgp_checkpoint_free", "tpgp_checkpoint_save", "tpgp_checkpoint_load",
                  "tpgp_checkpoint_same_configuration", "tpgp_payload_sha256",
                  "tpgp_evolution_snapshot", "tpgp_evolution_l2", "tpgp_liouvillian"]


def hip():
    """The device library with typed entry points."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load(HIP_SO)
    p_u64, p_d, p_c = C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_char_p
    sz = C.c_size_t

    def f(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    f("tpgp_hip_compiled", C.c_int)
    f("tpgp_hip_create", C.c_void_p, C.c_int, p_c, sz)
    f("tpgp_hip_destroy", None, C.c_void_p)
    f("tpgp_hip_device_name", C.c_char_p, C.c_void_p)
    f("tpgp_hip_branches", C.c_int, C.c_void_p, p_u64, p_d, sz, sz, sz, sz, sz, C.c_uint,
      C.c_uint, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, p_u64, p_d, p_d, p_c,
      sz)
    f("tpgp_hip_branches_noalloc", C.c_int, C.c_void_p, p_u64, p_d, sz, sz, sz, sz, sz,
      C.c_uint, C.c_uint, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, p_u64, p_d,
      p_d, p_c, sz)
    f("tpgp_hip_batch_create", C.c_void_p, C.c_void_p, sz, sz, C.POINTER(sz), C.POINTER(sz),
      p_c, sz)
    f("tpgp_hip_batch_destroy", None, C.c_void_p)
    f("tpgp_hip_batch_upload", C.c_int, C.c_void_p, sz, p_u64, p_d, sz, p_c, sz)
    f("tpgp_hip_batch_download", C.c_int, C.c_void_p, sz, p_u64, p_d, sz, C.POINTER(sz), p_c, sz)
    f("tpgp_hip_batch_rotate", C.c_int, C.c_void_p, C.POINTER(HipRotation), C.c_uint32, p_c, sz)
    f("tpgp_hip_batch_rotate_many", C.c_int, C.c_void_p, C.POINTER(HipRotation),
      C.POINTER(C.c_uint32), sz, p_c, sz)
    f("tpgp_hip_batch_readout", C.c_int, C.c_void_p, C.c_uint, C.POINTER(C.c_int), p_c, sz)
    f("tpgp_hip_batch_sync", C.c_int, C.c_void_p, C.POINTER(HipRecord), sz, C.POINTER(sz), p_d,
      sz, C.POINTER(sz), p_u64, C.POINTER(HipLogEntry), sz, C.POINTER(sz), p_c, sz)
    f("tpgp_hip_batch_commit", C.c_int, C.c_void_p, C.POINTER(C.c_int), p_c, sz)
    f("tpgp_hip_batch_rollback", C.c_int, C.c_void_p, p_c, sz)
    f("tpgp_hip_batch_reserve", C.c_int, C.c_void_p, sz, sz, p_c, sz)
    f("tpgp_hip_batch_info", C.c_int, C.c_void_p, sz, C.POINTER(sz), C.POINTER(sz))
    f("tpgp_hip_batch_stream", C.c_void_p, C.c_void_p)
    f("tpgp_hip_batch_assign", C.c_int, C.c_void_p, sz, sz, p_u64, p_d, sz, p_c, sz)
    f("tpgp_hip_batch_reserve_log", C.c_int, C.c_void_p, sz, p_c, sz)
    f("tpgp_hip_batch_reserve_many", C.c_int, C.c_void_p, C.POINTER(sz), p_c, sz)
    f("tpgp_hip_batch_resize", C.c_int, C.c_void_p, C.POINTER(sz), p_c, sz)
    f("tpgp_hip_batch_graph_replay", C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_int),
      C.POINTER(C.c_int), p_c, sz)
    f("tpgp_hip_batch_graph_begin", C.c_int, C.c_void_p, C.c_uint64, p_c, sz)
    f("tpgp_hip_batch_graph_end", C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int), p_c, sz)
    f("tpgp_hip_batch_sum_squares", C.c_int, C.c_void_p, sz, p_d, p_c, sz)
    f("tpgp_hip_liouvillian", C.c_int, C.c_void_p, p_u64, p_d, sz, sz, sz, C.POINTER(TpgpTerm),
      sz, p_u64, p_d, sz, C.POINTER(sz), p_c, sz)
    f("tpgp_hip_residual_vectors", C.c_int, C.c_void_p, p_u64, p_d, sz, p_u64, p_d, sz, sz, sz,
      C.POINTER(TpgpTerm), sz, C.c_double, C.c_int, C.POINTER(HipIntervalVector), p_c, sz)
    f("tpgp_hip_interval_vector_free", None, C.POINTER(HipIntervalVector))
    f("tpgp_hip_batch_profile", C.c_int, C.c_void_p, C.c_uint)
    f("tpgp_hip_batch_profile_read", C.c_int, C.c_void_p, C.c_uint, C.POINTER(C.c_double),
      C.POINTER(C.c_uint64))
    f("tpgp_hip_set_tuning", C.c_int, C.c_void_p, C.c_uint, C.c_uint, C.c_uint, C.c_uint)
    f("tpgp_hip_get_tuning", C.c_int, C.c_void_p, C.c_uint, C.c_uint, C.c_uint,
      C.POINTER(C.c_uint))
    _hip = L
    return L


def engine():
    """The C17 host engine (reference-shaped API) with typed entry points."""
    global _eng
    if _eng is not None:
        return _eng
    hip()  # resolve the dependency from the same directory first
    L = _load(ENGINE_SO)
    mp = C.POINTER(TpgpMap)
    op = C.POINTER(TpgpOptions)
    lp = C.POINTER(TpgpLedger)
    sz = C.c_size_t

    def f(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    f("tpgp_status_string", C.c_char_p, C.c_int)
    f("tpgp_has_mpfr", C.c_int)
    f("tpgp_has_openmp", C.c_int)
    f("tpgp_default_options", TpgpOptions)
    f("tpgp_ledger_init", None, lp, C.c_int)
    f("tpgp_map_init", C.c_int, mp, sz)
    f("tpgp_map_free", None, mp)
    f("tpgp_map_clone", C.c_int, mp, mp)
    f("tpgp_key_get", C.c_uint, C.POINTER(C.c_uint64), sz)
    f("tpgp_key_set", None, C.POINTER(C.c_uint64), sz, C.c_uint)
    f("tpgp_map_add", C.c_int, mp, C.POINTER(C.c_uint64), C.c_double)
    f("tpgp_map_add_single", C.c_int, mp, sz, C.c_uint, C.c_double)
    f("tpgp_map_get", C.c_double, mp, C.POINTER(C.c_uint64))
    f("tpgp_map_l2", C.c_double, mp)
    f("tpgp_map_single_sum", C.c_double, mp, C.c_uint)
    f("tpgp_readout", C.c_int, mp, C.c_uint, sz, C.POINTER(C.c_double), C.POINTER(C.c_double))
    f("tpgp_pauli_phase", C.c_uint, C.c_uint, C.c_uint)
    f("tpgp_rotate", C.c_int, mp, sz, sz, C.c_uint, C.c_uint, C.c_double, op, lp)
    f("tpgp_strang_step", C.c_int, mp, C.POINTER(TpgpTerm), sz, C.c_double, op, lp)
    f("tpgp_strang_bound", C.c_double, C.POINTER(TpgpTerm), sz, C.c_double)
    f("tpgp_strang_bound_taylor", C.c_double, C.POINTER(TpgpTerm), sz, C.c_double)
    f("tpgp_bound_add_up", C.c_double, C.c_double, C.c_double)
    f("tpgp_bound_fid_up", C.c_double, C.c_double, sz)
    for name in ("tpgp_residual_bound", "tpgp_residual_bound_reference"):
        f(name, C.c_int, mp, mp, C.POINTER(TpgpTerm), sz, C.c_double, C.c_uint,
          C.POINTER(C.c_double))
    f("tpgp_hip_evolve", C.c_int, mp, C.POINTER(TpgpTerm), sz, C.c_double, sz, op,
      C.POINTER(EvolveOptions), lp, C.POINTER(C.c_double), C.POINTER(C.c_double),
      C.POINTER(sz), C.POINTER(sz))
    f("tpgp_hip_last_error", C.c_char_p)
    pd, psz = C.POINTER(C.c_double), C.POINTER(sz)
    f("tpgp_evolution_open", C.c_int, C.POINTER(C.c_void_p), mp, C.POINTER(TpgpTerm), sz,
      C.c_double, op, C.POINTER(EvolveOptions))
    f("tpgp_evolution_run", C.c_int, C.c_void_p, sz, lp, pd, pd, psz, psz)
    f("tpgp_evolution_close", C.c_int, C.c_void_p, mp)
    f("tpgp_evolution_counters_get", None, C.c_void_p, C.POINTER(EvolutionCounters))
    f("tpgp_evolution_profile", C.c_int, C.c_void_p, C.c_uint)
    f("tpgp_evolution_profile_read", C.c_int, C.c_void_p, C.c_uint, pd, C.POINTER(C.c_uint64))
    f("tpgp_evolution_key_words", sz, C.c_void_p)
    f("tpgp_evolution_graph_counts", None, C.c_void_p, C.POINTER(C.c_uint64),
      C.POINTER(C.c_uint64))
    vp = C.c_void_p
    f("tpgp_fleet_open", C.c_int, C.POINTER(vp), vp, sz, sz, sz)
    f("tpgp_fleet_close", None, vp)
    f("tpgp_fleet_admit", C.c_int, vp, sz, C.POINTER(FidJob))
    f("tpgp_fleet_step", C.c_int, vp, psz)
    f("tpgp_fleet_slot", C.c_int, vp, sz, C.POINTER(C.c_int), psz, C.POINTER(C.c_int), lp)
    f("tpgp_fleet_fid", C.c_int, vp, sz, pd, pd, psz, sz, psz)
    f("tpgp_fleet_release", C.c_int, vp, sz, mp)
    f("tpgp_fleet_counters_get", None, vp, C.POINTER(EvolutionCounters))
    f("tpgp_fleet_slot_bytes", C.c_uint64, vp, sz)
    f("tpgp_fleet_set_budget", C.c_int, vp, C.c_uint64)
    f("tpgp_fleet_trim", C.c_int, vp)
    f("tpgp_fleet_graph_counts", None, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
    cp = C.POINTER(Checkpoint)
    f("tpgp_checkpoint_free", None, cp)
    f("tpgp_checkpoint_save", C.c_int, C.c_char_p, cp, C.c_char_p, sz)
    f("tpgp_checkpoint_load", C.c_int, C.c_char_p, cp, C.c_char_p, sz)
    f("tpgp_checkpoint_same_configuration", C.c_int, cp, cp)
    f("tpgp_payload_sha256", None, C.c_void_p, sz, C.POINTER(C.c_ubyte))
    f("tpgp_evolution_snapshot", C.c_int, vp, mp)
    f("tpgp_evolution_l2", C.c_int, vp, C.POINTER(C.c_double))
    f("tpgp_liouvillian", C.c_int, mp, mp, C.POINTER(TpgpTerm), sz)
    _eng = L
    return L


def as_u64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def as_dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def terms_array(terms):
    arr = (TpgpTerm * max(len(terms), 1))()
    for k, t in enumerate(terms):
        arr[k] = TpgpTerm(t[0], t[1], t[2], t[3], t[4], k)
    return arr

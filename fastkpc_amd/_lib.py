"""ctypes binding of libkpcmi.so (the gfx950 kernel-PC CI library).

The library is the product: importing this module never falls back to a CPU
implementation.  If the shared library is missing, `lib()` raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkpcmi.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U8 = ctypes.POINTER(ctypes.c_ubyte)


class KpcOptions(ctypes.Structure):
    _fields_ = [
        ("ci_method", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("max_conditioning_size", ctypes.c_int),
        ("index", ctypes.c_double),
        ("num_col", ctypes.c_int),
        ("hsic_sig", ctypes.c_double),
        ("replicates", ctypes.c_int),
        ("include_observed", ctypes.c_int),
        ("has_seed", ctypes.c_int),
        ("seed", ctypes.c_ulonglong),
        ("mode", ctypes.c_int),
        ("plain_test_limit", ctypes.c_int),
        ("perm_stream", ctypes.c_int),
        ("r_state", ctypes.POINTER(ctypes.c_int)),
        ("use_cache", ctypes.c_int),
    ]


class KpcResult(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_int),
        ("levels", ctypes.c_int),
        ("adjacency", _U8),
        ("pmax", _D),
        ("sepset_len", _I),
        ("sepset", _I),
        ("n_edgetests", ctypes.c_int * 64),
        ("n_tasks", ctypes.c_int),
        ("task_level", _I),
        ("task_x", _I),
        ("task_y", _I),
        ("task_slen", _I),
        ("task_S", _I),
        ("task_p", _D),
        ("task_deleted", _U8),
        ("wall_seconds", ctypes.c_double),
        ("stage_seconds", ctypes.c_double * 8),
        ("physical_tests", ctypes.c_longlong),
        ("residual_fits", ctypes.c_longlong),
        ("components", ctypes.c_longlong),
        ("setups", ctypes.c_longlong),
        ("magic_evaluations", ctypes.c_longlong),
        ("truncated", ctypes.c_int),
        ("level_tests", ctypes.c_longlong * 64),
        ("level_physical", ctypes.c_longlong * 64),
        ("level_fits", ctypes.c_longlong * 64),
        ("level_components", ctypes.c_longlong * 64),
        ("level_evaluations", ctypes.c_longlong * 64),
        ("level_lanczos_steps", ctypes.c_longlong * 64),
        ("level_seconds", ctypes.c_double * 64),
        ("plain_tests", ctypes.c_longlong * 64),
        ("plain_seconds", ctypes.c_double * 64),
        ("stage_launches", ctypes.c_longlong * 8),
        ("task_margin", _D),
        ("level_stage_seconds", (ctypes.c_double * 5) * 64),
        ("task_keys", ctypes.c_void_p),
        ("task_stat", _D),
        ("task_mean", _D),
        ("task_var", _D),
        ("level_tasks_planned", ctypes.c_longlong * 64),
        ("task_index", ctypes.POINTER(ctypes.c_int)),
        ("level_tasks_ignored", ctypes.c_longlong * 64),
        ("level_rounds", ctypes.c_longlong * 64),
        ("task_perm_seed", ctypes.POINTER(ctypes.c_ulonglong)),
        ("r_state_out", ctypes.c_int * 625),
        ("result_cache_hits", ctypes.c_longlong),
        ("target_cache_hits", ctypes.c_longlong),
    ]


class KpcOrientOptions(ctypes.Structure):
    _fields_ = [
        ("ci_method", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("num_col", ctypes.c_int),
        ("hsic_sig", ctypes.c_double),
        ("index", ctypes.c_double),
        ("replicates", ctypes.c_int),
        ("include_observed", ctypes.c_int),
        ("has_seed", ctypes.c_int),
        ("seed", ctypes.c_ulonglong),
        ("perm_stream", ctypes.c_int),
        ("r_state", ctypes.POINTER(ctypes.c_int)),
        ("orient_collider", ctypes.c_int),
        ("solve_confl", ctypes.c_int),
        ("rules", ctypes.c_int * 3),
    ]


class KpcOrientResult(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_int),
        ("pdag", _I),
        ("n_calls", ctypes.c_int),
        ("call_v", _I),
        ("call_slen", _I),
        ("call_S", _I),
        ("call_plen", _I),
        ("call_parents", _I),
        ("call_reject", _I),
        ("n_tests", ctypes.c_int),
        ("test_call", _I),
        ("test_other", _I),
        ("test_out", _D),
        ("r_state_out", ctypes.c_int * 625),
        ("wall_seconds", ctypes.c_double),
    ]


class KpcCacheInfo(ctypes.Structure):
    _fields_ = [(name, ctypes.c_longlong) for name in (
        "capacity", "entries", "dataset_entries", "requests", "hits", "misses", "inserts",
        "evictions", "generation", "target_capacity", "target_entries", "target_dataset_entries",
        "target_requests", "target_hits", "target_misses", "target_inserts", "target_evictions",
        "target_generation", "epoch")] + [("dataset_key", ctypes.c_char * 33)]


EXPORTS = (
    "kpc_default_options", "kpc_skeleton_run", "kpc_result_free", "kpc_ci_batch",
    "kpc_regrxons_batch", "kpc_regrxons_mixed", "kpc_gcv_grad_hess", "kpc_dcov_components",
    "kpc_last_fit_timing", "kpc_tuning_report", "kpc_task_seed", "kpc_permutation_table", "kpc_last_error",
    "kpc_device_count", "kpc_empty_cache", "kpc_dcov_components_noalloc", "kpc_ci_batch_rs",
    "kpc_cache_control", "kpc_cache_state", "kpc_default_orient_options", "kpc_wanpdag_run",
    "kpc_orient_result_free",
)

_lib = None


class KpcLibraryError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KpcLibraryError(
            "libkpcmi.so is not built; run `python -m fastkpc_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    L.kpc_default_options.argtypes = [ctypes.POINTER(KpcOptions)]
    L.kpc_default_options.restype = None
    L.kpc_skeleton_run.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(KpcOptions),
                                   ctypes.POINTER(ctypes.POINTER(KpcResult))]
    L.kpc_result_free.argtypes = [ctypes.POINTER(KpcResult)]
    L.kpc_result_free.restype = None
    L.kpc_ci_batch.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(KpcOptions),
                               ctypes.c_int, _I, _I, ctypes.c_int, _I, _D]
    L.kpc_ci_batch_rs.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(KpcOptions),
                                  ctypes.c_int, _I, _I, ctypes.c_int, _I, _D, _I]
    L.kpc_regrxons_batch.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I,
                                     ctypes.c_int, _I, _D, _D]
    L.kpc_regrxons_mixed.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, _I, _I, _I,
                                     _D, _D]
    L.kpc_last_fit_timing.argtypes = [_D]
    L.kpc_tuning_report.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.kpc_gcv_grad_hess.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int,
                                    _I, _D, _D, _D, _D]
    L.kpc_dcov_components.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]
    L.kpc_task_seed.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)]
    L.kpc_permutation_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_ulonglong), _I]
    L.kpc_cache_control.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.POINTER(KpcCacheInfo)]
    L.kpc_cache_state.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(KpcCacheInfo)]
    L.kpc_default_orient_options.argtypes = [ctypes.POINTER(KpcOrientOptions)]
    L.kpc_default_orient_options.restype = None
    L.kpc_wanpdag_run.argtypes = [_D, ctypes.c_int, ctypes.c_int, _U8, _I, _I,
                                  ctypes.POINTER(KpcOrientOptions),
                                  ctypes.POINTER(ctypes.POINTER(KpcOrientResult))]
    L.kpc_orient_result_free.argtypes = [ctypes.POINTER(KpcOrientResult)]
    L.kpc_orient_result_free.restype = None
    L.kpc_last_error.restype = ctypes.c_char_p
    L.kpc_device_count.restype = ctypes.c_int
    L.kpc_empty_cache.restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().kpc_last_error().decode("utf-8", "replace")
        raise KpcLibraryError("kpcmi error %d: %s" % (rc, msg))


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)

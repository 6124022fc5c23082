"""ctypes binding of libsfft_amd.so (the C ABI declared in include/sfft.h).

The library is built in-tree by ``sfft_amd/build.py``.  Loading fails loudly
when it is missing: there is no CPU fallback behind this package.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfft_amd.so")

# Every symbol include/sfft.h declares.
EXPORTS = [
    "sfft_malloc", "sfft_free", "sfft_make_plan", "sfft_free_plan", "sfft_exec",
    "sfft_exec_many", "sfft_plan_output_capacity", "sfft_exec_sparse",
    "sfft_exec_many_sparse", "sfft_plan_set_precision", "sfft_plan_seed",
    "sfft_plan_set_stream", "sfft_plan_set_reference_conjugate",
    "sfft_plan_memory", "sfft_last_error", "sfft_plan_params",
    "sfft_derive_params", "sfft_design_filter", "sfft_debug_buckets",
    "sfft_plan_enable_timing", "sfft_plan_stage_times", "sfft_plan_filter",
    "sfft_plan_last_status",
    "sfft_plan_set_schedule", "sfft_plan_get_schedule", "sfft_plan_autotune",
    "sfft_exec_many_plans", "sfft_plan_enable_graphs", "sfft_generate_input",
    "sfft_exec_sparse_noalloc", "sfft_plan_debug_comb_votes", "sfft_debug_comb_votes",
    "sfft_plan_debug_v3",
    "sfft_plan_debug_v3_inline_pre",
]

# Every symbol include/sfft_server.h declares.
SERVER_EXPORTS = [
    "sfft_output_capacity", "sfft_server_create", "sfft_server_prepare",
    "sfft_server_submit", "sfft_server_poll", "sfft_server_wait",
    "sfft_server_get_stats", "sfft_server_destroy",
]

SFFT_VERSION_1, SFFT_VERSION_2, SFFT_VERSION_3 = 0, 1, 2
SFFT_PRECISION_FP64, SFFT_PRECISION_FP32_REFINE = 0, 1


class SfftPlan(C.Structure):
    _fields_ = [("version", C.c_int), ("n", C.c_uint), ("k", C.c_uint),
                ("data", C.c_void_p)]


class SfftRequest(C.Structure):
    _fields_ = [("n", C.c_int), ("k", C.c_int), ("version", C.c_int),
                ("precision", C.c_int), ("inp", C.c_void_p), ("freq", C.c_void_p),
                ("val", C.c_void_p), ("capacity", C.c_int64),
                ("random_seed", C.c_uint), ("drand48_seed", C.c_long)]


class SfftServerStats(C.Structure):
    _fields_ = [(name, C.c_int64) for name in (
        "submitted", "completed", "rejected", "failed", "batches",
        "batched_requests", "plans_created", "plans_evicted", "plans_resident",
        "hbm_budget", "hbm_in_use", "hbm_peak", "max_concurrent_batches")] + [
        ("plan_build_ms", C.c_double)]


class SfftOutput(C.Structure):
    _fields_ = [("count", C.c_int64), ("capacity", C.c_int64),
                ("freq", C.c_void_p), ("val", C.c_void_p)]


_lib = None


def load():
    """Returns the loaded library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsfft_amd.so is not built (run sfft_amd/build.py "
                           "or __graft_entry__.build()); no fallback exists")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    PP = C.POINTER(SfftPlan)
    L.sfft_malloc.argtypes = [C.c_size_t]
    L.sfft_malloc.restype = P
    L.sfft_free.argtypes = [P]
    L.sfft_make_plan.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.sfft_make_plan.restype = PP
    L.sfft_free_plan.argtypes = [PP]
    L.sfft_exec.argtypes = [PP, P, P]
    L.sfft_exec_many.argtypes = [PP, C.c_int, P, P]
    L.sfft_plan_output_capacity.argtypes = [PP]
    L.sfft_plan_output_capacity.restype = C.c_int64
    L.sfft_exec_sparse.argtypes = [PP, P, C.POINTER(SfftOutput)]
    L.sfft_exec_many_sparse.argtypes = [PP, C.c_int, P, C.POINTER(SfftOutput)]
    L.sfft_exec_sparse_noalloc.argtypes = [PP, P, C.POINTER(SfftOutput)]
    L.sfft_plan_debug_comb_votes.argtypes = [PP, C.c_int]
    L.sfft_plan_debug_v3.argtypes = [PP, C.c_int, C.c_int]
    L.sfft_plan_debug_v3_inline_pre.argtypes = [PP, C.c_int]
    L.sfft_debug_comb_votes.argtypes = [PP, C.c_int, P, C.c_int64]
    L.sfft_debug_comb_votes.restype = C.c_int64
    L.sfft_plan_set_precision.argtypes = [PP, C.c_int]
    L.sfft_plan_seed.argtypes = [PP, C.c_uint, C.c_long]
    L.sfft_plan_set_stream.argtypes = [PP, P]
    L.sfft_plan_set_reference_conjugate.argtypes = [PP, C.c_int]
    L.sfft_plan_memory.argtypes = [PP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.sfft_last_error.argtypes = []
    L.sfft_last_error.restype = C.c_char_p
    L.sfft_plan_params.argtypes = [PP, P]
    L.sfft_derive_params.argtypes = [C.c_int, C.c_int, C.c_int, P]
    L.sfft_design_filter.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, P,
                                     C.c_int, P]
    L.sfft_debug_buckets.argtypes = [PP, C.c_int, P, C.c_int64]
    L.sfft_debug_buckets.restype = C.c_int64
    L.sfft_plan_enable_timing.argtypes = [PP, C.c_int]
    L.sfft_plan_filter.argtypes = [PP, C.c_int, P, C.c_int, P, C.c_int, C.POINTER(C.c_int)]
    L.sfft_plan_stage_times.argtypes = [PP, P, P]
    L.sfft_plan_last_status.argtypes = [PP, P, C.c_int]
    L.sfft_plan_enable_graphs.argtypes = [PP, C.c_int]
    L.sfft_generate_input.argtypes = [C.c_int, C.c_int, C.c_long, C.c_double, P, P,
                                      C.POINTER(C.c_double)]
    L.sfft_exec_many_plans.argtypes = [C.c_int, P, P, C.POINTER(SfftOutput)]
    L.sfft_plan_set_schedule.argtypes = [PP, C.c_int, C.c_int]
    L.sfft_plan_get_schedule.argtypes = [PP, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.sfft_plan_autotune.argtypes = [PP, C.c_int]
    L.sfft_output_capacity.argtypes = [C.c_int, C.c_int, C.c_int]
    L.sfft_output_capacity.restype = C.c_int64
    L.sfft_server_create.argtypes = [C.c_int64, C.c_int, C.c_int]
    L.sfft_server_create.restype = P
    L.sfft_server_prepare.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
    L.sfft_server_submit.argtypes = [P, C.POINTER(SfftRequest)]
    L.sfft_server_submit.restype = C.c_int64
    L.sfft_server_poll.argtypes = [P, C.c_int64]
    L.sfft_server_wait.argtypes = [P, C.c_int64, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double)]
    L.sfft_server_get_stats.argtypes = [P, C.POINTER(SfftServerStats)]
    L.sfft_server_destroy.argtypes = [P]
    _lib = L
    return L


def last_error():
    return load().sfft_last_error().decode()

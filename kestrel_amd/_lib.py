"""ctypes binding of libkestrel_amd.so (the gfx950 kernels behind the C ABI of
include/kestrel_amd.h).  There is no CPU fallback: loading fails loudly when
the library has not been built, and every call reports the library's error
text when the device is missing or a HIP call fails."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkestrel_amd.so")

OK = 0
EINVAL = -1
EDEVICE = -2
ENOMEM = -3
EDATA = -4

PAIR_OK = 0
PAIR_NO_SITES = 1

FP64 = 0
MIXED = 1
MAF_DEFAULT = -1.0  # KESTREL_MAF_DEFAULT: the reference default 1/(2N)


class KestrelError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kestrel error {code}: {msg}")
        self.code = code


class Tune(ctypes.Structure):
    """sqp.rs:16-29 Tuneables."""
    _fields_ = [
        ("sqp_max_iter", ctypes.c_uint64),
        ("sqp_conv_tol", ctypes.c_double),
        ("qp_max_iter", ctypes.c_uint64),
        ("qp_conv_tol", ctypes.c_double),
        ("qp_zero_search_tol", ctypes.c_double),
        ("bls_max_iter", ctypes.c_uint64),
        ("bls_sufficient_decrease", ctypes.c_double),
        ("bls_step_size_reduce", ctypes.c_double),
        ("epsilon", ctypes.c_double),
    ]


# kestrel_output: coefficients.rs:143-151 Output + n_sites + status
OUTPUT_DTYPE = np.dtype([
    ("x", np.uint32), ("y", np.uint32), ("jacquard", np.float64, 9), ("obj", np.float64),
    ("iters", np.uint64), ("cond", np.float64), ("n_sites", np.uint32), ("status", np.uint32),
])
assert OUTPUT_DTYPE.itemsize == 112

# every symbol include/kestrel_amd.h declares
EXPORTS = (
    "kestrel_tune_default", "kestrel_cohort_create", "kestrel_cohort_allele_freqs", "kestrel_cohort_n_kept",
    "kestrel_fit_pairs", "kestrel_evaluate", "kestrel_set_option", "kestrel_last_fit_stats",
    "kestrel_cohort_destroy", "kestrel_last_error", "kestrel_version", "kestrel_selftest",
    "kestrel_measure_fp64_peak", "kestrel_measure_copy", "kestrel_cohort_create_gt", "kestrel_fit_pairs_gt",
    "kestrel_serve_start", "kestrel_submit_queries", "kestrel_poll", "kestrel_wait", "kestrel_serve_get_stats",
    "kestrel_serve_stop", "kestrel_last_fit_busy", "kestrel_cohort_create_gl", "kestrel_cohort_create_gl_multi",
    "kestrel_cohort_multi_info", "kestrel_cohort_create_gt_multi", "kestrel_measure_fp32_peak", "kestrel_autotune",
    "kestrel_measure_lds", "kestrel_last_fit_recondition", "kestrel_evaluate_noalloc",
    "kestrel_measure_hessian_form", "kestrel_last_fit_form",
)


class AutotuneResult(ctypes.Structure):
    """kestrel_autotune_result."""
    _fields_ = [
        ("slots", ctypes.c_int64),
        ("streams", ctypes.c_int64),
        ("batch", ctypes.c_int64),
        ("pairs_per_s", ctypes.c_double),
        ("default_pairs_per_s", ctypes.c_double),
        ("trials", ctypes.c_uint32),
    ]


class ServeStats(ctypes.Structure):
    """kestrel_serve_stats."""
    _fields_ = [
        ("requests_done", ctypes.c_uint64),
        ("pairs_done", ctypes.c_uint64),
        ("sweeps", ctypes.c_uint64),
        ("busy_slot_sweeps", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("slots", ctypes.c_uint32),
        ("live", ctypes.c_uint32),
        ("pending", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

_lib = None
P = ctypes.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C kestrel_amd)")
    L = ctypes.CDLL(LIB_PATH)
    L.kestrel_last_error.restype = ctypes.c_char_p
    L.kestrel_version.restype = ctypes.c_char_p
    L.kestrel_tune_default.argtypes = [P, ctypes.c_uint64]
    L.kestrel_cohort_create.argtypes = [P, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t, P, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_int]
    L.kestrel_cohort_allele_freqs.argtypes = [P, P, P]
    L.kestrel_cohort_n_kept.argtypes = [P, P]
    L.kestrel_fit_pairs.argtypes = [P, P, P, ctypes.c_size_t, ctypes.c_uint32, P]
    L.kestrel_evaluate.argtypes = [P, P, P, ctypes.c_size_t, ctypes.c_double, P, P, P, P]
    L.kestrel_evaluate_noalloc.argtypes = [P, P, P, ctypes.c_size_t, ctypes.c_double, P, P, P, P]
    L.kestrel_set_option.argtypes = [P, ctypes.c_char_p, ctypes.c_int64]
    L.kestrel_last_fit_stats.argtypes = [P, P, P, P]
    L.kestrel_last_fit_busy.argtypes = [P, P, P]
    L.kestrel_last_fit_recondition.argtypes = [P, P, P]
    L.kestrel_last_fit_form.argtypes = [P, P]
    L.kestrel_selftest.argtypes = [ctypes.c_uint64, ctypes.c_uint64, P]
    L.kestrel_cohort_create_gt.argtypes = [P, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t, P]
    L.kestrel_cohort_create_gl.argtypes = [P, P, ctypes.c_uint32, ctypes.c_uint32, P, ctypes.c_double,
                                           ctypes.c_uint64]
    L.kestrel_cohort_create_gl_multi.argtypes = [P, ctypes.c_uint32, P, ctypes.c_uint32, P, ctypes.c_uint32, P,
                                                 ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64]
    L.kestrel_cohort_multi_info.argtypes = [P, P, P, P, P]
    L.kestrel_cohort_create_gt_multi.argtypes = [P, ctypes.c_uint32, P, ctypes.c_uint32, ctypes.c_size_t, P,
                                                 ctypes.c_uint32, P, ctypes.c_uint32, P, P, P]
    L.kestrel_fit_pairs_gt.argtypes = [P, ctypes.c_uint64, P, ctypes.c_size_t, ctypes.c_uint32, P, P]
    L.kestrel_measure_fp64_peak.argtypes = [P]
    L.kestrel_measure_fp32_peak.argtypes = [P]
    L.kestrel_measure_lds.argtypes = [P]
    L.kestrel_measure_hessian_form.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, P, P, P]
    L.kestrel_autotune.argtypes = [P, P, P, ctypes.c_size_t, ctypes.c_uint64, P]
    L.kestrel_measure_copy.argtypes = [ctypes.c_uint64, P]
    L.kestrel_serve_start.argtypes = [P, P, ctypes.c_uint32, ctypes.c_uint64]
    L.kestrel_submit_queries.argtypes = [P, P, ctypes.c_uint32, ctypes.c_size_t, P, ctypes.c_size_t, P, P]
    L.kestrel_poll.argtypes = [P, ctypes.c_uint64, P, P]
    L.kestrel_wait.argtypes = [P, ctypes.c_uint64, P]
    L.kestrel_serve_get_stats.argtypes = [P, P]
    L.kestrel_serve_stop.argtypes = [P]
    L.kestrel_cohort_destroy.argtypes = [P]
    L.kestrel_cohort_destroy.restype = None
    _lib = L
    return L


def check(rc: int):
    if rc != OK:
        raise KestrelError(rc, load().kestrel_last_error().decode())


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None

"""The C ABI library loads without a GPU and exports every symbol the header declares."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "kestrel_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kestrel_\w+)\s*\(", src)))


def test_header_and_binding_agree():
    from kestrel_amd import _lib
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from kestrel_amd import _lib
    L = _lib.load()
    for name in declared_functions():
        assert hasattr(L, name), name
        ctypes.cast(getattr(L, name), ctypes.c_void_p)


def test_host_only_calls_without_gpu():
    from kestrel_amd import _lib, tune_default
    t = tune_default(100)
    assert t.sqp_max_iter == 100 and t.qp_max_iter == 10 and t.epsilon == 1e-8
    assert t.bls_max_iter == 10 and t.bls_step_size_reduce == 0.9 and t.sqp_conv_tol == 1e-8
    L = _lib.load()
    assert b"gfx950" in L.kestrel_version()
    # invalid arguments are rejected before any device call
    h = ctypes.c_void_p()
    packed = np.zeros((2, 4), np.uint8)
    rc = L.kestrel_cohort_create(ctypes.byref(h), _lib.ptr(packed), 2, 0, 4, None, -1.0, 100, 0)
    assert rc == _lib.EINVAL
    rc = L.kestrel_cohort_create(ctypes.byref(h), _lib.ptr(packed), 2, 8, 4, None, 2.0, 100, 0)
    assert rc == _lib.EINVAL and b"Minimum allele frequency" in L.kestrel_last_error()
    assert L.kestrel_tune_default(ctypes.byref(t), 0) == _lib.EINVAL

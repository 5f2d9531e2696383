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


def test_every_entry_point_rejects_null_arguments():
    """Every int-returning entry point, called with null pointers and zero
    sizes (a null cohort where it takes one), returns an error code and
    leaves a message; none dereferences a null pointer (the checks come
    before any device call, so this runs without a GPU)."""
    from kestrel_amd import _lib
    L = _lib.load()
    skip = {"kestrel_version", "kestrel_last_error"}
    for name in declared_functions():
        if name in skip:
            continue
        fn = getattr(L, name)
        assert fn.argtypes is not None, name  # the binding declares every argument
        args = []
        for t in fn.argtypes or []:
            if t in (ctypes.c_double, ctypes.c_float):
                args.append(0.0)
            elif t is ctypes.c_char_p:
                args.append(None)
            elif isinstance(t, type) and issubclass(t, (ctypes._SimpleCData,)) and t not in (ctypes.c_void_p,):
                args.append(0)
            else:
                args.append(None)
        rc = fn(*args)
        assert rc != 0, name
        assert L.kestrel_last_error(), name

"""The Hessian accumulation in its vector and matrix-core forms
(kestrel_measure_hessian_form): on the same FP32 rows every form's 45 terms
and 9 gradient sums must agree with FP64 sums of those rows within the bound
of its arithmetic -- FP64 FMA, FP32 FMA over 4096-site segments, bf16 hi/lo
matrix-core products (truncation leaves |c' - hi| < 2^-7 |c'| and
|c' - hi - lo| < 2^-14 |c'|, and lo lo^T is dropped: a product misses less
than 3 2^-14 of itself, then FP32 sums) and the hi/mid/lo split (3 2^-21).
The matrix-core forms also check the operand and result lane layout of
v_mfma_f32_16x16x32_bf16 that they assume: a wrong layout gives errors of
order 1."""
import ctypes

import numpy as np
import pytest

from kestrel_amd import _lib

pytestmark = pytest.mark.gpu

N_PAIRS, N_SITES = 256, 8192
BOUNDS = {0: 1e-12, 1: 2e-5, 2: 2.5e-4, 3: 1e-5, 6: 2e-5}  # relative to sqrt(H_ii H_jj)


def say(*a):
    print(*a, flush=True)


def reference(table):
    w = (np.float32(1.0) + (np.arange(N_PAIRS) % 64).astype(np.float32) / np.float32(64.0)).astype(np.float32)
    sites = np.arange(N_SITES) % 256
    r = (table[sites, 9][None, :] * w[:, None]).astype(np.float32)
    cp = (table[sites, :9][None, :, :] * r[:, :, None]).astype(np.float32).astype(np.float64)
    return np.einsum("psi,psj->pij", cp, cp), cp.sum(1)


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4, 5, 6])
def test_hessian_form_matches_fp64_sums(form):
    say(f"case: hessian form {form}, {N_PAIRS} pairs x {N_SITES} sites")
    rng = np.random.default_rng(0x4E55 + form)
    table = np.zeros((256, 12), np.float32)
    table[:, :9] = rng.uniform(0.0, 1.0, (256, 9)).astype(np.float32)
    table[:, 9] = (1.0 / rng.uniform(0.01, 1.0, 256)).astype(np.float32)
    table[rng.integers(0, 256, 20), rng.integers(0, 9, 20)] = 0.0  # structural zeros of c
    nseg = N_SITES // 4096
    out = np.zeros((nseg, N_PAIRS, 54), np.float64 if form == 0 else np.float32)
    ms = ctypes.c_double()
    L = _lib.load()
    _lib.check(L.kestrel_measure_hessian_form(form, N_PAIRS, N_SITES, table.ctypes.data_as(ctypes.c_void_p),
                                              out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms)))
    href, gref = reference(table)
    tot = out.astype(np.float64).sum(0)
    il, jl = np.tril_indices(9)
    order = np.lexsort((jl, il))
    il, jl = il[order], jl[order]
    if form in (4, 5):
        assert (tot[:, :45] == 0).all()
    else:
        scale = np.sqrt(href[:, il, il] * href[:, jl, jl])
        err = np.max(np.abs(tot[:, :45] - href[:, il, jl]) / scale)
        say(f"form {form}: max relative Hessian error {err:.3g} ({ms.value:.3f} ms)")
        assert err <= BOUNDS[form], err
    gerr = np.max(np.abs(tot[:, 45:] - gref) / gref)
    assert gerr <= (1e-12 if form == 0 else 2e-5), gerr


def test_hessian_form_rejects_bad_shapes():
    L = _lib.load()
    table = np.zeros((256, 12), np.float32)
    out = np.zeros(54 * 256, np.float32)
    ms = ctypes.c_double()
    for form, n_pairs, n_sites in ((7, 256, 4096), (1, 100, 4096), (1, 256, 1000)):
        rc = L.kestrel_measure_hessian_form(form, n_pairs, n_sites, table.ctypes.data_as(ctypes.c_void_p),
                                            out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms))
        assert rc != 0

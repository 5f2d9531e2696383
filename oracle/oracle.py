"""ctypes binding of the oracle library (oracle/_build/libkestrel_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libkestrel_oracle.so")

ORDER_REF = 0
ORDER_SEG = 1
ORDER_FAST = 2  # ORDER_REF in the reference's AVX-512 loop shape (same bits; timing form)
ORDER_LANES = 3  # the segments of ORDER_SEG summed by 64 strided lanes and a butterfly
ORDER_CLASS = 4  # the class-sorted sweep's order: genotype-pair patterns per 8192-site block (ko_class.c)
CLASS_BLOCK = 8192


class Tune(ctypes.Structure):
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


OUTPUT_DTYPE = np.dtype([
    ("x", np.uint32), ("y", np.uint32), ("jacquard", np.float64, 9), ("obj", np.float64),
    ("iters", np.uint64), ("cond", np.float64), ("n_sites", np.uint32), ("pad", np.uint32),
])
assert OUTPUT_DTYPE.itemsize == 112

_lib = None

D = ctypes.c_double
P = ctypes.c_void_p
SZ = ctypes.c_size_t
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I = ctypes.c_int


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ko_log.restype = D
        L.ko_log.argtypes = [D]
        L.ko_log_getmant.restype = D
        L.ko_log_getmant.argtypes = [D]
        L.ko_lane_log8_fast.restype = D
        L.ko_lane_log8_fast.argtypes = [D, I]
        L.ko_has_avx512.restype = I
        L.ko_isa_name.restype = ctypes.c_char_p
        L.ko_tune_new.argtypes = [P, U64, U64]
        L.ko_mixture_row2.argtypes = [P, P, P, P]
        L.ko_mixture_generic2.argtypes = [P, P, P, P]
        L.ko_conditional_probabilities.argtypes = [D, D, D, D, I, P]
        L.ko_iis_mode.argtypes = [I, I, I, I]
        L.ko_iis_mode.restype = I
        L.ko_jacq_obj_ref.restype = D
        L.ko_jacq_obj_ref.argtypes = [P, SZ, P, D]
        L.ko_jacq_gh_ref.argtypes = [P, SZ, P, D, P, P]
        L.ko_jacq_obj_seg.restype = D
        L.ko_jacq_obj_seg.argtypes = [P, P, SZ, P, SZ, P, D]
        L.ko_jacq_gh_seg.argtypes = [P, P, SZ, P, SZ, P, D, P, P]
        L.ko_jacq_obj_lanes.restype = D
        L.ko_jacq_obj_lanes.argtypes = [P, P, SZ, P, SZ, P, D]
        L.ko_jacq_gh_lanes.argtypes = [P, P, SZ, P, SZ, P, D, P, P]
        L.ko_class_sums.argtypes = [P, P, P, P, SZ, P, D, P, P, P, P]
        L.ko_class_expand.argtypes = [P, P, P, P]
        L.ko_modify_gmw.argtypes = [I, P, P, P]
        L.ko_unmodified.argtypes = [I, P, P, P]
        L.ko_solve_ldl.argtypes = [I, P, P, P]
        L.ko_solve_qp.restype = U64
        L.ko_solve_qp.argtypes = [I, P, P, P, I, P, P]
        L.ko_condition.restype = D
        L.ko_condition.argtypes = [I, P, P, P, D]
        L.ko_fit_pair_rows.argtypes = [I, P, SZ, P, P, SZ, I, P, P]
        L.ko_pair_rows_codes.restype = SZ
        L.ko_pair_rows_codes.argtypes = [P, P, P, SZ, P, P]
        L.ko_allele_freqs_codes.restype = I
        L.ko_allele_freqs_codes.argtypes = [P, U32, U32, U64, D, I, P]
        L.ko_allele_freq2_rows.restype = I
        L.ko_allele_freq2_rows.argtypes = [P, SZ, U64, D, P, P]
        L.ko_fit_pairs_codes.restype = I
        L.ko_fit_pairs_codes.argtypes = [I, P, U32, U32, P, P, SZ, U32, P, SZ, U64, I, P]
        L.ko_allele_freqs_gl.restype = I
        L.ko_allele_freqs_gl.argtypes = [P, U32, U32, U64, D, I, P]
        L.ko_fit_pairs_gl.restype = I
        L.ko_fit_pairs_gl.argtypes = [I, P, U32, U32, P, P, SZ, U32, P, SZ, U64, I, P]
        L.ko_pair_rows_gl.restype = SZ
        L.ko_pair_rows_gl.argtypes = [P, P, P, SZ, P, P]
        L.ko_mixture_generic.restype = None
        L.ko_mixture_generic.argtypes = [I, P, P, P, P]
        L.ko_allele_freqA_rows.restype = I
        L.ko_allele_freqA_rows.argtypes = [I, P, SZ, U64, D, P, P]
        L.ko_pair_rows_generic.restype = SZ
        L.ko_pair_rows_generic.argtypes = [I, P, P, P, SZ, P, P]
        L.ko_fit_pair_class_parts.restype = None
        L.ko_fit_pair_class_parts.argtypes = [P, P, P, U32, P, P, P, SZ, I, P, P]
        L.ko_class_parts_eval.restype = None
        L.ko_class_parts_eval.argtypes = [P, P, P, U32, P, P, P, SZ, P, D, P, P, P]
        L.ko_fit_pairs_multi.restype = I
        L.ko_fit_pairs_multi.argtypes = [U32, P, P, U32, P, P, U32, P, P, U32, P, SZ, U64, I, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def tune(sqp_max_iter=100, qp_max_iter=10) -> Tune:
    t = Tune()
    lib().ko_tune_new(ctypes.byref(t), sqp_max_iter, qp_max_iter)
    return t


def isa() -> str:
    return lib().ko_isa_name().decode()


def log(d: float) -> float:
    return lib().ko_log(d)


def log_getmant(d: float) -> float:
    return lib().ko_log_getmant(d)


def mixture_row2(lx, ly, af) -> np.ndarray:
    out = np.zeros(9)
    lx, ly, af = f64(lx), f64(ly), f64(af)
    lib().ko_mixture_row2(_p(lx), _p(ly), _p(af), _p(out))
    return out


def mixture_generic2(lx, ly, af) -> np.ndarray:
    out = np.zeros(9)
    lx, ly, af = f64(lx), f64(ly), f64(af)
    lib().ko_mixture_generic2(_p(lx), _p(ly), _p(af), _p(out))
    return out


def conditional_probabilities(pi, pj, pk, pl, mode) -> np.ndarray:
    out = np.zeros(9)
    lib().ko_conditional_probabilities(pi, pj, pk, pl, mode, _p(out))
    return out


def iis_mode(i, j, k, l) -> int:
    return lib().ko_iis_mode(i, j, k, l)


def obj_ref(rows, x, eps=1e-8) -> float:
    rows, x = f64(rows), f64(x)
    return lib().ko_jacq_obj_ref(_p(rows), rows.shape[0], _p(x), eps)


def gh_ref(rows, x, eps=1e-8):
    rows, x = f64(rows), f64(x)
    g, h = np.zeros(9), np.zeros((9, 9))
    lib().ko_jacq_gh_ref(_p(rows), rows.shape[0], _p(x), eps, _p(g), _p(h))
    return g, h


def _seg(seg):
    return np.ascontiguousarray(seg, dtype=np.uint32)


def obj_seg(rows, valid, seg, x, eps=1e-8) -> float:
    rows, x, valid, seg = f64(rows), f64(x), np.ascontiguousarray(valid, np.uint8), _seg(seg)
    return lib().ko_jacq_obj_seg(_p(rows), _p(valid), rows.shape[0], _p(seg), len(seg) - 1, _p(x), eps)


def gh_seg(rows, valid, seg, x, eps=1e-8):
    rows, x, valid, seg = f64(rows), f64(x), np.ascontiguousarray(valid, np.uint8), _seg(seg)
    g, h = np.zeros(9), np.zeros((9, 9))
    lib().ko_jacq_gh_seg(_p(rows), _p(valid), rows.shape[0], _p(seg), len(seg) - 1, _p(x), eps, _p(g), _p(h))
    return g, h


def obj_lanes(rows, valid, seg, x, eps=1e-8) -> float:
    rows, x, valid, seg = f64(rows), f64(x), np.ascontiguousarray(valid, np.uint8), _seg(seg)
    return lib().ko_jacq_obj_lanes(_p(rows), _p(valid), rows.shape[0], _p(seg), len(seg) - 1, _p(x), eps)


def gh_lanes(rows, valid, seg, x, eps=1e-8):
    rows, x, valid, seg = f64(rows), f64(x), np.ascontiguousarray(valid, np.uint8), _seg(seg)
    g, h = np.zeros(9), np.zeros((9, 9))
    lib().ko_jacq_gh_lanes(_p(rows), _p(valid), rows.shape[0], _p(seg), len(seg) - 1, _p(x), eps, _p(g), _p(h))
    return g, h


def modify_gmw(a):
    a = f64(a)
    n = a.shape[0]
    l, d = np.zeros((n, n)), np.zeros(n)
    lib().ko_modify_gmw(n, _p(a), _p(l), _p(d))
    return l, d


def unmodified(a):
    a = f64(a)
    n = a.shape[0]
    l, d = np.zeros((n, n)), np.zeros(n)
    lib().ko_unmodified(n, _p(a), _p(l), _p(d))
    return l, d


def solve_ldl(l, d, b):
    l, d, b = f64(l), f64(d), f64(b).copy()
    lib().ko_solve_ldl(l.shape[0], _p(l), _p(d), _p(b))
    return b


def solve_qp(q, c, y0, modify=True, t: Tune | None = None):
    q, c, y0 = f64(q), f64(c), f64(y0)
    n = q.shape[0]
    y = np.zeros(n)
    t = t or tune(100, 10)
    it = lib().ko_solve_qp(n, _p(q), _p(c), _p(y0), int(modify), ctypes.byref(t), _p(y))
    return y, it


def condition(x, g, h, tol=1e-8) -> float:
    x, g, h = f64(x), f64(g), f64(h)
    return lib().ko_condition(len(x), _p(x), _p(g), _p(h), tol)


def fit_rows(rows, self_pair=False, order=ORDER_REF, valid=None, seg=None, iterations=100):
    """Fit one pair from its c rows: compacted (REF) or per-site with valid (SEG, LANES)."""
    rows = f64(rows)
    out = np.zeros(1, OUTPUT_DTYPE)
    t = tune(iterations, 10)
    if order == ORDER_REF:
        lib().ko_fit_pair_rows(ORDER_REF, _p(rows), rows.shape[0], None, None, 0, int(self_pair),
                               ctypes.byref(t), _p(out))
    else:
        valid = np.ascontiguousarray(valid, np.uint8)
        seg = _seg(seg)
        lib().ko_fit_pair_rows(order, _p(rows), rows.shape[0], _p(valid), _p(seg), len(seg) - 1,
                               int(self_pair), ctypes.byref(t), _p(out))
    return out[0]


def class_sums(gx, gy, af, x, eps=1e-8, seg=None):
    """One pair's sums in the class order: (F, G[9], H[45], n) with F = sum log L,
    G = sum c/L, H = sum c c^T / L^2 (upper triangle, row-major)."""
    gx, gy = np.ascontiguousarray(gx, np.uint8), np.ascontiguousarray(gy, np.uint8)
    af, x = f64(af), f64(x)
    if seg is None:
        seg = np.array([0, gx.shape[0]], np.uint32)
    seg = _seg(seg)
    F = np.zeros(1)
    G = np.zeros(9)
    H = np.zeros(45)
    n = np.zeros(1, np.uint64)
    lib().ko_class_sums(_p(gx), _p(gy), _p(af), _p(seg), len(seg) - 1, _p(x), eps, _p(F), _p(G), _p(H), _p(n))
    return float(F[0]), G, H, int(n[0])


def class_expand(R):
    """The expansion of a segment's 47 pattern sums into (F, G[9], H[45])."""
    R = f64(R)
    F = np.zeros(1)
    G = np.zeros(9)
    H = np.zeros(45)
    lib().ko_class_expand(_p(R), _p(F), _p(G), _p(H))
    return float(F[0]), G, H


def pair_rows_codes(gx, gy, af):
    gx, gy = np.ascontiguousarray(gx, np.uint8), np.ascontiguousarray(gy, np.uint8)
    af = f64(af)
    v = gx.shape[0]
    rows = np.zeros((v, 9))
    valid = np.zeros(v, np.uint8)
    n = lib().ko_pair_rows_codes(_p(gx), _p(gy), _p(af), v, _p(rows), _p(valid))
    return rows, valid, n


def allele_freqs(codes, iterations=100, maf=None, nthreads=0):
    """allele/mod.rs:37-134 on one-hot likelihoods of the codes (N, V)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    n, v = codes.shape
    if maf is None:
        maf = 1.0 / (2.0 * n)  # wrapper.rs:101
    af = np.zeros((v, 2))
    rc = lib().ko_allele_freqs_codes(_p(codes), n, v, iterations, maf, nthreads, _p(af))
    if rc != 0:
        raise ValueError("a site has no called sample")
    return af


def allele_freq_rows(gl, iterations=100, maf=0.0):
    gl = f64(gl)
    af = np.zeros(2)
    it = np.zeros(1, np.uint64)
    rc = lib().ko_allele_freq2_rows(_p(gl), gl.shape[0], iterations, maf, _p(af), _p(it))
    if rc != 0:
        raise ValueError("no called sample")
    return af, int(it[0])


def keep_sites(af) -> np.ndarray:
    """wrapper.rs:352-410: a biallelic site survives iff both frequencies are > 0."""
    return (af[:, 0] > 0.0) & (af[:, 1] > 0.0)


def fit_pairs(codes, af, pairs=None, nsel=None, order=ORDER_REF, seg=None, iterations=100, nthreads=0):
    """Pair fits over kept sites: codes (N, V) and af (V, 2) restricted to kept sites."""
    codes = np.ascontiguousarray(codes, np.uint8)
    af = f64(af)
    n, v = codes.shape
    if nsel is None:
        nsel = n
    if pairs is not None:
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        npairs = pairs.shape[0]
    else:
        npairs = nsel * (nsel + 1) // 2
    out = np.zeros(npairs, OUTPUT_DTYPE)
    if seg is None:
        seg = np.array([0, v], np.uint32)
    seg = _seg(seg)
    rc = lib().ko_fit_pairs_codes(order, _p(codes), n, v, _p(af), _p(pairs), npairs, nsel, _p(seg), len(seg) - 1,
                                  iterations, nthreads, _p(out))
    if rc != 0:
        raise ValueError("requested samples exceed total samples")
    return out


def allele_freqs_gl(gl, iterations=100, maf=None, nthreads=0):
    """allele/mod.rs:37-134 on likelihood rows gl (N, V, 3), NaN = missing."""
    gl = f64(gl)
    n, v, _ = gl.shape
    if maf is None:
        maf = 1.0 / (2.0 * n)  # wrapper.rs:101
    af = np.zeros((v, 2))
    rc = lib().ko_allele_freqs_gl(_p(gl), n, v, iterations, maf, nthreads, _p(af))
    if rc != 0:
        raise ValueError("a site has no called sample")
    return af


def pair_rows_gl(lx, ly, af):
    lx, ly, af = f64(lx), f64(ly), f64(af)
    v = lx.shape[0]
    rows = np.zeros((v, 9))
    valid = np.zeros(v, np.uint8)
    n = lib().ko_pair_rows_gl(_p(lx), _p(ly), _p(af), v, _p(rows), _p(valid))
    return rows, valid, n


def fit_pairs_gl(gl, af, pairs=None, nsel=None, order=ORDER_REF, seg=None, iterations=100, nthreads=0):
    """Pair fits over kept sites from likelihood rows gl (N, V, 3) and af (V, 2)."""
    gl = f64(gl)
    af = f64(af)
    n, v, _ = gl.shape
    if nsel is None:
        nsel = n
    if pairs is not None:
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        npairs = pairs.shape[0]
    else:
        npairs = nsel * (nsel + 1) // 2
    out = np.zeros(npairs, OUTPUT_DTYPE)
    if seg is None:
        seg = np.array([0, v], np.uint32)
    seg = _seg(seg)
    rc = lib().ko_fit_pairs_gl(order, _p(gl), n, v, _p(af), _p(pairs), npairs, nsel, _p(seg), len(seg) - 1,
                               iterations, nthreads, _p(out))
    if rc != 0:
        raise ValueError("requested samples exceed total samples")
    return out


def mixture_generic(A, lx, ly, af) -> np.ndarray:
    """mixture.rs:8-60 over IIS_A for one site."""
    out = np.zeros(9)
    lx, ly, af = f64(lx), f64(ly), f64(af)
    lib().ko_mixture_generic(A, _p(lx), _p(ly), _p(af), _p(out))
    return out


def allele_freqA_rows(A, gl, iterations=100, maf=0.0):
    """allele/mod.rs:78-134 for one site with A alleles: gl (n, A(A+1)/2)."""
    gl = f64(gl)
    af = np.zeros(A)
    it = np.zeros(1, np.uint64)
    rc = lib().ko_allele_freqA_rows(A, _p(gl), gl.shape[0], iterations, maf, _p(af), _p(it))
    if rc != 0:
        raise ValueError("no called sample")
    return af, int(it[0])


def pair_rows_generic(A, lx, ly, af):
    lx, ly, af = f64(lx), f64(ly), f64(af)
    v = lx.shape[0]
    rows = np.zeros((v, 9))
    valid = np.zeros(v, np.uint8)
    n = lib().ko_pair_rows_generic(A, _p(lx), _p(ly), _p(af), v, _p(rows), _p(valid))
    return rows, valid, n


def _gl_index(a, b):
    i, j = min(a, b), max(a, b)
    return j * (j + 1) // 2 + i


def multi_prepare(buckets, iterations=100, maf=None):
    """wrapper.rs:89-130 + split_gl_af / remove_zero_frequencies (wrapper.rs:237-410)
    for likelihood buckets {2: (V2, N, 3), 3: (V3, N, 6), 4: (V4, N, 10)}
    (site-major as vcf.rs builds them): allele frequencies per site, then
    alleles with frequency 0 removed and every site re-bucketed by its
    remaining allele count (processing the 2-, 3- then 4-allele buckets in
    site order; one allele left drops the site).  Returns {A: (af (V'_A, A),
    gl (N, V'_A, G'_A) sample-major)} and the frequencies per input bucket."""
    n = next(b.shape[1] for b in buckets.values() if b is not None and b.size)
    if maf is None:
        maf = 1.0 / (2.0 * n)  # wrapper.rs:101
    af_in, out_af, out_gl = {}, {2: [], 3: [], 4: []}, {2: [], 3: [], 4: []}
    for A in (2, 3, 4):
        b = buckets.get(A)
        if b is None or b.shape[0] == 0:
            af_in[A] = np.zeros((0, A))
            continue
        af_in[A] = np.array([allele_freqA_rows(A, b[s], iterations, maf)[0] for s in range(b.shape[0])])
        for s in range(b.shape[0]):
            nz = [a for a in range(A) if af_in[A][s, a] > 0.0]
            if len(nz) < 2:
                continue
            idx = [_gl_index(nz[i], nz[j]) for j in range(len(nz)) for i in range(j + 1)]
            out_af[len(nz)].append(af_in[A][s, nz])
            out_gl[len(nz)].append(b[s][:, idx])
    res = {}
    for A in (2, 3, 4):
        G = A * (A + 1) // 2
        res[A] = (np.array(out_af[A]).reshape(-1, A),
                  np.ascontiguousarray(np.array(out_gl[A]).reshape(-1, n, G).transpose(1, 0, 2)))
    return res, af_in


def multi_pair_rows(parts, x, y, order, pads=None):
    """The c rows of pair (x, y) over the re-bucketed sites.  REF: the order of
    coefficients.rs:97-107 (the 8-lane parts of the 2-, 3-, 4-allele sites,
    then their remainders), compacted to the usable rows.  SEG: the device's
    order, each part padded to pads[A] sites; returns (rows, valid)."""
    chunks = {}
    for A in (2, 3, 4):
        af, gl = parts[A]
        if af.shape[0] == 0:
            chunks[A] = (np.zeros((0, 9)), np.zeros(0, np.uint8))
            continue
        rows, valid, _ = pair_rows_generic(A, gl[x], gl[y], af) if A > 2 else pair_rows_gl(gl[x], gl[y], af)
        chunks[A] = (rows, valid)
    if order == ORDER_REF:
        seq = []
        for A in (2, 3, 4):
            r, v = chunks[A]
            m = (len(v) // 8) * 8
            seq.append((r[:m], v[:m]))
        for A in (2, 3, 4):
            r, v = chunks[A]
            m = (len(v) // 8) * 8
            seq.append((r[m:], v[m:]))
        rows = np.vstack([r for r, _ in seq])
        valid = np.concatenate([v for _, v in seq])
        return rows[valid == 1], None
    rows, valid = [], []
    for A in (2, 3, 4):
        r, v = chunks[A]
        pad = pads[A] if pads and A in pads else len(v)
        if pad == 0:
            continue
        rr = np.zeros((pad, 9))
        vv = np.zeros(pad, np.uint8)
        rr[:len(v)] = r
        vv[:len(v)] = v
        rows.append(rr)
        valid.append(vv)
    return np.vstack(rows), np.concatenate(valid)


def multi_fit_pairs(parts, pairs, order=ORDER_REF, pads=None, seg=None, iterations=100):
    """Pair fits over re-bucketed multi-allelic sites (coefficients.rs:40-141)."""
    out = np.zeros(len(pairs), OUTPUT_DTYPE)
    for i, (x, y) in enumerate(pairs):
        rows, valid = multi_pair_rows(parts, int(x), int(y), order, pads)
        n = len(rows) if order == ORDER_REF else int(valid.sum())
        if n == 0:
            o = out[i]
            o["x"], o["y"], o["n_sites"] = x, y, 0
            o["jacquard"], o["obj"], o["cond"] = np.nan, np.nan, np.nan
            continue
        if order == ORDER_REF:
            o = fit_rows(rows, self_pair=x == y, order=ORDER_REF, iterations=iterations)
        else:
            o = fit_rows(rows, self_pair=x == y, order=ORDER_SEG, valid=valid, seg=seg, iterations=iterations)
        out[i] = o
        out[i]["x"], out[i]["y"] = x, y
    return out


def _class_parts_inputs(parts, pads, seg):
    af2, gl2 = parts[2]
    called = ~np.isnan(gl2).any(2)
    one_hot = (np.nan_to_num(gl2) == 1.0).sum(2) == 1
    if not (one_hot | ~called).all():
        raise ValueError("the class form takes 2-allele rows that are one-hot where called")
    codes = np.where(called, np.argmax(np.nan_to_num(gl2), 2), 3).astype(np.uint8)
    p2 = int(pads[2])
    seg = np.asarray(seg, np.int64)
    i0 = int(np.nonzero(seg == p2)[0][0])
    return codes, f64(af2), p2, _seg(seg[i0:] - p2)


def multi_class_eval(parts, x, y, pads, seg, pt, eps=1e-8):
    """f, g, h of pair (x, y) at pt in multi_class_fit_pairs' order."""
    codes, af2, p2, pseg = _class_parts_inputs(parts, pads, seg)
    rows, valid = multi_pair_rows(parts, int(x), int(y), ORDER_SEG, pads)
    prows, pvalid = f64(rows[p2:]), np.ascontiguousarray(valid[p2:], np.uint8)
    gx, gy = np.ascontiguousarray(codes[int(x)]), np.ascontiguousarray(codes[int(y)])
    f, g, h = np.zeros(1), np.zeros(9), np.zeros(81)
    lib().ko_class_parts_eval(_p(gx), _p(gy), _p(af2), af2.shape[0], _p(prows), _p(pvalid), _p(pseg), len(pseg) - 1,
                              _p(f64(pt)), eps, _p(f), _p(g), _p(h))
    return float(f[0]), g, h.reshape(9, 9)


def multi_class_fit_pairs(parts, pairs, pads, seg, iterations=100):
    """Pair fits over re-bucketed multi-allelic sites of hard calls (one-hot
    rows) in the device's class form: the 2-allele part as codes (the one-hot
    row's genotype index, 3 = missing) in the class order (ORDER_CLASS, one
    segment), then the 3- and 4-allele parts' segments as ORDER_SEG takes them
    (pads and seg of the whole SEG layout; the 2-allele part's padded rows
    come first), folded left to right after it."""
    codes, af2, p2, pseg = _class_parts_inputs(parts, pads, seg)
    t = tune(iterations, 10)
    out = np.zeros(len(pairs), OUTPUT_DTYPE)
    for i, (x, y) in enumerate(pairs):
        rows, valid = multi_pair_rows(parts, int(x), int(y), ORDER_SEG, pads)
        prows, pvalid = f64(rows[p2:]), np.ascontiguousarray(valid[p2:], np.uint8)
        gx, gy = np.ascontiguousarray(codes[int(x)]), np.ascontiguousarray(codes[int(y)])
        o = np.zeros(1, OUTPUT_DTYPE)
        lib().ko_fit_pair_class_parts(_p(gx), _p(gy), _p(af2), af2.shape[0], _p(prows), _p(pvalid), _p(pseg),
                                      len(pseg) - 1, int(x == y), ctypes.byref(t), _p(o))
        if o[0]["n_sites"] == 0:
            o[0]["jacquard"], o[0]["obj"], o[0]["cond"] = np.nan, np.nan, np.nan
        out[i] = o[0]
        out[i]["x"], out[i]["y"] = x, y
    return out


def multi_fit_pairs_threads(parts, pairs, iterations=100, nthreads=0):
    """multi_fit_pairs in the reference's order, in C on nthreads threads
    (ko_fit_pairs_multi); parts as multi_prepare returns them."""
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    n = parts[2][1].shape[0]  # every part's rows are (N, V_A, G_A), V_A possibly 0
    args = []
    for A in (2, 3, 4):
        af, gl = parts[A]
        af, gl = f64(af), f64(gl)
        args += [gl, af, af.shape[0]]
    out = np.zeros(len(pairs), OUTPUT_DTYPE)
    rc = lib().ko_fit_pairs_multi(n, _p(args[0]), _p(args[1]), args[2], _p(args[3]), _p(args[4]), args[5],
                                  _p(args[6]), _p(args[7]), args[8], _p(pairs), len(pairs), iterations, nthreads,
                                  _p(out))
    if rc != 0:
        raise ValueError("a pair names a sample outside the cohort")
    return out


def gt_allele_counts(codes):
    """wrapper.rs:412-428 allele frequencies by counting (GT path)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    n, v = codes.shape
    af = np.zeros((v, 2))
    L = lib()
    L.ko_gt_allele_counts.argtypes = [P, U32, U32, P]
    L.ko_gt_allele_counts(_p(codes), n, v, _p(af))
    return af


def cls_fit_pairs(codes, af=None, pairs=None, nsel=None, iterations=100, nthreads=0):
    """coefficients.rs:166-278 (CLS / Brier estimator of -f GT) for biallelic codes
    without missing calls.  Returns (outputs, Q)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    n, v = codes.shape
    if nsel is None:
        nsel = n
    if pairs is not None:
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        npairs = pairs.shape[0]
    else:
        npairs = nsel * (nsel + 1) // 2
    afp = f64(af) if af is not None else None
    out = np.zeros(npairs, OUTPUT_DTYPE)
    q = np.zeros((9, 9))
    L = lib()
    L.ko_cls_fit_pairs.restype = I
    L.ko_cls_fit_pairs.argtypes = [P, U32, U32, P, P, SZ, U32, U64, I, P, P]
    rc = L.ko_cls_fit_pairs(_p(codes), n, v, _p(afp), _p(pairs), npairs, nsel, iterations, nthreads, _p(out), _p(q))
    if rc != 0:
        raise ValueError("requested samples exceed total samples")
    return out, q


def gt_allele_counts_generic(A, idx):
    """wrapper.rs:412-428 calculate_allele_frequencies::<A> over genotype
    indices (N x V, VCF order).  Returns V x A."""
    idx = np.ascontiguousarray(idx, np.uint8)
    n, v = idx.shape
    af = np.zeros((v, A))
    L = lib()
    L.ko_gt_allele_counts_generic.argtypes = [I, P, U32, U32, P]
    L.ko_gt_allele_counts_generic(A, _p(idx), n, v, _p(af))
    return af


def cls_m_generic(A, af):
    """cls.rs:6-26 m-matrices over IIS_A: V x G x G x 9."""
    af = f64(af).reshape(-1, A)
    v = af.shape[0]
    G = A * (A + 1) // 2
    m = np.zeros((v, G, G, 9))
    L = lib()
    L.ko_cls_m_generic.argtypes = [I, P, U32, P]
    L.ko_cls_m_generic(A, _p(af), v, _p(m))
    return m


def cls_gram_multi(m2, m3, m4):
    """coefficients.rs:280-320: Gram matrix of every m row (2-, then 3-, then
    4-allele sites) in the 8-lane order, divided by the number of sites."""
    ms = [f64(m) for m in (m2, m3, m4)]
    q = np.zeros((9, 9))
    L = lib()
    L.ko_cls_gram_multi.argtypes = [P, U32, P, U32, P, U32, P]
    L.ko_cls_gram_multi(_p(ms[0]), ms[0].shape[0], _p(ms[1]), ms[1].shape[0], _p(ms[2]), ms[2].shape[0], _p(q))
    return q


def cls_fit_pairs_multi(idx2, idx3, idx4, af=None, pairs=None, nsel=None, iterations=100, nthreads=0):
    """coefficients.rs:166-278 with 2-, 3- and 4-allele sites: idxA are N x V_A
    genotype indices (VCF order, no missing calls); af = (af2, af3, af4) or
    None (counted per bucket).  Returns (outputs, Q)."""
    idx = [np.ascontiguousarray(a, np.uint8) for a in (idx2, idx3, idx4)]
    n = idx[0].shape[0]
    assert all(a.shape[0] == n for a in idx)
    if nsel is None:
        nsel = n
    if pairs is not None:
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        npairs = pairs.shape[0]
    else:
        npairs = nsel * (nsel + 1) // 2
    afs = [None, None, None] if af is None else [f64(a) for a in af]
    out = np.zeros(npairs, OUTPUT_DTYPE)
    q = np.zeros((9, 9))
    L = lib()
    L.ko_cls_fit_pairs_multi.restype = I
    L.ko_cls_fit_pairs_multi.argtypes = [U32, P, U32, P, U32, P, U32, P, P, P, P, SZ, U32, U64, I, P, P]
    rc = L.ko_cls_fit_pairs_multi(n, _p(idx[0]), idx[0].shape[1], _p(idx[1]), idx[1].shape[1], _p(idx[2]),
                                  idx[2].shape[1], _p(afs[0]), _p(afs[1]), _p(afs[2]), _p(pairs), npairs, nsel,
                                  iterations, nthreads, _p(out), _p(q))
    if rc != 0:
        raise ValueError("requested samples exceed total samples")
    return out, q

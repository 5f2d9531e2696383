"""Pins of the per-site reductions: objective (jacquard/objective.rs:8-52),
gradient and Hessian (jacquard/grad_hess.rs:12-449), in the reference's
8-lane/tile order and in the segmented order, against compensated sums and
finite differences."""
import math

import numpy as np
import pytest

EPS = 1e-8


def pair_rows(O, rng, n, V=None):
    """Rows c_s of a random pair of hard calls (one-hot mixture rows)."""
    p = np.clip(rng.beta(0.5, 0.5, n), 0.01, 0.99)
    af = np.stack([p, 1 - p], 1)
    gx, gy = rng.integers(0, 3, n), rng.integers(0, 3, n)
    rows, valid, m = O.pair_rows_codes(gx.astype(np.uint8), gy.astype(np.uint8), af)
    assert m == n
    return rows


def exact_fgh(rows, x):
    L = [math.fsum([float(x[k] * r[k]) for k in range(9)] + [EPS]) for r in rows]
    n = len(rows)
    f = -math.fsum(math.log(v) for v in L) / n
    g = np.array([-math.fsum(r[k] / v for r, v in zip(rows, L)) / n for k in range(9)])
    h = np.array([[math.fsum(r[a] * r[b] / (v * v) for r, v in zip(rows, L)) / n for b in range(9)]
                  for a in range(9)])
    return f, g, h


@pytest.mark.parametrize("n", [1, 5, 8, 9, 64, 255, 256, 257, 1000, 2311])
def test_obj_grad_hess_vs_compensated_sums(O, n):
    rng = np.random.default_rng(n)
    rows = pair_rows(O, rng, n)
    x = rng.dirichlet(np.ones(9))
    f, g, h = exact_fgh(rows, x)
    for order in ("ref", "seg"):
        if order == "ref":
            fo = O.obj_ref(rows, x)
            go, ho = O.gh_ref(rows, x)
        else:
            seg = np.array([0, n // 3, n], np.uint32)
            valid = np.ones(n, np.uint8)
            fo = O.obj_seg(rows, valid, seg, x)
            go, ho = O.gh_seg(rows, valid, seg, x)
        assert abs(fo - f) <= 1e-13 * abs(f)
        np.testing.assert_allclose(go, g, rtol=1e-12, atol=1e-14 * np.abs(g).max())
        np.testing.assert_allclose(ho, h, rtol=1e-12, atol=1e-14 * np.abs(h).max())


def test_gradient_and_hessian_are_derivatives_of_objective(O):
    rng = np.random.default_rng(7)
    rows = pair_rows(O, rng, 3000)
    x = rng.dirichlet(np.ones(9) * 3)
    g, h = O.gh_ref(rows, x)
    for k in range(9):
        e = np.zeros(9)
        e[k] = 1e-6
        fd = (O.obj_ref(rows, x + e) - O.obj_ref(rows, x - e)) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k]))
        gp, _ = O.gh_ref(rows, x + e)
        gm, _ = O.gh_ref(rows, x - e)
        np.testing.assert_allclose((gp - gm) / 2e-6, h[:, k], rtol=1e-5, atol=1e-6 * np.abs(h).max())


def test_hessian_symmetric_psd(O):
    rng = np.random.default_rng(8)
    rows = pair_rows(O, rng, 777)
    _, h = O.gh_ref(rows, rng.dirichlet(np.ones(9)))
    assert np.array_equal(h, h.T)
    assert np.linalg.eigvalsh(h).min() >= -1e-12 * np.abs(h).max()


def test_masked_sites_are_skipped_exactly(O):
    """Sites with a missing call contribute nothing: the gradient and Hessian
    equal the sums over the usable sites alone, and whatever the rows of
    masked sites hold does not change the objective."""
    rng = np.random.default_rng(9)
    rows = pair_rows(O, rng, 600)
    valid = (rng.random(600) > 0.2).astype(np.uint8)
    x = rng.dirichlet(np.ones(9))
    seg = np.array([0, 600], np.uint32)
    garbage = rows.copy()
    garbage[valid == 0] = rng.random((int((valid == 0).sum()), 9))
    assert O.obj_seg(rows, valid, seg, x) == O.obj_seg(garbage, valid, seg, x)
    g1, h1 = O.gh_seg(garbage, valid, seg, x)
    kept = rows[valid == 1]
    g2, h2 = O.gh_seg(kept, np.ones(len(kept), np.uint8), np.array([0, len(kept)], np.uint32), x)
    assert np.array_equal(g1, g2) and np.array_equal(h1, h2)


def test_grouped_log_objective_accuracy(O):
    """The segmented order sums log over 16-site groups as log(prod L); it is
    as accurate as separate logs (checked against a compensated sum)."""
    rng = np.random.default_rng(10)
    for n in (16, 17, 1000, 4099):
        rows = pair_rows(O, rng, n)
        x = rng.dirichlet(np.ones(9))
        f, _, _ = exact_fgh(rows, x)
        fs = O.obj_seg(rows, np.ones(n, np.uint8), np.array([0, n // 2 // 32 * 32, n], np.uint32), x)
        assert abs(fs - f) <= 2e-15 * abs(f) + 1e-16


def seg_of(n):
    return np.array([0, n], np.uint32)


def test_seg_objective_outside_the_positive_normal_range(O):
    """The segmented order's 16-site groups: where a group's product leaves
    the positive normal range, its logs are taken one site at a time, so the
    objective is the per-site sum (compensated reference within 1e-13) and
    not the log of an underflowed or overflowed product."""
    rng = np.random.default_rng(70)
    rows = pair_rows(O, rng, 200)
    valid = np.ones(200, np.uint8)
    for x, eps in ((np.eye(9)[8] * 1e-30, 1e-300), (np.full(9, 1e200), 1e-8), (np.full(9, 1e-200), 1e-300)):
        L = [math.fsum([float(x[k] * r[k]) for k in range(9)] + [eps]) for r in rows]
        exact = -math.fsum(math.log(v) for v in L) / len(rows)
        got = O.obj_seg(rows, valid, seg_of(200), x, eps)
        assert abs(got - exact) <= 1e-13 * abs(exact), (x[0], got, exact)
        ref = O.obj_ref(rows, x, eps)
        assert abs(ref - exact) <= 1e-13 * abs(exact)


def test_seg_objective_special_values(O):
    """L = 0 gives +inf, a negative L NaN (log/scalar.rs:49-57) in both
    orders, also with two negative L in one 16-site group (whose product is
    positive)."""
    rng = np.random.default_rng(71)
    rows = pair_rows(O, rng, 64)
    valid = np.ones(64, np.uint8)
    zero = int(np.nonzero(rows[:, 0] == 0)[0][0])  # a site where c_0 = 0
    x = np.eye(9)[0]
    assert np.isposinf(O.obj_seg(rows, valid, seg_of(64), x, 0.0))
    assert np.isposinf(O.obj_ref(rows, x, 0.0))
    assert rows[zero, 0] == 0
    neg = rows.copy()
    neg[:, :] = np.abs(neg)
    x = np.full(9, 1 / 9)
    neg[3] = -neg[3]   # two negative L in the first group: product positive
    neg[5] = -neg[5]
    assert np.isnan(O.obj_seg(neg, valid, seg_of(64), x, 0.0))
    assert np.isnan(O.obj_ref(neg, x, 0.0))

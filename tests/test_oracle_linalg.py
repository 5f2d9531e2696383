"""Pins of the solver pieces: GMW81 / LDL^T (cholesky.rs), the active-set QP
(sqp.rs:111-268) and the condition estimate (sqp.rs:312-391)."""
import itertools

import numpy as np
import pytest


def rand_spd(rng, n, cond=1e3):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    return (q * ev) @ q.T


def test_unmodified_ldl_reconstructs(O):
    rng = np.random.default_rng(0)
    for n in range(1, 10):
        a = rand_spd(rng, n)
        l, d = O.unmodified(a)
        assert np.allclose(np.triu(l, 1), 0) and np.allclose(np.diag(l), 1)
        np.testing.assert_allclose(l @ np.diag(d) @ l.T, a, rtol=0, atol=1e-10 * np.abs(a).max())


def test_gmw_is_exact_for_well_conditioned_pd(O):
    rng = np.random.default_rng(1)
    for n in range(1, 10):
        a = rand_spd(rng, n, cond=10)
        l, d = O.modify_gmw(a)
        l0, d0 = O.unmodified(a)
        np.testing.assert_allclose(d, d0, rtol=1e-12)
        np.testing.assert_allclose(l, l0, rtol=1e-12, atol=1e-14)


def test_gmw_adds_nonnegative_diagonal(O):
    """GMW81 (Practical Optimization 4.4.2.2): L D L^T = A + E, E diagonal >= 0,
    D > 0 even for indefinite A."""
    rng = np.random.default_rng(2)
    for n in range(2, 10):
        m = rng.standard_normal((n, n))
        a = (m + m.T) / 2
        l, d = O.modify_gmw(a)
        assert (d > 0).all()
        e = l @ np.diag(d) @ l.T - a
        off = e - np.diag(np.diag(e))
        assert np.abs(off).max() <= 1e-10 * max(1.0, np.abs(a).max())
        assert (np.diag(e) >= -1e-10 * max(1.0, np.abs(a).max())).all()


def test_solve_ldl(O):
    rng = np.random.default_rng(3)
    for n in range(1, 10):
        a = rand_spd(rng, n)
        b = rng.standard_normal(n)
        l, d = O.unmodified(a)
        np.testing.assert_allclose(O.solve_ldl(l, d, b), np.linalg.solve(a, b), rtol=1e-9, atol=1e-12)


def qp_bruteforce(q, c):
    """min 1/2 y'Qy + c'y, y >= 0, sum y = 1: enumerate the supports, solve each
    equality-constrained KKT system, keep the feasible KKT point."""
    n = len(c)
    best = None
    for k in range(1, n + 1):
        for s in itertools.combinations(range(n), k):
            s = list(s)
            kkt = np.zeros((k + 1, k + 1))
            kkt[:k, :k] = q[np.ix_(s, s)]
            kkt[:k, k] = 1
            kkt[k, :k] = 1
            rhs = np.concatenate([-c[s], [1.0]])
            try:
                sol = np.linalg.solve(kkt, rhs)
            except np.linalg.LinAlgError:
                continue
            y = np.zeros(n)
            y[s] = sol[:k]
            if (y < -1e-12).any():
                continue
            lam = -sol[k]
            g = q @ y + c
            if (g - lam < -1e-9).any():
                continue
            val = 0.5 * y @ q @ y + c @ y
            if best is None or val < best[0]:
                best = (val, y)
    return best[1]


@pytest.mark.parametrize("n", [2, 3, 5, 9])
def test_qp_matches_support_enumeration(O, n):
    rng = np.random.default_rng(10 + n)
    t = O.tune(100, 200)
    checked = 0
    for _ in range(40):
        q = rand_spd(rng, n, cond=50)
        c = rng.standard_normal(n) * 2
        y0 = np.full(n, 1.0 / n)
        y, it = O.solve_qp(q, c, y0, True, t)
        if it >= 200:
            continue
        np.testing.assert_allclose(y, qp_bruteforce(q, c), atol=1e-10)
        checked += 1
    assert checked >= 30


def test_condition_estimate_tracks_eigenvalues(O):
    """The estimate is a lower bound of the same order of magnitude as the
    eigenvalue condition number of the reduced Hessian."""
    rng = np.random.default_rng(4)
    for n in (3, 9):
        for cond in (1e1, 1e3, 1e6):
            h = rand_spd(rng, n, cond=cond)
            x = np.full(n, 1.0 / n)
            g = np.zeros(n)
            est = O.condition(x, g, h)
            ev = np.linalg.eigvalsh(h)
            k = ev.max() / ev.min()
            assert est <= k * (1 + 1e-8)
            assert np.log10(est) >= np.log10(k) - 2.0


def test_condition_drops_strict_complementarity_and_flags_singular(O):
    h = np.diag([1.0, 2.0, 0.0])
    x = np.array([0.5, 0.5, 0.0])
    g = np.array([1.0, 1.0, 5.0])  # x3 == 0 and g3 >= lambda + tol: projected out
    assert np.isfinite(O.condition(x, g, h))
    g2 = np.array([1.0, 1.0, 1.0])  # x3 == 0 but not strictly complementary: singular
    assert O.condition(x, g2, h) == np.inf

"""Closed forms and invariants that pin the oracle independently of the
reference: eigenvalues of tridiag(-1,2,-1), the Chebyshev recurrence,
symmetry and zero row sums of the operator, a dense direct solve on tiny grids,
and conservation of sum(vol*u) by exact solves."""
import math

import numpy as np
import pytest

import tl_oracle as O
from helpers import interior, random_chunk


def test_tqli_toeplitz():
    n = 17
    d = np.full(n, 2.0)
    e = np.zeros(n)
    e[1:] = -1.0  # e(2:n) is the sub-diagonal on entry (eoshift'ed inside)
    info, ev = O.tqli(d, e)
    assert info == 0
    exact = 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    assert np.allclose(np.sort(ev), np.sort(exact), rtol=0, atol=1e-13)


def test_tqli_random_symmetric_tridiagonal():
    rng = np.random.default_rng(0)
    n = 25
    d = rng.uniform(1, 5, n)
    off = rng.uniform(-1, 1, n - 1)
    e = np.zeros(n)
    e[1:] = off
    info, ev = O.tqli(d, e)
    T = np.diag(d) + np.diag(off, 1) + np.diag(off, -1)
    assert info == 0
    assert np.allclose(np.sort(ev), np.linalg.eigvalsh(T), atol=1e-12)


def test_eigen_from_cg_coefficients():
    """CG on an SPD matrix: the Lanczos tridiagonal built from alpha/beta
    (tea_leaf_cheby.f90:156-160) has eigenvalues inside the spectrum and the
    extreme ones converge to the matrix's."""
    rng = np.random.default_rng(1)
    n = 40
    lam = np.linspace(1.0, 50.0, n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(lam) @ Q.T
    b = rng.standard_normal(n)
    x = np.zeros(n)
    r = b.copy()
    p = r.copy()
    rro = r @ r
    al, be = [], []
    for _ in range(n):
        w = A @ p
        a = rro / (p @ w)
        x += a * p
        r -= a * w
        rrn = r @ r
        al.append(a)
        be.append(rrn / rro)
        p = r + be[-1] * p
        rro = rrn
    info, mn, mx = O.calc_eigenvalues(al[:30], be[:30])
    assert info == 0
    assert 1.0 - 1e-8 <= mn and mx <= 50.0 + 1e-8
    assert abs(mx - 50.0) < 1e-6 and abs(mn - 1.0) < 1e-3


def test_chebyshev_coefficients_closed_form():
    """rho_n from the recurrence equals T_{n}(sigma)/T_{n+1}(sigma)... checked
    through the three-term identity alpha_n = rho_n rho_{n-1},
    beta_n = 2 rho_n / delta (tea_leaf_cheby.f90:187-203)."""
    emin, emax, n = 0.7, 33.0, 50
    a, b, theta = O.calc_ch_coefs(emin, emax, n)
    th, de = (emax + emin) / 2, (emax - emin) / 2
    sigma = th / de
    assert theta == th
    T = [1.0, sigma]
    for _ in range(n + 2):
        T.append(2 * sigma * T[-1] - T[-2])
    for k in range(1, n + 1):
        rho_k = T[k] / T[k + 1]
        rho_km1 = T[k - 1] / T[k]
        assert math.isclose(a[k - 1], rho_k * rho_km1, rel_tol=1e-12)
        assert math.isclose(b[k - 1], 2 * rho_k / de, rel_tol=1e-12)


def _chunk_after_init(nx, ny, seed, coef=1, reflective=False):
    rho, e1, vol, rx, ry = random_chunk(nx, ny, 1, seed)
    o = O.OracleChunk(nx, ny, 1)
    o.field("density")[:] = rho
    o.field("energy1")[:] = e1
    o.set_rxry(rx, ry)
    o.common_init(reflective, coef, 1)
    return o


def _apply(o, x):
    o.field("p")[:] = 0.0
    interior(o.field("p"), 1)[:] = x
    o.cg_calc_w()
    return interior(o.field("w"), 1).copy()


@pytest.mark.parametrize("coef", [1, 2])
def test_operator_symmetric_and_conservative(coef):
    o = _chunk_after_init(13, 11, 5, coef)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((11, 13))
    y = rng.standard_normal((11, 13))
    Ax, Ay = _apply(o, x), _apply(o, y)
    assert math.isclose(np.sum(x * Ay), np.sum(Ax * y), rel_tol=1e-12)
    # zero-flux faces: A * const = const (Di = 1 + sum of off-diagonals)
    c = np.full((11, 13), 3.25)
    assert np.allclose(_apply(o, c), c, rtol=1e-13, atol=0)
    # column sums equal 1 as well (symmetry), so sum(A x) = sum(x)
    assert math.isclose(np.sum(Ax), np.sum(x), rel_tol=1e-11, abs_tol=1e-11)


def test_cg_converges_to_dense_solve():
    nx, ny = 9, 7
    o = _chunk_after_init(nx, ny, 9)
    # dense matrix of the operator, column by column
    A = np.zeros((nx * ny, nx * ny))
    for i in range(nx * ny):
        e = np.zeros(nx * ny)
        e[i] = 1.0
        A[:, i] = _apply(o, e.reshape(ny, nx)).ravel()
    u0 = interior(o.field("u0"), 1).ravel().copy()
    x = np.linalg.solve(A, u0)
    # re-init and run CG through the oracle kernels
    o = _chunk_after_init(nx, ny, 9)
    o.calc_residual()
    rro = o.cg_init(1)
    for _ in range(200):
        pw = o.cg_calc_w()
        a = rro / pw
        rrn = o.cg_calc_ur(a, 1)
        o.cg_calc_p(rrn / rro, 1, 0)
        rro = rrn
        if math.sqrt(abs(rrn)) < 1e-14:
            break
    assert np.allclose(interior(o.field("u"), 1).ravel(), x, rtol=1e-11, atol=1e-11)


def test_energy_column_conserved_by_solves():
    """sum(mass*e1) = sum(vol*u) is invariant under exact solves (A has unit
    column sums), so the ie column stays constant across steps."""
    import fixtures_util as FX
    fx = FX.load("cg_circles_128")
    recs = O.run(O.parse_deck(fx["deck"]))
    ie = [r["summary"][3] for r in recs]
    assert max(ie) - min(ie) < 1e-12 * abs(ie[0]) * 100
    vol = [r["summary"][0] for r in recs]
    assert all(math.isclose(v, 100.0, rel_tol=1e-14) for v in vol)

"""The CPU reference adjoint (adjoint_ref.py) against central finite
differences of its own forward with exact sparse solves: pins the
face-gradient formula and the first-step term independently of the GPU."""
import numpy as np
import pytest

import adjoint_ref as R


@pytest.mark.parametrize("recip", [False, True])
def test_reference_adjoint_matches_fd(recip):
    rng = np.random.default_rng(5 + recip)
    ny, nx, steps, rx, ry = 10, 12, 3, 3.0, 2.0
    theta = np.log(np.where(rng.random((ny, nx)) < 0.3, 0.1, 100.0)) + 0.1 * rng.standard_normal((ny, nx))
    e0 = rng.uniform(0.1, 25.0, (ny, nx))
    target = rng.uniform(0.0, 1.0, (ny, nx))

    def J(th):
        u = R.forward(th, e0, rx, ry, steps, recip)[-1]
        return float(((u - target) ** 2).sum())

    u = R.forward(theta, e0, rx, ry, steps, recip)[-1]
    g = R.gradient(theta, e0, rx, ry, steps, recip, 2.0 * (u - target))
    h = 1e-5
    fd = np.zeros_like(theta)
    for idx in np.ndindex(theta.shape):
        tp, tm = theta.copy(), theta.copy()
        tp[idx] += h
        tm[idx] -= h
        fd[idx] = (J(tp) - J(tm)) / (2 * h)
    err = np.max(np.abs(fd - g)) / np.max(np.abs(g))
    assert err < 1e-6, err


def test_ppcg_helpers_of_the_adjoint():
    """The adjoint's PPCG helpers: Lanczos bounds from CG coefficients agree
    with the oracle's calc_eigenvalues (tea_leaf_cheby.f90:142-176) and reach
    the spectrum's ends; the Chebyshev coefficients equal the oracle's
    calc_ch_coefs bit for bit."""
    import tl_oracle as O
    from tealeaf_ref_amd.adjoint import chebyshev_coefs, lanczos_bounds
    rng = np.random.default_rng(0)
    n = 60
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = q @ np.diag(np.linspace(1.0, 50.0, n)) @ q.T
    x, r = np.zeros(n), rng.standard_normal(n)
    p, rro = r.copy(), r @ r
    al, be = [], []
    for _ in range(40):
        w = a @ p
        alpha = rro / (p @ w)
        x += alpha * p
        r -= alpha * w
        rrn = r @ r
        al.append(alpha)
        be.append(rrn / rro)
        p = r + be[-1] * p
        rro = rrn
    mn, mx = lanczos_bounds(al[:30], be[:30])
    info, omn, omx = O.calc_eigenvalues(al[:30], be[:30])
    assert info == 0
    assert abs(mn - omn) <= 1e-10 * omx and abs(mx - omx) <= 1e-10 * omx
    assert abs(mx - 50.0) < 1e-6 and abs(mn - 1.0) < 1e-3
    ca, cb, th = chebyshev_coefs(0.7, 33.0, 25)
    oa, ob, oth = O.calc_ch_coefs(0.7, 33.0, 25)
    assert th == oth and np.array_equal(ca, np.asarray(oa)) and np.array_equal(cb, np.asarray(ob))

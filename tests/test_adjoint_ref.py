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

"""Independent FP64 reference of the N-step forward and its adjoint, written
directly from the mathematics with scipy sparse direct solves (test
infrastructure only).

Face coefficients follow kernels/tea_leaf_common_kernel.f90:75-101:
K = (w_a + w_b) / (2 w_a w_b), w = rho (conductivity) or 1/rho (inverse
density), zero on external faces; A = I + rx Lx + ry Ly
(kernels/tea_leaf_cg_kernel.f90:145-147).  A step solves A u = rho * e and
sets e = u / rho (tea_leaf_common.f90:67-72, tea_leaf_finalise).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def faces(rho, recip):
    w = 1.0 / rho if recip else rho
    kx = np.zeros((rho.shape[0], rho.shape[1] + 1))
    ky = np.zeros((rho.shape[0] + 1, rho.shape[1]))
    kx[:, 1:-1] = (w[:, :-1] + w[:, 1:]) / (2.0 * w[:, :-1] * w[:, 1:])
    ky[1:-1, :] = (w[:-1, :] + w[1:, :]) / (2.0 * w[:-1, :] * w[1:, :])
    return kx, ky


def operator(kx, ky, rx, ry):
    ny, nx = kx.shape[0], kx.shape[1] - 1
    kl, kr = kx[:, :-1], kx[:, 1:]
    kd, ku = ky[:-1, :], ky[1:, :]
    main = 1.0 + rx * (kl + kr) + ry * (kd + ku)
    n = nx * ny
    diags, offs = [main.ravel()], [0]
    if n > 1:
        diags += [-rx * kr.ravel()[:-1], -rx * kl.ravel()[1:]]
        offs += [1, -1]
    if ny > 1:
        diags += [-ry * ku.ravel()[:-nx], -ry * kd.ravel()[nx:]]
        offs += [nx, -nx]
    return sp.diags(diags, offs, shape=(n, n), format="csc")


def forward(theta, e0, rx, ry, steps, recip):
    """u after each step (list), with exact solves."""
    rho = np.exp(theta)
    A = spla.splu(operator(*faces(rho, recip), rx, ry))
    e = e0.copy()
    us = []
    for _ in range(steps):
        u = A.solve((rho * e).ravel()).reshape(rho.shape)
        us.append(u)
        e = u / rho
    return us


def gradient(theta, e0, rx, ry, steps, recip, grad_u):
    """dJ/dtheta for dJ/du_N = grad_u, by the discrete adjoint."""
    rho = np.exp(theta)
    A = spla.splu(operator(*faces(rho, recip), rx, ry))
    us = forward(theta, e0, rx, ry, steps, recip)
    dk = 0.5 * rho if recip else -0.5 / rho
    g = grad_u.copy()
    out = np.zeros_like(theta)
    for n in range(steps, 0, -1):
        lam = A.solve(g.ravel()).reshape(rho.shape)
        u = us[n - 1]
        s = np.zeros_like(theta)
        dlx = (lam[:, 1:] - lam[:, :-1]) * (u[:, 1:] - u[:, :-1]) * rx
        dly = (lam[1:, :] - lam[:-1, :]) * (u[1:, :] - u[:-1, :]) * ry
        s[:, 1:] += dlx
        s[:, :-1] += dlx
        s[1:, :] += dly
        s[:-1, :] += dly
        out += -s * dk
        if n == 1:
            out += lam * rho * e0
        g = lam
    return out

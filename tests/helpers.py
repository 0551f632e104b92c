"""Shared helpers of the parity tests: seeded random chunks in the reference
layout and comparison utilities."""
import numpy as np


def random_chunk(nx, ny, h, seed, rho_range=(0.1, 100.0)):
    """Density (log-uniform, with a few sharp blocks), energy1, volume and
    rx/ry for an nx*ny chunk with halo h; arrays are (ny+2h, nx+2h)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(rho_range[0]), np.log(rho_range[1])
    rho = np.exp(rng.uniform(lo, hi, size=(ny + 2 * h, nx + 2 * h)))
    # add piecewise-constant structure like the decks' states
    for _ in range(3):
        x0, y0 = rng.integers(0, nx), rng.integers(0, ny)
        rho[y0:y0 + ny // 3, x0:x0 + nx // 3] = rng.choice([0.1, 100.0])
    e1 = rng.uniform(1e-4, 25.0, size=(ny + 2 * h, nx + 2 * h))
    dx, dy = 10.0 / nx, 10.0 / ny
    vol = np.full((ny + 4, nx + 4), dx * dy)
    dt = 0.004
    return rho, e1, vol, dt / (dx * dx), dt / (dy * dy)


def interior(a, h):
    return a[h:a.shape[0] - h, h:a.shape[1] - h]


def relerr(a, b):
    """max |a-b| / max |b| (field-scale relative error)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = np.max(np.abs(b))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


def assert_close(a, b, tol, what=""):
    e = relerr(a, b)
    assert e <= tol, f"{what}: relative error {e:.3e} > {tol:.1e}"


def assert_scalar(a, b, tol, what=""):
    d = abs(a - b) / max(abs(b), 1e-300)
    assert d <= tol, f"{what}: {a!r} vs {b!r} (rel {d:.3e})"

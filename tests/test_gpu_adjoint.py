"""Adjoint of N implicit steps (tl_adjoint.hip, tealeaf_ref_amd.adjoint)
against the CPU reference adjoint (adjoint_ref.py, itself pinned by finite
differences), against central finite differences of the GPU forward, and
for checkpoint invariance, the torch binding and the mixed-precision form."""
import numpy as np
import pytest

import adjoint_ref as R
from tealeaf_ref_amd import _lib as L
from tealeaf_ref_amd.adjoint import TeaLeafSteps

pytestmark = pytest.mark.gpu


def case(seed, ny=40, nx=48):
    rng = np.random.default_rng(seed)
    base = np.where(rng.random((ny // 8, nx // 8)) < 0.3, 0.1, 100.0)
    rho = np.kron(base, np.ones((8, 8)))
    theta = np.log(rho) + 0.1 * rng.standard_normal((ny, nx))
    e0 = rng.uniform(0.1, 25.0, (ny, nx))
    target = rng.uniform(0.0, 2.0, (ny, nx))
    win = (slice(8, 24), slice(10, 34))
    return theta, e0, target, win


def seed_grad(u, target, win):
    g = np.zeros_like(u)
    g[win] = 2.0 * (u[win] - target[win])
    return g


def objective(u, target, win):
    return float(((u[win] - target[win]) ** 2).sum())


@pytest.fixture
def problems():
    made = []

    def make(*a, **k):
        p = TeaLeafSteps(*a, **k)
        made.append(p)
        return p

    yield make
    for p in made:
        p.close()


@pytest.mark.parametrize("coef", [L.CONDUCTIVITY, L.RECIP_CONDUCTIVITY])
def test_forward_and_gradient_match_reference(problems, coef):
    theta, e0, target, win = case(1 + coef)
    steps, rx = 4, 0.5 if coef == L.RECIP_CONDUCTIVITY else 20.0
    p = problems(e0, rx, rx, steps, coefficient=coef, eps=1e-15, checkpoint_every=3)
    print(f"adjoint vs reference coef={coef}", flush=True)
    u = p.forward(theta)
    recip = coef == L.RECIP_CONDUCTIVITY
    u_ref = R.forward(theta, e0, rx, rx, steps, recip)[-1]
    assert np.max(np.abs(u - u_ref)) <= 1e-11 * np.max(np.abs(u_ref))
    g = p.backward(seed_grad(u, target, win))
    g_ref = R.gradient(theta, e0, rx, rx, steps, recip, seed_grad(u_ref, target, win))
    err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    print(f"  max rel gradient error vs reference {err:.2e}", flush=True)
    assert err <= 1e-9


def test_gradient_matches_finite_differences(problems):
    """Central differences with h = 1e-5 in log-density on cells inside the
    window of J.  The target is the forward of a log-normally perturbed
    density (sigma 0.1), as in the benchmark's adjoint case.  The difference
    quotient carries the solver's own error (|du| / (h |du/dtheta|), |du| up
    to eps * cond(A) |u|), so cells that barely couple to the window, where
    |du/dtheta| is small, are not a sharp test."""
    theta, e0, _, win = case(7)
    steps, rx, h = 3, 20.0, 1e-5
    p = problems(e0, rx, rx, steps, coefficient=L.CONDUCTIVITY, eps=1e-15, checkpoint_every=2)
    rng = np.random.default_rng(3)
    target = p.forward(theta + 0.1 * rng.standard_normal(theta.shape))
    u = p.forward(theta)
    g = p.backward(seed_grad(u, target, win))
    cells = [(int(rng.integers(9, 23)), int(rng.integers(11, 33))) for _ in range(8)]
    worst = 0.0
    for c in cells:
        tp, tm = theta.copy(), theta.copy()
        tp[c] += h
        tm[c] -= h
        fd = (objective(p.forward(tp), target, win) - objective(p.forward(tm), target, win)) / (2 * h)
        rel = abs(fd - g[c]) / max(abs(g[c]), 1e-300)
        worst = max(worst, rel)
        assert rel <= 1e-5, (c, fd, g[c])
    print(f"  FD check on {len(cells)} cells: max rel err {worst:.2e}", flush=True)


def test_checkpointing_is_bit_identical(problems):
    theta, e0, target, win = case(11)
    grads = []
    for k in (1, 2, 5):
        p = problems(e0, 10.0, 10.0, 5, checkpoint_every=k)
        u = p.forward(theta)
        grads.append(p.backward(seed_grad(u, target, win)))
        p.close()
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


@pytest.mark.parametrize("dev", ["cpu", "cuda"])
def test_torch_function(problems, dev):
    """The autograd binding against the numpy calls.  A CPU tensor takes the
    numpy path (identical bits); a GPU tensor stays on the GPU (exp on the
    GPU, device-to-device copies), where exp may round a density one unit
    differently from numpy's, hence a tolerance."""
    import torch
    theta, e0, target, win = case(13)
    p = problems(e0, 10.0, 10.0, 3, checkpoint_every=2)
    th = torch.tensor(theta, device=dev, requires_grad=True)
    tg = torch.tensor(target, device=dev)
    u = p.apply(th)
    assert u.device == th.device and u.dtype == th.dtype
    J = ((u[win] - tg[win]) ** 2).sum()
    J.backward()
    assert th.grad.device == th.device
    u_np = p.forward(theta)
    g_np = p.backward(seed_grad(u_np, target, win))
    g = th.grad.cpu().numpy()
    if dev == "cpu":
        assert np.array_equal(g, g_np)
        assert np.array_equal(u.detach().numpy(), u_np)
    else:
        assert np.allclose(u.detach().cpu().numpy(), u_np, rtol=1e-12, atol=0.0)
        assert np.linalg.norm(g - g_np) <= 1e-9 * np.linalg.norm(g_np)
    assert abs(J.item() - objective(u_np, target, win)) <= 1e-11 * abs(J.item())


def test_device_copies():
    """tl_upload_device / tl_download_device and the snapshot forms copy the
    interior of a field bit for bit to and from torch tensors on the GPU
    (row stride ld), and refuse a host pointer or a short row stride."""
    import ctypes as C
    import torch
    nx, ny = 37, 29
    d = L.Device(nx, ny, 2)
    try:
        rng = np.random.default_rng(3)
        a = rng.standard_normal((ny, nx))
        ta = torch.tensor(a, device="cuda")
        L.call("tl_upload_device", L.FIELD["u0"], ta.data_ptr(), nx)
        assert np.array_equal(d.download("u0")[2:-2, 2:-2], a)
        # a strided source: every other column of a (ny, 2 nx) tensor
        wide = torch.zeros((ny, 2 * nx), dtype=torch.float64, device="cuda")
        wide[:, :nx] = ta * 2.0
        L.call("tl_upload_device", L.FIELD["r"], wide.data_ptr(), 2 * nx)
        assert np.array_equal(d.download("r")[2:-2, 2:-2], 2.0 * a)
        out = torch.full((ny, nx), -1.0, dtype=torch.float64, device="cuda")
        L.call("tl_download_device", L.FIELD["u0"], out.data_ptr(), nx)
        assert np.array_equal(out.cpu().numpy(), a)
        L.call("tl_snapshot_reserve", 1)
        L.call("tl_snapshot_fill", 0, 0.0)
        L.call("tl_snapshot_upload_device", 0, ta.data_ptr(), nx)
        back = torch.zeros((ny, nx), dtype=torch.float64, device="cuda")
        L.call("tl_snapshot_download_device", 0, back.data_ptr(), nx)
        assert torch.equal(back, ta)
        host = np.zeros((ny, nx))
        with pytest.raises(L.TeaLeafError, match="not device memory"):
            L.call("tl_upload_device", L.FIELD["u0"], host.ctypes.data, nx)
        with pytest.raises(L.TeaLeafError, match="row stride"):
            L.call("tl_download_device", L.FIELD["u0"], out.data_ptr(), nx - 1)
        with pytest.raises(L.TeaLeafError, match="unknown field"):
            L.call("tl_upload_device", 999, ta.data_ptr(), nx)
    finally:
        d.close()


def test_mixed_precision_adjoint(problems):
    theta, e0, target, win = case(17)
    grads = {}
    for prec in ("fp64", "mixed"):
        p = problems(e0, 20.0, 20.0, 4, checkpoint_every=2, precision=prec,
                     eps=1e-15 if prec == "fp64" else 1e-11)
        u = p.forward(theta)
        grads[prec] = p.backward(seed_grad(u, target, win))
        p.close()
    rel = np.linalg.norm(grads["mixed"] - grads["fp64"]) / np.linalg.norm(grads["fp64"])
    print(f"  mixed vs FP64 gradient rel L2 {rel:.2e}", flush=True)
    assert rel <= 1e-7


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_ppcg_adjoint(problems, prec):
    """solver="ppcg": the forward and the gradient agree with FP64 CG's (the
    same linear systems solved to a tight tolerance), PPCG really ran (inner
    sweeps counted) and checkpointing stays bit-identical."""
    theta, e0, target, win = case(19)
    ref = problems(e0, 20.0, 20.0, 4, checkpoint_every=2, eps=1e-15)
    u_ref = ref.forward(theta)
    g_ref = ref.backward(seed_grad(u_ref, target, win))
    ref.close()
    grads = []
    for k in (1, 4):
        p = problems(e0, 20.0, 20.0, 4, checkpoint_every=k, precision=prec, solver="ppcg",
                     inner_steps=6, presteps=10, eps=1e-15 if prec == "fp64" else 1e-12)
        print(f"ppcg adjoint {prec} checkpoint {k}", flush=True)
        u = p.forward(theta)
        assert p.sweeps > 0
        du = np.max(np.abs(u - u_ref)) / np.max(np.abs(u_ref))
        grads.append(p.backward(seed_grad(u, target, win)))
        assert p.sweeps > 0
        p.close()
        rel = np.linalg.norm(grads[-1] - g_ref) / np.linalg.norm(g_ref)
        print(f"  u max rel {du:.2e}, gradient rel L2 vs FP64 CG {rel:.2e}", flush=True)
        assert du <= (1e-12 if prec == "fp64" else 1e-9)
        assert rel <= (1e-10 if prec == "fp64" else 1e-7)
    assert np.array_equal(grads[0], grads[1])


def test_reflective_boundary_matches_reference(problems):
    """Reflective faces (the halo mirrors the cell beside it, so no flux
    crosses an external face, as with zeroed faces): forward and gradient
    against the same reference."""
    theta, e0, target, win = case(23)
    steps, rx = 3, 20.0
    p = problems(e0, rx, rx, steps, coefficient=L.CONDUCTIVITY, eps=1e-15, checkpoint_every=2,
                 reflective=True)
    print("adjoint reflective", flush=True)
    u = p.forward(theta)
    u_ref = R.forward(theta, e0, rx, rx, steps, False)[-1]
    assert np.max(np.abs(u - u_ref)) <= 1e-11 * np.max(np.abs(u_ref))
    g = p.backward(seed_grad(u, target, win))
    g_ref = R.gradient(theta, e0, rx, rx, steps, False, seed_grad(u_ref, target, win))
    err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    print(f"  max rel gradient error vs reference {err:.2e}", flush=True)
    assert err <= 1e-9


def test_mixed_ppcg_adjoint_at_scale():
    """The bench's 2048^2 adjoint case (20 steps, checkpoint every 5) with
    PPCG solves: the mixed gradient within 1e-8 of the FP64 one.  Later steps
    start from residuals near the true residual's rounding floor, where a
    mixed solve must tell a stall of the eigenvalue-estimate CG (which ends
    when its estimate settles) from the floor itself; ending at the first
    stall put this gradient 1.5e-3 off."""
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import bench
    print("adjoint 2048^2 PPCG, fp64 and mixed", flush=True)
    out = bench.adjoint_bench(2048, 20, 5, 0, 1e-6, "both", solver="ppcg")
    rel = out["mixed_vs_fp64_gradient_rel_l2"]
    dj = abs(out["mixed"]["J"] - out["fp64"]["J"]) / abs(out["fp64"]["J"])
    print(f"  mixed vs FP64 gradient rel L2 {rel:.2e}, J rel {dj:.2e}", flush=True)
    assert rel <= 1e-8
    assert dj <= 1e-9

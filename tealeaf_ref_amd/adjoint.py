"""Gradient of N TeaLeaf timesteps with respect to the per-cell log-density.

`TeaLeafSteps` runs the implicit heat-conduction step of the reference
(tea_leaf.f90 -> tea_leaf_init_common, the CG solve of tea_solve.f90,
tea_leaf_finalise) N times on one GPU through libtealeaf_hip.so, starting
from a given energy field and a per-cell density, and returns u after the
last solve.  `TeaLeafSteps.apply(theta)` is a torch.autograd.Function: its
backward is the hand-written adjoint (tl_adjoint.hip): one solve with the
same symmetric operator per step, plus the face-gradient kernel chained to
log-density.  The forward keeps the energy every `checkpoint_every` steps in
device snapshot slots; the backward recomputes each segment from its
checkpoint and walks it in reverse.  Solves are FP64 CG or PPCG, or the
FP32-inner, FP64-refined CG or PPCG (precision="mixed"); solver="ppcg" takes
its eigenvalue bounds from `presteps` CG iterations of every solve, so a
recomputed step repeats the forward's arithmetic exactly.

The library holds one chunk per process: one TeaLeafSteps at a time.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L

FIELDS = 11  # FIELD_* slots of tl_update_halo


def _halo_flags(*names):
    order = ("density", "energy0", "energy1", "u", "p", "sd", "r", "z", "kx", "ky", "di")
    f = [0] * FIELDS
    for n in names:
        f[order.index(n)] = 1
    return f


def _pad(a):
    """Interior (ny, nx) -> (ny+2, nx+2) with the edge values copied into the halo."""
    return np.pad(np.asarray(a, dtype=np.float64), 1, mode="edge")


class TeaLeafSteps:
    """N implicit steps on an nx*ny chunk with a per-cell density.

    energy0: initial energy1 field (ny, nx); rx, ry: dt/dx^2, dt/dy^2;
    coefficient: _lib.CONDUCTIVITY or _lib.RECIP_CONDUCTIVITY; eps: relative
    residual of every solve (recurrence residual for FP64 CG and PPCG as in
    tea_solve.f90:410, true residual for the mixed solver); solver: "cg" or
    "ppcg" (inner_steps Chebyshev sweeps per outer iteration, bounds from
    presteps CG iterations).
    """

    def __init__(self, energy0, rx, ry, steps, coefficient=L.CONDUCTIVITY, eps=1e-15,
                 max_iters=20000, checkpoint_every=5, precision="fp64", reflective=False,
                 device=0, mixed_inner_tol=1e-5, solver="cg", inner_steps=10, presteps=30):
        e0 = np.asarray(energy0, dtype=np.float64)
        self.ny, self.nx = e0.shape
        self.rx, self.ry = float(rx), float(ry)
        self.steps = int(steps)
        self.k = max(1, int(checkpoint_every))
        self.eps, self.max_iters = float(eps), int(max_iters)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        if solver not in ("cg", "ppcg"):
            raise ValueError("solver must be 'cg' or 'ppcg'")
        self.solver = solver
        self.inner_steps, self.presteps = int(inner_steps), int(presteps)
        if solver == "ppcg" and (self.inner_steps < 1 or self.presteps < 2):
            raise ValueError("ppcg needs inner_steps >= 1 and presteps >= 2")
        self.sweeps = 0                # PPCG inner sweeps of the last forward/backward
        # solves that did not meet eps: at the iteration limit (max_iters; the
        # reference carries on silently, tea_solve.f90:152) or, mixed
        # precision, at the true residual's rounding floor
        self.unconverged = self.maxiter_stops = self.floor_stops = 0
        self.mixed_inner_tol = mixed_inner_tol
        self.reflective = bool(reflective)
        self.iterations = 0            # solver iterations of the last forward/backward
        self.progress = None           # optional callable(str), told of every step
        self.device = int(device)
        self.dev = L.Device(self.nx, self.ny, 1, device)
        self.dev.set_options(coefficient, L.PREC_NONE, self.reflective)
        self.dev.upload("energy1", _pad(e0))
        nck = (self.steps + self.k - 1) // self.k
        # slots: e_0, adjoint right-hand side, gradient, zeros, checkpoints, segment
        self.E0, self.RHS, self.GRAD, self.ZERO = 0, 1, 2, 3
        self.CK = 4
        self.SEG = self.CK + nck
        L.call("tl_snapshot_reserve", self.SEG + self.k)
        L.call("tl_snapshot_save", self.E0, L.FIELD["energy1"])
        L.call("tl_snapshot_fill", self.ZERO, 0.0)
        self._forward_done = False

    def close(self):
        self.dev.close()

    # -- one step ---------------------------------------------------------------
    def _solve(self):
        """Solve A u = u0 from the current u (tea_solve.f90:37-414, CG branch)."""
        d = self.dev
        if self.reflective:
            d.update_halo(_halo_flags("u"), 1)
        if self.precision == "mixed":
            ppcg = self.solver == "ppcg"
            st = d.mixed_solve(solver=L.MIXED_PPCG if ppcg else L.MIXED_CG, rel_tol=self.eps,
                               inner_tol=self.mixed_inner_tol, max_inner=self.max_iters,
                               inner_steps=self.inner_steps, presteps=self.presteps)
            self.iterations += st.inner_iters
            self.sweeps += st.sweeps
            self._count(st.reason)
            return
        L.call("tl_set_ppcg_active", 0)
        L.call("tl_calc_residual")
        rr = d.scalar("tl_calc_2norm", 1)
        if rr == 0.0:
            return
        rro = d.scalar("tl_cg_init")
        eps_abs = self.eps * math.sqrt(abs(rr))
        if self.solver == "cg":
            st = L.LoopState(n_first=1, n_last=self.max_iters, eps_abs=eps_abs, check_switch=0,
                             epslim=0.0, presteps=0, rro=rro, old_error=rr)
            L.call("tl_cg_loop", C.byref(st), None, None)
            self.iterations += st.n_end
            self._count(st.reason)
            return
        self._solve_ppcg(rro, rr, eps_abs)

    def _count(self, reason):
        if reason == L.LOOP_CONVERGED:
            return
        self.unconverged += 1
        if reason == L.LOOP_FLOOR:
            self.floor_stops += 1
        else:
            self.maxiter_stops += 1

    def _solve_ppcg(self, rro, rr, eps_abs):
        """FP64 PPCG as tea_solve.f90:160-330 runs it: CG until the switch at
        iteration `presteps`, eigenvalue bounds from the CG coefficients
        (tea_leaf_cheby.f90:142-176, x0.95 / x1.05), Chebyshev coefficients for
        the inner sweeps (tea_leaf_cheby.f90:178-205), then flexible CG with the
        polynomial preconditioner to the exit test."""
        d = self.dev
        n_max = self.max_iters
        al = np.zeros(n_max + 2)
        be = np.zeros(n_max + 2)
        pd = C.POINTER(C.c_double)
        st = L.LoopState(n_first=1, n_last=n_max, eps_abs=eps_abs, check_switch=1,
                         epslim=math.inf, presteps=self.presteps, rro=rro, old_error=rr)
        L.call("tl_cg_loop", C.byref(st), al.ctypes.data_as(pd), be.ctypes.data_as(pd))
        if st.reason != L.LOOP_SWITCH:
            self.iterations += st.n_end
            self._count(st.reason)
            return
        n = st.n_end
        self.iterations += n - 1
        emin, emax = lanczos_bounds(al[1:n], be[1:n])
        ca, cb, theta = chebyshev_coefs(emin * 0.95, emax * 1.05, self.inner_steps)
        L.call("tl_set_ppcg_active", 1)
        d.scalar("tl_ppcg_init", 2)
        L.call("tl_ppcg_init_sd", theta)
        if self.reflective:
            d.update_halo(_halo_flags("sd", "r"), 1)
        L.call("tl_ppcg_inner", ca.ctypes.data_as(pd), cb.ctypes.data_as(pd), 1, self.inner_steps)
        L.call("tl_ppcg_update_z")
        rro = d.scalar("tl_ppcg_init", 3)
        if self.reflective:
            d.update_halo(_halo_flags("p"), 1)
        sweeps = C.c_int(self.inner_steps)
        st2 = L.LoopState(n_first=n, n_last=n_max, eps_abs=eps_abs, check_switch=0, epslim=0.0,
                          presteps=0, rro=rro, old_error=st.old_error)
        L.call("tl_ppcg_loop", C.byref(st2), ca.ctypes.data_as(pd), cb.ctypes.data_as(pd),
               self.inner_steps, theta, C.byref(sweeps))
        L.call("tl_set_ppcg_active", 0)
        self._count(st2.reason)
        self.iterations += st2.n_end - n + 1
        self.sweeps += sweeps.value

    def _step(self):
        d = self.dev
        d.update_halo(_halo_flags("energy1", "density"), 1)
        d.common_init(self.rx, self.ry)
        self._solve()
        L.call("tl_finalise")

    # -- forward / backward --------------------------------------------------------
    def forward(self, theta):
        """u after `steps` steps for log-density theta (ny, nx); numpy in, numpy out."""
        th = np.asarray(theta, dtype=np.float64)
        if th.shape != (self.ny, self.nx):
            raise ValueError(f"theta must have shape {(self.ny, self.nx)}")
        self.dev.upload("density", _pad(np.exp(th)))
        self._run_forward()
        return self.dev.download("u")[1:-1, 1:-1].copy()

    def forward_device(self, theta):
        """forward() on a torch tensor theta (ny, nx) on this chunk's GPU: exp
        on the GPU, density and u copied device to device; returns u as a new
        float64 tensor on the same GPU."""
        import torch
        th = self._device_tensor(theta, "theta")
        rho = torch.exp(th).contiguous()
        torch.cuda.current_stream(th.device).synchronize()
        L.call("tl_upload_device", L.FIELD["density"], rho.data_ptr(), self.nx)
        self._run_forward()
        u = torch.empty((self.ny, self.nx), dtype=torch.float64, device=th.device)
        L.call("tl_download_device", L.FIELD["u"], u.data_ptr(), self.nx)
        return u

    def _device_tensor(self, t, what):
        import torch
        if t.shape != (self.ny, self.nx):
            raise ValueError(f"{what} must have shape {(self.ny, self.nx)}")
        if not t.is_cuda or t.device.index != self.device:
            raise ValueError(f"{what} must be on cuda:{self.device}")
        return t.detach().to(torch.float64).contiguous()

    def _run_forward(self):
        L.call("tl_snapshot_load", self.E0, L.FIELD["energy1"])
        self.iterations = 0
        self.sweeps = 0
        self.unconverged = self.maxiter_stops = self.floor_stops = 0
        for n in range(1, self.steps + 1):
            if (n - 1) % self.k == 0:
                L.call("tl_snapshot_save", self.CK + (n - 1) // self.k, L.FIELD["energy1"])
            self._step()
            if self.progress:
                self.progress(f"forward step {n}/{self.steps}, {self.iterations} iterations")
        self._forward_done = True

    def backward(self, grad_u):
        """dJ/dtheta from dJ/du_N (ny, nx) for the theta of the last forward."""
        if not self._forward_done:
            raise RuntimeError("backward needs a forward of the same TeaLeafSteps first")
        g = np.asarray(grad_u, dtype=np.float64)
        if g.shape != (self.ny, self.nx):
            raise ValueError(f"grad_u must have shape {(self.ny, self.nx)}")
        host = np.zeros((self.ny + 2, self.nx + 2))
        host[1:-1, 1:-1] = g
        # the slot's halo stays zero: g_N lives on the interior
        L.call("tl_snapshot_fill", self.RHS, 0.0)
        L.call("tl_snapshot_upload", self.RHS, host.ctypes.data_as(C.POINTER(C.c_double)), 1)
        self._run_backward()
        out = np.zeros((self.ny + 2, self.nx + 2))
        L.call("tl_snapshot_download", self.GRAD, out.ctypes.data_as(C.POINTER(C.c_double)), 1)
        return out[1:-1, 1:-1].copy()

    def backward_device(self, grad_u):
        """backward() on a torch tensor dJ/du_N on this chunk's GPU, copied
        device to device; returns dJ/dtheta as a new float64 tensor there."""
        import torch
        if not self._forward_done:
            raise RuntimeError("backward needs a forward of the same TeaLeafSteps first")
        g = self._device_tensor(grad_u, "grad_u")
        torch.cuda.current_stream(g.device).synchronize()
        L.call("tl_snapshot_fill", self.RHS, 0.0)
        L.call("tl_snapshot_upload_device", self.RHS, g.data_ptr(), self.nx)
        self._run_backward()
        out = torch.empty((self.ny, self.nx), dtype=torch.float64, device=g.device)
        L.call("tl_snapshot_download_device", self.GRAD, out.data_ptr(), self.nx)
        return out

    def _run_backward(self):
        L.call("tl_snapshot_fill", self.GRAD, 0.0)
        self.iterations = 0
        self.sweeps = 0
        self.unconverged = self.maxiter_stops = self.floor_stops = 0
        nseg = (self.steps + self.k - 1) // self.k
        for seg in range(nseg - 1, -1, -1):
            a, b = seg * self.k, min(seg * self.k + self.k, self.steps)
            # recompute u_{a+1..b} from the checkpoint (bit-identical to the forward)
            L.call("tl_snapshot_load", self.CK + seg, L.FIELD["energy1"])
            for n in range(a + 1, b + 1):
                self._step()
                L.call("tl_snapshot_save", self.SEG + (n - a - 1), L.FIELD["u"])
            for n in range(b, a, -1):
                # lambda_n = A^-1 g_n, from zero
                L.call("tl_snapshot_load", self.RHS, L.FIELD["u0"])
                L.call("tl_snapshot_load", self.ZERO, L.FIELD["u"])
                self._solve()
                L.call("tl_adjoint_face_grad", L.FIELD["u"], self.SEG + (n - a - 1),
                       self.E0 if n == 1 else -1, self.GRAD)
                L.call("tl_snapshot_save", self.RHS, L.FIELD["u"])   # g_{n-1} = lambda_n
                if self.progress:
                    self.progress(f"adjoint step {n}, {self.iterations} iterations")

    # -- torch -------------------------------------------------------------------------
    def apply(self, theta):
        """u_N as a differentiable function of a torch tensor theta (ny, nx)."""
        return _function()(theta, self)


def lanczos_bounds(alphas, betas):
    """Smallest and largest eigenvalue of the Lanczos tridiagonal of m CG
    iterations with coefficients alpha_1..m, beta_1..m: diagonal
    1/alpha_n + beta_{n-1}/alpha_{n-1}, off-diagonal sqrt(beta_n)/alpha_n
    (tea_leaf_cheby.f90:156-160)."""
    a = np.asarray(alphas, dtype=np.float64)
    b = np.asarray(betas, dtype=np.float64)
    diag = 1.0 / a
    diag[1:] += b[:-1] / a[:-1]
    off = np.sqrt(b[:-1]) / a[:-1]
    t = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    ev = np.linalg.eigvalsh(t)
    if not (ev[0] > 0.0):
        raise RuntimeError("eigenvalue estimate is not positive")
    return float(ev[0]), float(ev[-1])


def chebyshev_coefs(eigmin, eigmax, n):
    """alpha_i, beta_i (i = 1..n) and theta of the Chebyshev recurrence
    (tea_leaf_cheby.f90:178-205); arrays start at i = 1."""
    theta = (eigmax + eigmin) / 2.0
    delta = (eigmax - eigmin) / 2.0
    sigma = theta / delta
    rho_old = 1.0 / sigma
    ca, cb = np.zeros(n), np.zeros(n)
    for i in range(n):
        rho_new = 1.0 / (2.0 * sigma - rho_old)
        ca[i] = rho_new * rho_old
        cb[i] = 2.0 * rho_new / delta
        rho_old = rho_new
    return ca, cb, theta


_FN = None


def _function():
    global _FN
    if _FN is None:
        import torch

        class TeaLeafFunction(torch.autograd.Function):
            """theta (log-density) -> u after N steps; backward by the adjoint."""

            @staticmethod
            def forward(ctx, theta, problem):
                ctx.problem = problem
                if theta.is_cuda:
                    # stays on the GPU: device-to-device copies only
                    return problem.forward_device(theta).to(theta.dtype)
                u = problem.forward(theta.detach().to(torch.float64).numpy())
                return torch.from_numpy(u).to(dtype=theta.dtype)

            @staticmethod
            def backward(ctx, grad_u):
                if grad_u.is_cuda:
                    return ctx.problem.backward_device(grad_u).to(grad_u.dtype), None
                g = ctx.problem.backward(grad_u.detach().to(torch.float64).numpy())
                return torch.from_numpy(g).to(dtype=grad_u.dtype), None

        _FN = TeaLeafFunction.apply
    return _FN

"""tl_cg_loop_noalloc (the fused CG loop with its scalars in a file-scope
device array, for callers that cannot allocate) against tl_cg_loop on the
same chunk: the same kernels on the same data, so the iteration count, the
exit reason and error, the recorded alphas and betas and the field u must be
identical to the last bit.  Two calls in a row, the second on a smaller mesh
(the array keeps the first call's contents), a converging and a max_iters
case, and the size limit."""
import ctypes as C
import math

import numpy as np
import pytest

from helpers import random_chunk
from tealeaf_ref_amd import _lib as L

pytestmark = pytest.mark.gpu

MAX_ITERS = 16384  # the entry point's history size


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _solve(d, rx, ry, fn, n_last, eps):
    """One CG solve of the current step from tl_common_init, through `fn`."""
    d.common_init(rx, ry)
    L.call("tl_calc_residual")
    r0 = d.scalar("tl_calc_2norm", 1)
    rro = d.scalar("tl_cg_init")
    st = L.LoopState(n_first=1, n_last=n_last, eps_abs=eps * math.sqrt(r0), check_switch=0,
                     epslim=1.0, presteps=25, rro=rro, old_error=r0)
    al = np.zeros(n_last + 2)
    be = np.zeros(n_last + 2)
    L.check(getattr(L.lib(), fn)(C.byref(st), _dp(al), _dp(be)), fn)
    return st, al, be, d.download("u")


@pytest.mark.parametrize("eps,n_last", [(1e-12, 3000), (1e-300, 150)])
def test_noalloc_matches_cg_loop(eps, n_last):
    # the second mesh is smaller than the first
    for (nx, ny, seed) in ((600, 520, 3), (97, 61, 4)):
        print(f"[no_alloc] {nx}x{ny} eps {eps} n_last {n_last}", flush=True)
        rho, e1, vol, rx, ry = random_chunk(nx, ny, 1, seed)
        d = L.Device(nx, ny, 1)
        try:
            d.set_options(L.CONDUCTIVITY, L.PREC_NONE, False)
            d.upload("density", rho)
            d.upload("energy1", e1)
            d.upload("volume", vol, halo=2)
            ref = _solve(d, rx, ry, "tl_cg_loop", n_last, eps)
            got = _solve(d, rx, ry, "tl_cg_loop_noalloc", n_last, eps)
        finally:
            d.close()
        (st0, al0, be0, u0), (st1, al1, be1, u1) = ref, got
        assert (st1.n_end, st1.reason) == (st0.n_end, st0.reason)
        if eps > 1e-200:
            assert st0.reason == L.LOOP_CONVERGED and st0.n_end < n_last
        else:
            assert st0.reason == L.LOOP_MAXITERS and st0.n_end == n_last
        assert (st1.error, st1.rro, st1.old_error) == (st0.error, st0.rro, st0.old_error)
        assert np.array_equal(al1, al0) and np.array_equal(be1, be0)
        assert np.array_equal(u1, u0)


def test_noalloc_rejects_long_loops():
    d = L.Device(40, 30, 1)
    try:
        st = L.LoopState(n_first=1, n_last=MAX_ITERS + 1, eps_abs=1.0, rro=1.0, old_error=1.0)
        rc = L.lib().tl_cg_loop_noalloc(C.byref(st), None, None)
        assert rc == L.ERR_ARG
        assert b"16384" in L.lib().tl_last_error()
    finally:
        d.close()

"""Error returns of the C ABI on a live device: every entry point checks its
arguments on the host before anything reaches the GPU, returns the
documented TL_ERR_* code with a message in tl_last_error(), and leaves the
library usable (the reference aborts through report_error, report.f90:29-58;
the C ABI returns a status instead)."""
import ctypes as C

import numpy as np
import pytest

from tealeaf_ref_amd import _lib

pytestmark = pytest.mark.gpu

ERR_STATE, ERR_ARG = 2, 3
NX, NY, H = 64, 48, 2


@pytest.fixture
def lib():
    L = _lib.lib()
    print(f"[abi errors] tl_init {NX}x{NY} halo {H}", flush=True)
    assert L.tl_init(NX, NY, H, 0) == 0, L.tl_last_error()
    try:
        yield L
    finally:
        L.tl_finalize()


def expect(rc, code, L, text=None):
    msg = L.tl_last_error().decode()
    assert rc == code, (rc, msg)
    assert msg
    if text:
        assert text in msg, msg


def test_rejected_arguments(lib):
    L = lib
    d = C.c_double()
    n = C.c_long()
    buf = np.zeros((NY + 2 * H, NX + 2 * H))
    ptr = buf.ctypes.data_as(C.POINTER(C.c_double))
    expect(L.tl_init(NX, NY, H, 0), ERR_STATE, L, "twice")
    expect(L.tl_upload(99, ptr, H), ERR_ARG, L, "field id")
    expect(L.tl_upload(_lib.FIELD["u"], None, H), ERR_ARG, L, "null")
    expect(L.tl_download(_lib.FIELD["u"], ptr, -1), ERR_ARG, L, "halo")
    expect(L.tl_set_options(7, 1, 0), ERR_ARG, L, "coefficient")
    expect(L.tl_set_options(1, 4, 0), ERR_ARG, L, "preconditioner")
    expect(L.tl_set_options(1, 3, 0), ERR_ARG, L, "block jacobi")  # halo 2
    expect(L.tl_calc_2norm(2, C.byref(d)), ERR_ARG, L, "norm_array")
    expect(L.tl_calc_2norm(1, None), ERR_ARG, L)
    fields = (C.c_int * 11)(*([1] * 11))
    expect(L.tl_update_halo(fields, 0), ERR_ARG, L, "depth")
    expect(L.tl_update_halo(fields, H + 1), ERR_ARG, L, "depth")
    expect(L.tl_snapshot_save(0, _lib.FIELD["u"]), ERR_ARG, L, "not reserved")
    expect(L.tl_snapshot_reserve(-1), ERR_ARG, L)
    st = _lib.LoopState(n_first=5, n_last=4)
    expect(L.tl_cg_loop(C.byref(st), None, None), ERR_ARG, L, "n_first")
    expect(L.tl_cg_loop(None, None, None), ERR_ARG, L, "null")
    ms = _lib.MixedState(solver=9)
    expect(L.tl_mixed_solve(C.byref(ms)), ERR_ARG, L, "solver")
    expect(L.tl_profile_enable(-1), ERR_ARG, L)
    expect(L.tl_profile_get(8, C.byref(n), C.byref(d)), ERR_ARG, L, "class")
    expect(L.tl_profile_idle(-1, C.byref(n)), ERR_ARG, L, "class")


def test_usable_after_errors(lib):
    """A rejected call changes nothing: a field round trip afterwards gives
    back what went up, halo cells included."""
    L = lib
    d = C.c_double()
    assert L.tl_calc_2norm(3, C.byref(d)) == ERR_ARG
    rng = np.random.default_rng(5)
    up = rng.random((NY + 2 * H, NX + 2 * H))
    down = np.zeros_like(up)
    assert L.tl_upload(_lib.FIELD["energy0"], up.ctypes.data_as(C.POINTER(C.c_double)), H) == 0
    assert L.tl_upload(99, up.ctypes.data_as(C.POINTER(C.c_double)), H) == ERR_ARG
    assert L.tl_download(_lib.FIELD["energy0"], down.ctypes.data_as(C.POINTER(C.c_double)),
                         H) == 0
    assert np.array_equal(up, down)
    assert L.tl_synchronize() == 0


def test_state_errors_after_finalize():
    L = _lib.lib()
    print("[abi errors] init / finalize / use", flush=True)
    assert L.tl_init(NX, NY, 1, 0) == 0, L.tl_last_error()
    assert L.tl_finalize() == 0
    d = C.c_double()
    expect(L.tl_calc_2norm(1, C.byref(d)), ERR_STATE, L, "tl_init")
    # and it initialises again
    assert L.tl_init(NX, NY, 1, 0) == 0, L.tl_last_error()
    assert L.tl_finalize() == 0

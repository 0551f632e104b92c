"""ctypes binding of libtealeaf_hip.so (the C ABI declared in include/tealeaf_hip.h).

There is no CPU fallback: if the library is missing or no HIP device is
present, every call raises.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libtealeaf_hip.so")

TL_OK = 0
CONDUCTIVITY, RECIP_CONDUCTIVITY = 1, 2
PREC_NONE, PREC_JAC_DIAG, PREC_JAC_BLOCK = 1, 2, 3
LOOP_CONVERGED, LOOP_SWITCH, LOOP_MAXITERS, LOOP_NONFINITE, LOOP_FLOOR = 1, 2, 3, 4, 5
ERR_HIP, ERR_STATE, ERR_ARG, ERR_NONFINITE, ERR_EIGEN = 1, 2, 3, 4, 5

FIELD = {"density": 1, "energy0": 2, "energy1": 3, "u": 4, "p": 5, "sd": 6, "r": 7, "z": 8,
         "kx": 9, "ky": 10, "di": 11, "u0": 12, "w": 13, "r_store": 14, "mi": 15,
         "utemp": 16, "rtemp": 17, "volume": 18, "cp": 19, "bfp": 20}

MIXED_CG, MIXED_PPCG = 1, 2

# kernel classes of tl_profile_get
(PROF_CG_W, PROF_CG_UR, PROF_PPCG_SWEEP, PROF_CHEBY, PROF_JACOBI, PROF_PPCG_W, PROF_PPCG_UR,
 PROF_MIXED_REFINE) = range(8)


class TeaLeafError(RuntimeError):
    pass


class LoopState(C.Structure):
    _fields_ = [("n_first", C.c_int), ("n_last", C.c_int), ("eps_abs", C.c_double),
                ("check_switch", C.c_int), ("epslim", C.c_double), ("presteps", C.c_int),
                ("rro", C.c_double), ("old_error", C.c_double), ("n_end", C.c_int),
                ("reason", C.c_int), ("error", C.c_double)]


class TLState(C.Structure):
    _fields_ = [("geometry", C.c_int), ("density", C.c_double), ("energy", C.c_double),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("radius", C.c_double)]


class MixedState(C.Structure):
    _fields_ = [("solver", C.c_int), ("max_outer", C.c_int), ("max_inner", C.c_int),
                ("inner_steps", C.c_int), ("presteps", C.c_int), ("rel_tol", C.c_double),
                ("inner_tol", C.c_double), ("eigmin", C.c_double), ("eigmax", C.c_double),
                ("outer", C.c_int), ("inner_iters", C.c_int), ("sweeps", C.c_int),
                ("reason", C.c_int), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("eig_iters", C.c_int)]


_lib = None


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TeaLeafError(
                "libtealeaf_hip.so is not built; run __graft_entry__.build() or "
                "make -C tealeaf_ref_amd/csrc")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        I, D, P = C.c_int, C.c_double, C.POINTER
        PD = P(D)
        sig = {
            "tl_init": [I, I, I, I], "tl_finalize": [], "tl_set_options": [I, I, I],
            "tl_set_ppcg_active": [I], "tl_upload": [I, PD, I], "tl_download": [I, PD, I],
            "tl_synchronize": [], "tl_common_init": [D, D, P(I)], "tl_calc_residual": [],
            "tl_calc_2norm": [I, PD], "tl_finalise": [], "tl_set_field": [],
            "tl_update_halo": [P(I), I], "tl_field_summary": [PD, PD, PD, PD],
            "tl_cg_init": [PD], "tl_cg_calc_w": [PD], "tl_cg_calc_ur": [D, PD],
            "tl_cg_calc_p": [D], "tl_cheby_init": [D], "tl_cheby_iterate": [D, D],
            "tl_ppcg_init": [I, PD], "tl_ppcg_init_sd": [D], "tl_ppcg_inner": [PD, PD, I, I],
            "tl_ppcg_store_r": [], "tl_ppcg_calc_zrnorm": [PD], "tl_ppcg_calc_rrn": [PD],
            "tl_ppcg_update_z": [], "tl_jacobi_solve": [PD],
            "tl_cg_loop": [P(LoopState), PD, PD], "tl_cg_loop_noalloc": [P(LoopState), PD, PD],
            "tl_ppcg_loop": [P(LoopState), PD, PD, I, D, P(I)],
            "tl_jacobi_loop": [P(LoopState)], "tl_loop_errors": [I, I, PD],
            "tl_cheby_run": [PD, PD, I, I, I, PD],
            "tl_mixed_solve": [P(MixedState)],
            "tl_generate": [D, D, D, D, I, P(TLState)],
            "tl_snapshot_reserve": [I], "tl_snapshot_save": [I, I], "tl_snapshot_load": [I, I],
            "tl_snapshot_fill": [I, D], "tl_snapshot_upload": [I, PD, I],
            "tl_snapshot_download": [I, PD, I], "tl_adjoint_face_grad": [I, I, I, I],
            "tl_device_bytes": [PD], "tl_profile_enable": [I], "tl_profile_get": [I, P(C.c_long), PD],
            "tl_profile_idle": [I, P(C.c_long)],
            "tl_profile_reset": [], "tl_hbm_stream": [C.c_long, I, PD, PD],
            # device addresses (e.g. torch's data_ptr()) and a row stride
            "tl_upload_device": [I, C.c_void_p, C.c_long],
            "tl_download_device": [I, C.c_void_p, C.c_long],
            "tl_snapshot_upload_device": [I, C.c_void_p, C.c_long],
            "tl_snapshot_download_device": [I, C.c_void_p, C.c_long],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I
        L.tl_last_error.argtypes = []
        L.tl_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != TL_OK:
        msg = lib().tl_last_error().decode(errors="replace")
        raise TeaLeafError(f"{what}: error {rc}: {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def download_field(name, nx, ny, halo=1):
    """Copy a field of the chunk the library currently holds (whoever set it
    up) to a new host array (ny+2*halo, nx+2*halo)."""
    a = np.zeros((ny + 2 * halo, nx + 2 * halo))
    call("tl_download", FIELD[name], _dptr(a), halo)
    return a


class Device:
    """A chunk resident on one GPU, driven through the C ABI."""

    def __init__(self, nx, ny, halo=1, device=0):
        self.nx, self.ny, self.h = nx, ny, halo
        call("tl_init", nx, ny, halo, device)

    def close(self):
        lib().tl_finalize()

    def set_options(self, coefficient=CONDUCTIVITY, precond=PREC_NONE, reflective=False):
        call("tl_set_options", coefficient, precond, int(bool(reflective)))

    def upload(self, name, arr, halo=None):
        halo = self.h if halo is None else halo
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (self.ny + 2 * halo, self.nx + 2 * halo), a.shape
        call("tl_upload", FIELD[name], _dptr(a), halo)

    def download(self, name, halo=None, out=None):
        halo = self.h if halo is None else halo
        shape = (self.ny + 2 * halo, self.nx + 2 * halo)
        a = np.zeros(shape) if out is None else out
        call("tl_download", FIELD[name], _dptr(a), halo)
        return a

    def scalar(self, name, *args):
        out = C.c_double()
        call(name, *args, C.byref(out))
        return out.value

    def field_summary(self):
        v = [C.c_double() for _ in range(4)]
        call("tl_field_summary", *[C.byref(x) for x in v])
        return [x.value for x in v]

    def update_halo(self, fields, depth):
        f = (C.c_int * 11)(*fields)
        call("tl_update_halo", f, depth)

    def common_init(self, rx, ry, zero_boundary=(1, 1, 1, 1)):
        zb = (C.c_int * 4)(*zero_boundary)
        call("tl_common_init", rx, ry, zb)

    def mixed_solve(self, solver=MIXED_PPCG, rel_tol=1e-10, inner_tol=1e-5, max_outer=50,
                    max_inner=20000, inner_steps=10, presteps=30, eigmin=0.0, eigmax=0.0):
        """FP32-inner / FP64-refined solve of the current step (tl_mixed_solve)."""
        st = MixedState(solver=solver, max_outer=max_outer, max_inner=max_inner,
                        inner_steps=inner_steps, presteps=presteps, rel_tol=rel_tol,
                        inner_tol=inner_tol, eigmin=eigmin, eigmax=eigmax)
        call("tl_mixed_solve", C.byref(st))
        return st

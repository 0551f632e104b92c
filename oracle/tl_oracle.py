"""TEST INFRASTRUCTURE ONLY: ctypes binding of the C restatement (liboracle.so).

Imported by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
bench.py, as the checker.  The product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
# the same source built with multiply-add contraction (a second rounding)
_LIB_FMA = os.path.join(_HERE, "liboracle_fma.so")

PREC = {"none": 1, "jac_diag": 2, "jac_block": 3}
FIELD_IDS = {"density": 1, "energy0": 2, "energy1": 3, "u": 4, "p": 5, "sd": 6,
             "r": 7, "z": 8, "kx": 9, "ky": 10, "di": 11, "u0": 12, "w": 13,
             "r_store": 14, "mi": 15, "utemp": 16, "rtemp": 17}
MAX_STATES = 64


class State(C.Structure):
    _fields_ = [("defined", C.c_int), ("geometry", C.c_int), ("density", C.c_double),
                ("energy", C.c_double), ("xmin", C.c_double), ("xmax", C.c_double),
                ("ymin", C.c_double), ("ymax", C.c_double), ("radius", C.c_double)]


class Params(C.Structure):
    _fields_ = [("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double),
                ("ymax", C.c_double), ("x_cells", C.c_int), ("y_cells", C.c_int),
                ("end_time", C.c_double), ("dtinit", C.c_double), ("eps", C.c_double),
                ("tl_ch_cg_epslim", C.c_double), ("end_step", C.c_int),
                ("visit_frequency", C.c_int), ("summary_frequency", C.c_int),
                ("max_iters", C.c_int), ("tl_ch_cg_presteps", C.c_int),
                ("tl_ppcg_inner_steps", C.c_int), ("tl_check_result", C.c_int),
                ("precond", C.c_int), ("coefficient", C.c_int), ("reflective", C.c_int),
                ("halo_depth", C.c_int), ("test_problem", C.c_int), ("use_cg", C.c_int),
                ("use_cheby", C.c_int), ("use_ppcg", C.c_int), ("use_jacobi", C.c_int),
                ("profiler_on", C.c_int), ("verbose_on", C.c_int),
                ("number_of_states", C.c_int), ("states", State * (MAX_STATES + 1))]


class Chunk(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("h", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("rx", C.c_double), ("ry", C.c_double)] + \
               [(n, C.POINTER(C.c_double)) for n in (
                   "density", "energy0", "energy1", "u", "u0", "p", "r", "r_store", "Mi",
                   "w", "z", "utemp", "rtemp", "Di", "Kx", "Ky", "sd", "cp", "bfp",
                   "volume", "vertexx", "vertexy", "cellx", "celly")]


class StepRecord(C.Structure):
    _fields_ = [("step", C.c_int), ("iterations", C.c_int), ("ppcg_inner_iters", C.c_int),
                ("switch_n", C.c_int), ("est_itc", C.c_int),
                ("conduction_error", C.c_double), ("initial_residual", C.c_double),
                ("exact_error", C.c_double), ("eigmin", C.c_double), ("eigmax", C.c_double),
                ("has_summary", C.c_int), ("summary", C.c_double * 5)]


_lib = None
_lib_fma = None


def lib(contracted=False):
    """The plain build, or (contracted=True) the multiply-add build."""
    global _lib, _lib_fma
    if contracted:
        if _lib_fma is None:
            if not os.path.exists(_LIB_FMA):
                raise RuntimeError("oracle/liboracle_fma.so is not built (run make -C oracle)")
            _lib_fma = _configure(C.CDLL(_LIB_FMA))
        return _lib_fma
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _lib = _configure(C.CDLL(_LIB))
    return _lib


def _configure(L):
    P, D, I = C.POINTER, C.c_double, C.c_int
    PC = P(Chunk)
    L.tlo_parse_deck.argtypes = [C.c_char_p, P(Params), C.c_char_p, I]
    L.tlo_run.argtypes = [P(Params), P(StepRecord), I, I, C.c_char_p, I]
    L.tlo_chunk_create.argtypes = [I, I, I]
    L.tlo_chunk_create.restype = PC
    L.tlo_chunk_destroy.argtypes = [PC]
    L.tlo_field.argtypes = [PC, I]
    L.tlo_field.restype = P(D)
    L.tlo_initialise_chunk.argtypes = [PC, P(Params)]
    L.tlo_generate_chunk.argtypes = [PC, P(Params)]
    L.tlo_set_field.argtypes = [PC]
    L.tlo_common_init.argtypes = [PC, P(I), I, I, I]
    for name in ("tlo_finalise", "tlo_calc_residual", "tlo_diag_init", "tlo_block_init",
                 "tlo_ppcg_update_z", "tlo_ppcg_store_r"):
        getattr(L, name).argtypes = [PC]
    for name in ("tlo_calc_2norm",):
        getattr(L, name).argtypes = [PC, I]
        getattr(L, name).restype = D
    L.tlo_cg_init.argtypes = [PC, I]
    L.tlo_cg_init.restype = D
    L.tlo_cg_calc_w.argtypes = [PC]
    L.tlo_cg_calc_w.restype = D
    L.tlo_cg_calc_ur.argtypes = [PC, D, I]
    L.tlo_cg_calc_ur.restype = D
    L.tlo_cg_calc_p.argtypes = [PC, D, I, I]
    L.tlo_cheby_init.argtypes = [PC, D, I]
    L.tlo_cheby_iterate.argtypes = [PC, D, D, I]
    L.tlo_ppcg_init_sd.argtypes = [PC, D, I]
    L.tlo_ppcg_init.argtypes = [PC, I, I, I]
    L.tlo_ppcg_init.restype = D
    L.tlo_ppcg_inner.argtypes = [PC, D, D, I]
    L.tlo_ppcg_calc_zrnorm.argtypes = [PC, I, I]
    L.tlo_ppcg_calc_zrnorm.restype = D
    L.tlo_ppcg_calc_rrn.argtypes = [PC]
    L.tlo_ppcg_calc_rrn.restype = D
    L.tlo_jacobi_solve.argtypes = [PC]
    L.tlo_jacobi_solve.restype = D
    L.tlo_update_halo.argtypes = [PC, P(I), I, I]
    L.tlo_field_summary.argtypes = [PC, P(D)]
    L.tlo_tqli.argtypes = [P(D), P(D), I]
    L.tlo_calc_eigenvalues.argtypes = [P(D), P(D), I, P(D), P(D)]
    L.tlo_calc_ch_coefs.argtypes = [P(D), P(D), D, D, P(D), I]
    return L


class OracleError(RuntimeError):
    pass


def parse_deck(text):
    p = Params()
    msg = C.create_string_buffer(512)
    rc = lib().tlo_parse_deck(text.encode(), C.byref(p), msg, 512)
    if rc != 0:
        raise OracleError(msg.value.decode())
    return p


def run(params, max_steps=0, contracted=False):
    """Run a parsed deck; returns a list of dict records (record 0 = step 0).
    contracted=True runs the multiply-add build of the same source."""
    if isinstance(params, str):
        params = parse_deck(params)
    n = (params.end_step if params.end_step < 100000 else 1000) + 2
    if max_steps:
        n = min(n, max_steps + 2)
    recs = (StepRecord * n)()
    msg = C.create_string_buffer(512)
    rc = lib(contracted).tlo_run(C.byref(params), recs, n, max_steps, msg, 512)
    if rc < 0:
        raise OracleError(msg.value.decode())
    out = []
    for i in range(rc):
        r = recs[i]
        out.append({"step": r.step, "iterations": r.iterations,
                    "ppcg_inner_iters": r.ppcg_inner_iters, "switch_n": r.switch_n,
                    "est_itc": r.est_itc, "conduction_error": r.conduction_error,
                    "initial_residual": r.initial_residual, "exact_error": r.exact_error,
                    "eigmin": r.eigmin, "eigmax": r.eigmax,
                    "summary": list(r.summary) if r.has_summary else None})
    return out


class OracleChunk:
    """One haloed chunk in the reference layout, for per-kernel checks."""

    def __init__(self, nx, ny, h):
        self.nx, self.ny, self.h = nx, ny, h
        self._c = lib().tlo_chunk_create(nx, ny, h)

    def __del__(self):
        if getattr(self, "_c", None):
            lib().tlo_chunk_destroy(self._c)
            self._c = None

    @property
    def c(self):
        return self._c.contents

    def field(self, name):
        """numpy view (ny+2h, nx+2h) of a haloed field; x is the fast index."""
        ptr = lib().tlo_field(self._c, FIELD_IDS[name])
        shape = (self.ny + 2 * self.h, self.nx + 2 * self.h)
        return np.ctypeslib.as_array(ptr, shape=shape)

    def set_rxry(self, rx, ry):
        self.c.rx = rx
        self.c.ry = ry

    # kernels -------------------------------------------------------------
    def common_init(self, reflective=False, coef=1, precond=1, zero_boundary=(1, 1, 1, 1)):
        zb = (C.c_int * 4)(*zero_boundary)
        lib().tlo_common_init(self._c, zb, int(reflective), coef, precond)

    def __getattr__(self, name):
        fn = getattr(lib(), "tlo_" + name)
        return lambda *a: fn(self._c, *a)

    def update_halo(self, fields, depth, reflective=True):
        f = (C.c_int * 11)(*fields)
        lib().tlo_update_halo(self._c, f, depth, int(reflective))

    def field_summary(self):
        out = (C.c_double * 4)()
        lib().tlo_field_summary(self._c, out)
        return list(out)


def tqli(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    e = np.ascontiguousarray(e, dtype=np.float64).copy()
    dp = d.ctypes.data_as(C.POINTER(C.c_double))
    ep = e.ctypes.data_as(C.POINTER(C.c_double))
    info = lib().tlo_tqli(dp, ep, len(d))
    return info, d


def calc_eigenvalues(cg_alphas, cg_betas):
    """cg_alphas/betas: sequences of length n (1-based in C, index 0 padded)."""
    n = len(cg_alphas)
    a = np.concatenate([[0.0], np.asarray(cg_alphas, float)])
    b = np.concatenate([[0.0], np.asarray(cg_betas, float)])
    mn, mx = C.c_double(), C.c_double()
    info = lib().tlo_calc_eigenvalues(a.ctypes.data_as(C.POINTER(C.c_double)),
                                      b.ctypes.data_as(C.POINTER(C.c_double)), n,
                                      C.byref(mn), C.byref(mx))
    return info, mn.value, mx.value


def calc_ch_coefs(eigmin, eigmax, n):
    a = np.zeros(n + 1)
    b = np.zeros(n + 1)
    th = C.c_double()
    lib().tlo_calc_ch_coefs(a.ctypes.data_as(C.POINTER(C.c_double)),
                            b.ctypes.data_as(C.POINTER(C.c_double)), eigmin, eigmax,
                            C.byref(th), n)
    return a[1:], b[1:], th.value

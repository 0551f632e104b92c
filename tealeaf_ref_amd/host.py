"""Python handle on the native Fortran host (libtealeaf_host.so).

The host reads a tea.in deck, writes tea.out in the reference's formats and
drives libtealeaf_hip.so; this module only calls its C entry points.  A run
can be stepped (``step()``) so callers can time individual timesteps.
"""
import ctypes as C
import os

from . import _lib
from . import teaout

_PKG = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(_PKG, "libtealeaf_host.so")
TEA_LEAF_BIN = os.path.join(_PKG, "bin", "tea_leaf")

_host = None


def host_lib():
    global _host
    if _host is None:
        _lib.lib()  # the solver library first, so both resolve to one instance
        if not os.path.exists(HOST_LIB):
            raise _lib.TeaLeafError("libtealeaf_host.so is not built; run __graft_entry__.build()")
        L = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL)
        for name in ("tea_host_init", "tea_host_guarded_init"):
            getattr(L, name).argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int]
            getattr(L, name).restype = C.c_int
        for name in ("tea_host_step", "tea_host_run", "tea_host_finalize", "tea_host_guarded_step",
                     "tea_host_guarded_run", "tea_host_guarded_finalize"):
            getattr(L, name).argtypes = []
            getattr(L, name).restype = C.c_int
        L.tea_host_last.argtypes = [C.POINTER(C.c_double)]
        L.tea_host_last.restype = C.c_int
        L.tea_host_error.argtypes = []
        L.tea_host_error.restype = C.c_char_p
        _host = L
    return _host


# returned by a guarded entry point when the host reported an error (the
# reference's report_error, report.f90:29-58, which ends the program there)
HOST_ERROR = -2


def _guarded(rc):
    """Turn a report_error inside a guarded host call into TeaLeafError.  The
    run is abandoned: the host is finalized so that another can start."""
    if rc != HOST_ERROR:
        return rc
    L = host_lib()
    msg = L.tea_host_error().decode(errors="replace")
    L.tea_host_guarded_finalize()
    raise _lib.TeaLeafError(msg)


class TeaHost:
    """One run of a deck.  Only one may be open per process."""

    def __init__(self, deck_path, out_path, device=0):
        L = host_lib()
        d = os.fsencode(deck_path)
        o = os.fsencode(out_path)
        self.out_path = out_path
        rc = _guarded(L.tea_host_guarded_init(d, len(d), o, len(o), device))
        if rc != 0:
            raise _lib.TeaLeafError(f"tea_host_init failed ({rc})")
        self.complete = False

    def step(self):
        """One timestep; returns True when the run is complete.  A deck or
        device error raises TeaLeafError and ends the run."""
        rc = _guarded(host_lib().tea_host_guarded_step())
        if rc < 0:
            raise _lib.TeaLeafError("tea_host_step: host not initialised")
        self.complete = rc == 1
        return self.complete

    def run(self):
        rc = _guarded(host_lib().tea_host_guarded_run())
        if rc < 0:
            raise _lib.TeaLeafError("tea_host_run: host not initialised")
        return rc

    def last(self):
        out = (C.c_double * 14)()
        host_lib().tea_host_last(out)
        keys = ("step", "time", "iterations", "ppcg_inner", "conduction_error", "exact_error",
                "solve_time", "switch_n", "summary_step", "volume", "mass", "energy",
                "temperature", "refinements")
        return dict(zip(keys, list(out)))

    def close(self):
        _guarded(host_lib().tea_host_guarded_finalize())

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run_deck(deck_text, workdir):
    """Write deck_text to workdir/tea.in, run it to completion in-process and
    return the parsed tea.out."""
    os.makedirs(workdir, exist_ok=True)
    deck = os.path.join(workdir, "tea.in")
    out = os.path.join(workdir, "tea.out")
    with open(deck, "w") as f:
        f.write(deck_text)
    with TeaHost(deck, out) as h:
        h.run()
    with open(out) as f:
        return teaout.parse(f.read())

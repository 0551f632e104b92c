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
        L.tea_host_init.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int]
        L.tea_host_init.restype = C.c_int
        L.tea_host_step.argtypes = []
        L.tea_host_step.restype = C.c_int
        L.tea_host_run.argtypes = []
        L.tea_host_run.restype = C.c_int
        L.tea_host_last.argtypes = [C.POINTER(C.c_double)]
        L.tea_host_last.restype = C.c_int
        L.tea_host_finalize.argtypes = []
        L.tea_host_finalize.restype = C.c_int
        _host = L
    return _host


class TeaHost:
    """One run of a deck.  Only one may be open per process."""

    def __init__(self, deck_path, out_path, device=0):
        L = host_lib()
        d = os.fsencode(deck_path)
        o = os.fsencode(out_path)
        self.out_path = out_path
        rc = L.tea_host_init(d, len(d), o, len(o), device)
        if rc != 0:
            raise _lib.TeaLeafError(f"tea_host_init failed ({rc})")
        self.complete = False

    def step(self):
        """One timestep; returns True when the run is complete."""
        rc = host_lib().tea_host_step()
        if rc < 0:
            raise _lib.TeaLeafError("tea_host_step: host not initialised")
        self.complete = rc == 1
        return self.complete

    def run(self):
        return host_lib().tea_host_run()

    def last(self):
        out = (C.c_double * 14)()
        host_lib().tea_host_last(out)
        keys = ("step", "time", "iterations", "ppcg_inner", "conduction_error", "exact_error",
                "solve_time", "switch_n", "summary_step", "volume", "mass", "energy",
                "temperature", "refinements")
        return dict(zip(keys, list(out)))

    def close(self):
        host_lib().tea_host_finalize()

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

"""Non-finite reductions in the fused loops.

The reference does not check for NaN or Inf: a state with density 0 makes the
face coefficients 0/0 or x/0 (kernels/tea_leaf_common_kernel.f90:97-98), the
initial residual is NaN, and every solver runs to tl_max_iters (the oracle
reproduces that: conduction error NaN after max_iters).  The library stops the
loop at the first non-finite p.w, r.r, z.r or Jacobi error instead and returns
TL_ERR_NONFINITE, which the host turns into the reference's report_error
abort (in process: TeaLeafError).  These are ordinary NaN values in double
arithmetic, not faults."""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

import fixtures_util as FX
import tl_oracle as O
from tealeaf_ref_amd import _lib as L
from tealeaf_ref_amd import host
from test_gpu_kernels import pair_factory  # noqa: F401  (fixture)
from test_gpu_loops import _setup

pytestmark = pytest.mark.gpu

ZERO_DENSITY = """*tea
state 1 density=100.0 energy=0.0001
state 2 density=0.0 energy=25.0 geometry=rectangle xmin=2.0 xmax=5.0 ymin=2.0 ymax=5.0
x_cells=96
y_cells=80
xmin=0.0
ymin=0.0
xmax=10.0
ymax=10.0
initial_timestep=0.004
end_step=2
tl_max_iters=10000
tl_eps=1.0e-12
{solver}
*endtea
"""


@pytest.mark.parametrize("solver", ["tl_use_cg", "tl_use_chebyshev", "tl_use_ppcg",
                                    "tl_use_jacobi",
                                    "tl_use_cg\ntl_preconditioner_type jac_diag"])
def test_zero_density_deck_raises(solver):
    print(f"[nonfinite] deck {solver!r}", flush=True)
    deck = ZERO_DENSITY.format(solver=solver)
    # the oracle (the reference's arithmetic) runs to the limit on NaN
    recs = O.run(O.parse_deck(deck.replace("tl_max_iters=10000", "tl_max_iters=40")))
    assert math.isnan(recs[1]["conduction_error"])
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "tea.in"), "w") as f:
            f.write(deck)
        with pytest.raises(L.TeaLeafError, match="non-finite"):
            h = host.TeaHost(os.path.join(d, "tea.in"), os.path.join(d, "tea.out"))
            h.run()
        with open(os.path.join(d, "tea.out")) as f:
            text = f.read()
        assert "TEA is terminating." in text and "non-finite" in text
        # stopped at the first iterations, not after 10000
        assert "Iteration count" not in text
    # the library is usable again in the same process
    fx = FX.load("tea_bm_1")
    with tempfile.TemporaryDirectory() as d:
        got = host.run_deck(fx["deck"], d)
    FX.compare_summaries(got["summaries"], fx["summaries"], 1e-10)


def _poison(p, name, value=float("nan")):
    a = p.d.download(name)
    a[p.h + p.ny // 2, p.h + p.nx // 3] = value
    p.d.upload(name, a)


def _loop_state(n_last, rro, r0):
    return L.LoopState(n_first=1, n_last=n_last, eps_abs=1e-300, check_switch=0, epslim=1.0,
                       presteps=25, rro=rro, old_error=r0)


@pytest.mark.parametrize("field,n_end", [("p", 1), ("r", 1)])
def test_cg_loop_stops(pair_factory, field, n_end):
    """NaN in p: p.w of iteration 1 (update pass); NaN in r: r.r of iteration
    1, found by the head of iteration 2.  Both stop at iteration 1."""
    p = pair_factory(96, 80, 1, seed=17)
    r0, rro = _setup(p)
    _poison(p, field)
    st = _loop_state(5000, rro, r0)
    rc = L.lib().tl_cg_loop(C.byref(st), None, None)
    assert rc == L.ERR_NONFINITE, (rc, L.lib().tl_last_error())
    assert st.reason == L.LOOP_NONFINITE and st.n_end == n_end, (st.reason, st.n_end)
    assert b"non-finite" in L.lib().tl_last_error()


@pytest.mark.parametrize("value", [float("nan"), float("inf")])
def test_jacobi_loop_stops(pair_factory, value):
    p = pair_factory(96, 80, 1, seed=19)
    _setup(p)
    _poison(p, "u0", value)
    st = _loop_state(5000, 0.0, 0.0)
    rc = L.lib().tl_jacobi_loop(C.byref(st))
    assert rc == L.ERR_NONFINITE, (rc, L.lib().tl_last_error())
    assert st.reason == L.LOOP_NONFINITE and st.n_end == 1, (st.reason, st.n_end)


def test_ppcg_loop_stops(pair_factory):
    p = pair_factory(77, 58, 1, seed=31)
    _setup(p)
    ch_a, ch_b, theta = O.calc_ch_coefs(0.9, 40.0, 6)
    L.call("tl_set_ppcg_active", 1)
    rro = p.d.scalar("tl_ppcg_init", 2)
    _poison(p, "p")
    st = _loop_state(3000, rro, 1.0)
    a = np.ascontiguousarray(ch_a[:6])
    b = np.ascontiguousarray(ch_b[:6])
    inner = C.c_int(0)
    rc = L.lib().tl_ppcg_loop(C.byref(st), a.ctypes.data_as(C.POINTER(C.c_double)),
                              b.ctypes.data_as(C.POINTER(C.c_double)), 6, theta,
                              C.byref(inner))
    assert rc == L.ERR_NONFINITE, (rc, L.lib().tl_last_error())
    assert st.reason == L.LOOP_NONFINITE and st.n_end == 1, (st.reason, st.n_end)


def test_cheby_run_stops(pair_factory):
    p = pair_factory(64, 48, 1, seed=5)
    _setup(p)
    ch_a, ch_b, theta = O.calc_ch_coefs(0.9, 40.0, 30)
    L.call("tl_cheby_init", theta)
    _poison(p, "u")
    a = np.ascontiguousarray(ch_a)
    b = np.ascontiguousarray(ch_b)
    norm = C.c_double(0.0)
    rc = L.lib().tl_cheby_run(a.ctypes.data_as(C.POINTER(C.c_double)),
                              b.ctypes.data_as(C.POINTER(C.c_double)), 2, 10, 1, C.byref(norm))
    assert rc == L.ERR_NONFINITE, (rc, L.lib().tl_last_error())
    assert not math.isfinite(norm.value)

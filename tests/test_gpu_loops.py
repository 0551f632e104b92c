"""Fused device-side solver loops (tl_cg_loop, tl_ppcg_loop, tl_jacobi_loop,
tl_cheby_run) against the oracle's routines driven one iteration at a time as
tea_solve.f90:152-414 does.  Iteration counts must be identical; fields and
scalars agree to 1e-10 relative."""
import ctypes as C
import math

import numpy as np
import pytest

import tl_oracle as O
from helpers import assert_scalar
from tealeaf_ref_amd import _lib as L
from test_gpu_kernels import Pair, pair_factory  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def _setup(p):
    p.init()
    p.o.calc_residual()
    L.call("tl_calc_residual")
    r0 = p.o.calc_2norm(1)
    rro = p.o.cg_init(p.precond)
    p.d.scalar("tl_cg_init")
    return r0, rro


def oracle_cg_loop(o, precond, n_first, n_last, rro, old_error, eps_abs, check_switch=0,
                   epslim=1.0, presteps=25):
    al, be = {}, {}
    err = 0.0
    for n in range(n_first, n_last + 1):
        if n > n_first and check_switch and abs(old_error) <= epslim and n >= presteps:
            return dict(n_end=n, reason=L.LOOP_SWITCH, rro=rro, error=err, al=al, be=be)
        pw = o.cg_calc_w()
        a = rro / pw
        al[n] = a
        rrn = o.cg_calc_ur(a, precond)
        b = rrn / rro
        be[n] = b
        o.cg_calc_p(b, precond, 0)
        rro = rrn
        err = math.sqrt(abs(rrn))
        if abs(err) < eps_abs:
            return dict(n_end=n, reason=L.LOOP_CONVERGED, rro=rro, error=err, al=al, be=be)
        old_error = err
    return dict(n_end=n_last, reason=L.LOOP_MAXITERS, rro=rro, error=err, al=al, be=be)


def run_gpu_cg_loop(n_first, n_last, rro, old_error, eps_abs, check_switch=0, epslim=1.0,
                    presteps=25):
    st = L.LoopState(n_first=n_first, n_last=n_last, eps_abs=eps_abs, check_switch=check_switch,
                     epslim=epslim, presteps=presteps, rro=rro, old_error=old_error)
    al = np.zeros(n_last + 2)
    be = np.zeros(n_last + 2)
    L.call("tl_cg_loop", C.byref(st), al.ctypes.data_as(C.POINTER(C.c_double)),
           be.ctypes.data_as(C.POINTER(C.c_double)))
    return st, al, be


@pytest.mark.parametrize("precond", [1, 2, 3])
@pytest.mark.parametrize("batch_case", ["converge", "maxiters"])
def test_cg_loop(pair_factory, precond, batch_case):
    p = pair_factory(96, 80, 1, seed=17, precond=precond)
    r0, rro = _setup(p)
    r0s = math.sqrt(r0)
    eps_abs = 1e-9 * r0s if batch_case == "converge" else 1e-300
    n_last = 400 if batch_case == "converge" else 37
    ref = oracle_cg_loop(p.o, precond, 1, n_last, rro, r0, eps_abs)
    st, al, be = run_gpu_cg_loop(1, n_last, rro, r0, eps_abs)
    assert st.reason == ref["reason"]
    assert st.n_end == ref["n_end"], (st.n_end, ref["n_end"])
    assert_scalar(st.rro, ref["rro"], 1e-8, "rro")
    for n in range(1, ref["n_end"] + 1):
        assert_scalar(al[n], ref["al"][n], 1e-9, f"alpha {n}")
        assert_scalar(be[n], ref["be"][n], 1e-8, f"beta {n}")
    p.cmp("u", tol=1e-10)


def test_cg_loop_switch_and_resume(pair_factory):
    """Switch test (tea_solve.f90:167) stops at the top of the iteration, and a
    second call continues with the state the first left."""
    p = pair_factory(64, 64, 1, seed=23)
    r0, rro = _setup(p)
    eps_abs = 1e-12 * math.sqrt(r0)
    # epslim large: the switch fires as soon as n >= presteps
    ref = oracle_cg_loop(p.o, 1, 1, 300, rro, r0, eps_abs, check_switch=1, epslim=1e30,
                         presteps=7)
    st, al, be = run_gpu_cg_loop(1, 300, rro, r0, eps_abs, check_switch=1, epslim=1e30,
                                 presteps=7)
    assert ref["reason"] == L.LOOP_SWITCH and ref["n_end"] == 7
    assert st.reason == L.LOOP_SWITCH and st.n_end == 7
    for n in range(1, 7):
        assert_scalar(al[n], ref["al"][n], 1e-9, f"alpha {n}")
    p.cmp("u", tol=1e-10)
    # resume plain CG from iteration 7 on both sides
    ref2 = oracle_cg_loop(p.o, 1, 7, 300, ref["rro"], ref["error"], eps_abs)
    st2, _, _ = run_gpu_cg_loop(7, 300, st.rro, st.old_error, eps_abs)
    assert st2.n_end == ref2["n_end"] and st2.reason == ref2["reason"]
    p.cmp("u", tol=1e-10)


# (preconditioner, inner steps): the fused sweeps (two per pass, or one per
# pass with the block preconditioner, an odd count ending on a single one),
# and the per-routine path of a preconditioned solve with two inner steps
@pytest.mark.parametrize("precond,inner", [(1, 6), (2, 6), (3, 6), (3, 5), (2, 2), (3, 2)])
@pytest.mark.parametrize("reflective", [False, True])
def test_ppcg_loop(pair_factory, precond, inner, reflective):
    print(f"ppcg loop precond={precond} inner={inner} reflective={reflective}", flush=True)
    p = pair_factory(77, 58, 1, seed=31, precond=precond, reflective=reflective)
    r0, rro = _setup(p)
    eigmin, eigmax = 0.9, 40.0
    ch_a, ch_b, theta = O.calc_ch_coefs(eigmin, eigmax, inner)
    # the switch block (tea_solve.f90:229-250) on both sides
    p.o.ppcg_init(2, precond, 1)
    L.call("tl_set_ppcg_active", 1)
    p.d.scalar("tl_ppcg_init", 2)
    sdr = [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0]
    sd_only = [0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0]
    p.o.ppcg_init_sd(theta, precond)
    p.o.update_halo(sdr, 1, reflective)
    for i in range(inner):
        p.o.ppcg_inner(ch_a[i], ch_b[i], precond)
        p.o.update_halo(sd_only, 1, reflective)
    p.o.ppcg_update_z()
    rro_o = p.o.ppcg_init(3, precond, 1)
    p_only = [0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0]
    p.o.update_halo(p_only, 1, reflective)  # tea_solve.f90:250
    a = np.ascontiguousarray(ch_a)
    b = np.ascontiguousarray(ch_b)
    pa, pb = a.ctypes.data_as(C.POINTER(C.c_double)), b.ctypes.data_as(C.POINTER(C.c_double))
    L.call("tl_ppcg_init_sd", theta)
    L.call("tl_ppcg_inner", pa, pb, 1, inner)
    L.call("tl_ppcg_update_z")
    rro_d = p.d.scalar("tl_ppcg_init", 3)
    assert_scalar(rro_d, rro_o, 1e-10, "rro step 3")
    eps_abs = 1e-10 * math.sqrt(r0)
    # oracle outer loop
    n_end, err = None, 0.0
    rro = rro_o
    for n in range(2, 200):
        pw = p.o.cg_calc_w()
        p.o.ppcg_store_r()
        rro = p.o.ppcg_calc_zrnorm(precond, 1)
        p.o.cg_calc_ur(rro / pw, precond)
        p.o.ppcg_init_sd(theta, precond)
        p.o.update_halo(sdr, 1, reflective)
        for i in range(inner):
            p.o.ppcg_inner(ch_a[i], ch_b[i], precond)
            p.o.update_halo(sd_only, 1, reflective)
        p.o.ppcg_update_z()
        rrn = p.o.ppcg_calc_rrn()
        p.o.cg_calc_p(rrn / rro, precond, 1)
        p.o.update_halo(p_only, 1, reflective)  # tea_solve.f90:384 with fields = P
        rro = rrn
        err = math.sqrt(abs(rrn))
        if err < eps_abs:
            n_end = n
            break
    assert n_end is not None
    st = L.LoopState(n_first=2, n_last=199, eps_abs=eps_abs, rro=rro_d, old_error=0.0)
    inner_iters = C.c_int(0)
    L.call("tl_ppcg_loop", C.byref(st), pa, pb, inner, theta, C.byref(inner_iters))
    assert st.reason == L.LOOP_CONVERGED
    assert st.n_end == n_end, (st.n_end, n_end)
    assert inner_iters.value == inner * (n_end - 1)
    assert_scalar(st.error, err, 1e-6, "error")
    p.cmp("u", tol=1e-10)
    p.cmp("z", tol=1e-8)


@pytest.mark.parametrize("reflective", [False, True])
def test_jacobi_loop(pair_factory, reflective):
    p = pair_factory(45, 50, 1, seed=37, reflective=reflective)
    p.init()
    errs = []
    eps_abs = 1e-3
    n_end = None
    for n in range(1, 500):
        e = math.sqrt(abs(p.o.jacobi_solve()))
        errs.append(e)
        if e < eps_abs:
            n_end = n
            break
    assert n_end is not None
    st = L.LoopState(n_first=1, n_last=499, eps_abs=eps_abs)
    L.call("tl_jacobi_loop", C.byref(st))
    assert st.reason == L.LOOP_CONVERGED and st.n_end == n_end
    # sum |u - un| cancels: each term is ~1e-7 of u, so rounding of u at 1e-16
    # moves it at ~1e-9 relative and the sum at up to ~1e-6
    assert_scalar(st.error, errs[-1], 1e-5, "error")
    p.cmp("u", tol=1e-10)


@pytest.mark.parametrize("precond,reflective", [(1, False), (1, True), (2, True), (3, True)])
def test_cheby_run(pair_factory, precond, reflective):
    """tl_cheby_run against the oracle's cheby_iterate.  With reflective faces
    the fused pass reads u's halo cells as they are (the reference refreshes
    only p's halo between Chebyshev iterations), one iteration per pass; with
    reflective faces and the block preconditioner it runs per routine."""
    print(f"cheby run precond={precond} reflective={reflective}", flush=True)
    p = pair_factory(64, 40, 1, seed=41, precond=precond, reflective=reflective)
    _setup(p)
    theta = 20.0
    ch_a, ch_b, _ = O.calc_ch_coefs(1.0, 39.0, 30)
    p.o.cheby_init(theta, precond)
    L.call("tl_cheby_init", theta)
    for s in range(1, 21):
        p.o.cheby_iterate(ch_a[s - 1], ch_b[s - 1], precond)
    rr_o = p.o.calc_2norm(1)
    a = np.ascontiguousarray(ch_a)
    b = np.ascontiguousarray(ch_b)
    rr = C.c_double()
    L.call("tl_cheby_run", a.ctypes.data_as(C.POINTER(C.c_double)),
           b.ctypes.data_as(C.POINTER(C.c_double)), 1, 20, 1, C.byref(rr))
    assert_scalar(rr.value, rr_o, 1e-8, "cheby residual")
    p.cmp("u", tol=1e-10)


@pytest.mark.parametrize("nx,ny,count", [(64, 40, 20), (64, 42, 19), (1100, 70, 11), (37, 5, 6),
                                         (530, 259, 8)])
def test_cheby_run_block(pair_factory, nx, ny, count):
    """Chebyshev iterations with the block-Jacobi preconditioner fused into
    one pass each (k_cheby_blk: the 4-row Thomas solve on the rows as the
    walk passes them) against the oracle's cheby_iterate with
    tea_block_solve, including meshes whose height is not a multiple of the
    4-row blocks and the fields the last pass of a run stores."""
    print(f"cheby block {nx}x{ny} count={count}", flush=True)
    p = pair_factory(nx, ny, 1, seed=47, precond=L.PREC_JAC_BLOCK)
    _setup(p)
    theta = 20.0
    ch_a, ch_b, _ = O.calc_ch_coefs(1.0, 39.0, count + 2)
    p.o.cheby_init(theta, L.PREC_JAC_BLOCK)
    L.call("tl_cheby_init", theta)
    for s_ in range(1, count + 1):
        p.o.cheby_iterate(ch_a[s_ - 1], ch_b[s_ - 1], L.PREC_JAC_BLOCK)
    rr_o = p.o.calc_2norm(1)
    a = np.ascontiguousarray(ch_a)
    b = np.ascontiguousarray(ch_b)
    rr = C.c_double()
    L.call("tl_cheby_run", a.ctypes.data_as(C.POINTER(C.c_double)),
           b.ctypes.data_as(C.POINTER(C.c_double)), 1, count, 1, C.byref(rr))
    assert_scalar(rr.value, rr_o, 1e-8, "cheby residual")
    for f in ("u", "p", "w", "r", "z"):
        p.cmp(f, tol=1e-10)


@pytest.mark.parametrize("precond,count,nx,ny", [(1, 20, 1100, 70), (2, 21, 1100, 70),
                                                (1, 7, 1100, 1400), (1, 9, 37, 5)])
def test_cheby_pairs_bit_identical(precond, count, nx, ny, monkeypatch):
    """Two Chebyshev iterations per pass (temporal blocking with ghost rows and
    columns) compute what two single passes compute; an odd count ends with a
    single pass.  The same operations on the same operands, so a difference is
    the compiler's choice of fused multiply-adds (last-bit level); a wrong
    ghost cell would show as O(1).  The first case is also checked against the
    oracle."""
    results = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("TL_SWEEP_PAIRS", pairs)
        print(f"cheby pairs={pairs} precond={precond} count={count} {nx}x{ny}", flush=True)
        p = Pair(nx, ny, 1, seed=43, precond=precond)
        try:
            _setup(p)
            theta = 20.0
            ch_a, ch_b, _ = O.calc_ch_coefs(1.0, 39.0, count + 2)
            L.call("tl_cheby_init", theta)
            a = np.ascontiguousarray(ch_a)
            b = np.ascontiguousarray(ch_b)
            rr = C.c_double()
            L.call("tl_cheby_run", a.ctypes.data_as(C.POINTER(C.c_double)),
                   b.ctypes.data_as(C.POINTER(C.c_double)), 1, count, 1, C.byref(rr))
            if pairs == "1" and nx == 1100 and ny == 70 and precond == 1:
                p.o.cheby_init(theta, precond)
                for s_ in range(1, count + 1):
                    p.o.cheby_iterate(ch_a[s_ - 1], ch_b[s_ - 1], precond)
                assert_scalar(rr.value, p.o.calc_2norm(1), 1e-8, "cheby residual")
                p.cmp("u", tol=1e-10)
            results.append((rr.value, [p.d.download(f) for f in ("u", "p", "w", "r")]))
        finally:
            p.close()
    (r0, f0), (r1, f1) = results
    assert abs(r0 - r1) <= 1e-12 * abs(r0), (r0, r1)
    for name, x, y in zip(("u", "p", "w", "r"), f0, f1):
        d = np.max(np.abs(x - y)) / np.max(np.abs(x))
        print(f"  pairs vs single passes: {name} max rel diff {d:.1e}", flush=True)
        assert d <= 1e-12, (name, d)


@pytest.mark.parametrize("precond,reflective,inner,ny", [(1, False, 6, 70), (2, False, 7, 70),
                                                          (1, True, 5, 70), (1, False, 4, 1400)])
def test_ppcg_sweep_pairs_bit_identical(precond, reflective, inner, ny, monkeypatch):
    """Two inner sweeps per pass (temporal blocking with ghost rows and
    columns) compute what two single passes compute: the same operations on
    the same operands, so any difference is the compiler's choice of fused
    multiply-adds (last-bit level); a wrong ghost cell would show as O(1)."""
    results = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("TL_SWEEP_PAIRS", pairs)
        p = Pair(1100, ny, 1, seed=41, precond=precond, reflective=reflective)
        try:
            r0, rro = _setup(p)
            ch_a, ch_b, theta = O.calc_ch_coefs(0.9, 40.0, inner)
            a = np.ascontiguousarray(ch_a)
            b = np.ascontiguousarray(ch_b)
            pa, pb = a.ctypes.data_as(C.POINTER(C.c_double)), b.ctypes.data_as(C.POINTER(C.c_double))
            L.call("tl_set_ppcg_active", 1)
            p.d.scalar("tl_ppcg_init", 2)
            L.call("tl_ppcg_init_sd", theta)
            L.call("tl_ppcg_inner", pa, pb, 1, inner)
            L.call("tl_ppcg_update_z")
            rro_d = p.d.scalar("tl_ppcg_init", 3)
            st = L.LoopState(n_first=2, n_last=40, eps_abs=1e-12 * math.sqrt(r0), rro=rro_d,
                             old_error=0.0)
            it = C.c_int(0)
            L.call("tl_ppcg_loop", C.byref(st), pa, pb, inner, theta, C.byref(it))
            results.append((st.n_end, st.error, [p.d.download(f) for f in ("u", "z", "r", "p")]))
        finally:
            p.close()
    (n0, e0, f0), (n1, e1, f1) = results
    assert n0 == n1
    assert abs(e0 - e1) <= 1e-12 * abs(e0), (e0, e1)
    for name, a, b in zip(("u", "z", "r", "p"), f0, f1):
        d = np.max(np.abs(a - b)) / np.max(np.abs(a))
        print(f"  pairs vs single sweeps: {name} max rel diff {d:.1e}", flush=True)
        assert d <= 1e-12, (name, d)


def test_profile_sampling():
    """tl_profile_enable(N) times every N-th launch of a kernel class: the
    same CG loop queues the same launches, of which ceil(n / N) are timed."""
    counts = []
    for period in (1, 4):
        print(f"profile sampling period {period}", flush=True)
        p = Pair(200, 150, 1, seed=47)
        try:
            _setup(p)
            r0 = p.d.scalar("tl_calc_2norm", 1)
            rro = p.d.scalar("tl_cg_init")
            L.call("tl_profile_reset")
            L.call("tl_profile_enable", period)
            st = L.LoopState(n_first=1, n_last=300, eps_abs=1e-10 * math.sqrt(r0), rro=rro,
                             old_error=r0)
            L.call("tl_cg_loop", C.byref(st), None, None)
            L.call("tl_profile_enable", 0)
            n, ms, idle = C.c_long(), C.c_double(), C.c_long()
            L.call("tl_profile_get", 0, C.byref(n), C.byref(ms))
            L.call("tl_profile_idle", 0, C.byref(idle))
            counts.append((n.value + idle.value, ms.value, st.n_end, n.value))
        finally:
            p.close()
    (n1, ms1, e1, a1), (n4, ms4, e4, a4) = counts
    assert e1 == e4 and n1 > 8
    # launches timed, with or without work
    assert n4 == (n1 + 3) // 4, (n1, n4)
    # every launch up to the exit did work and is kept
    assert a1 >= e1 and a4 >= (e1 + 3) // 4 - 1, (a1, a4, e1)
    assert ms1 > 0.0 and ms4 > 0.0


@pytest.mark.parametrize("nx,ny,reflective", [(45, 50, False), (1100, 70, False),
                                               (300, 1400, False), (45, 50, True),
                                               (1100, 70, True)])
def test_jacobi_pairs_match_singles(nx, ny, reflective, monkeypatch):
    """Two Jacobi iterations per pass (with both exit tests per pass, the
    first iteration's u kept for an exit there) give what one iteration per
    pass gives: the same exit iteration and reason, u to the last bits, for
    exits at either iteration of a pair and for odd and even limits; with
    reflective faces both read zero outside the mesh."""
    cases = [(1e-3, 499), (3e-4, 499), (1e-30, 7), (1e-30, 8), (1e-30, 1)]
    got = {}
    for pairs in ("0", "1"):
        monkeypatch.setenv("TL_SWEEP_PAIRS", pairs)
        for eps_abs, n_last in cases:
            print(f"jacobi pairs={pairs} {nx}x{ny} reflective={reflective} eps {eps_abs} "
                  f"n_last {n_last}", flush=True)
            p = Pair(nx, ny, 1, seed=53, reflective=reflective)
            try:
                p.init()
                st = L.LoopState(n_first=1, n_last=n_last, eps_abs=eps_abs)
                L.call("tl_jacobi_loop", C.byref(st))
                got[(pairs, eps_abs, n_last)] = (st.n_end, st.reason, st.error, p.d.download("u"))
            finally:
                p.close()
    parities = set()
    for eps_abs, n_last in cases:
        (n0, r0, e0, u0), (n1, r1, e1, u1) = got[("0", eps_abs, n_last)], got[("1", eps_abs, n_last)]
        assert (n0, r0) == (n1, r1), (eps_abs, n_last, n0, r0, n1, r1)
        assert abs(e0 - e1) <= 1e-9 * abs(e0), (e0, e1)
        d = np.max(np.abs(u0 - u1)) / np.max(np.abs(u0))
        print(f"  eps {eps_abs} n_last {n_last}: exit {n0} reason {r0}, u max rel diff {d:.1e}",
              flush=True)
        assert d <= 1e-12, d
        parities.add(n0 % 2)
    assert parities == {0, 1}

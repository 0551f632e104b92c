#!/usr/bin/env python3
"""Benchmark: seconds per timestep and HBM GB/s of the implicit solve at 4096^2.

Workload (one GPU): the seeded 4096x4096 deck (3 rectangles + 2 circles on
[0,10]^2, numpy PCG64 seed 2), FP64 CG, tl_eps 1e-15, tl_max_iters 10000,
dt 0.004, a field summary every step.  A "step" is one pass of the reference's
diffuse() loop body: timestep + tea_leaf() solve + field_summary, run by the
native host with every field resident in HBM.

Prints one JSON line.  `value` is seconds per timestep (lower is better);
`precision_comparison` times FP64 CG / PPCG against the FP32-inner,
FP64-refined forms at a relative residual of 1e-10 on the inverse-density
deck;
`roofline` uses the algorithmic bytes of the CG calc_w kernel (48 B per cell:
read r, p, Kx, Ky; write p, w) over its HIP-event-timed launches (every 16th
launch; launches queued past a loop's exit, which return at once, are left
out) and takes `traffic` from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE summary; `solvers` times Chebyshev, PPCG and Jacobi steps with
the same per-kernel rates, on algorithmic and on counter-measured bytes;
`cpu_baseline` times the reference Fortran/OpenMP build (oracle/_ref) on the
host cores on a bounded sample of the same deck.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# every EVENT_PERIOD-th launch of each kernel class in the timed steps records
# its own start and end in HIP events (timing every launch slows a CG step
# by a few per cent)
EVENT_PERIOD = 16
CG_W_BYTES_PER_CELL = 48.0      # fused calc_w: r, p, Kx, Ky in; p, w out
CG_UR_BYTES_PER_CELL = 48.0     # calc_ur: u, p, r, w in; u, r out
CG_ITER_BYTES_PER_CELL = 96.0
# algorithmic bytes per cell at the fusion level the reference's dataflow
# permits (each pass reads each input once and writes each output once, no
# reuse across iterations): Chebyshev iteration, PPCG outer iteration without
# its sweeps, PPCG sweep, and the fixed work of a step (init, residual,
# cg_init, finalise, summary)
CHEBY_ITER_BYTES_PER_CELL = 56.0
JACOBI_ITER_BYTES_PER_CELL = 40.0
PPCG_OUTER_BYTES_PER_CELL = 112.0
PPCG_SWEEP_BYTES_PER_CELL = 64.0
STEP_FIXED_BYTES_PER_CELL = 200.0


def model_bytes_per_cell(solver, rows):
    """Algorithmic bytes per cell of the steps in `rows` (TeaHost.last()):
    the CG iterations before the switch, then Chebyshev iterations or PPCG
    outer iterations and sweeps.  Temporal blocking (two sweeps or iterations
    per pass) moves fewer bytes than this, so the resulting fraction of the
    roofline can exceed 1."""
    total = 0.0
    for r in rows:
        its = r["iterations"]
        cg = its if solver == "cg" or r["switch_n"] <= 0 else r["switch_n"] - 1
        rest = its - cg
        total += STEP_FIXED_BYTES_PER_CELL + CG_ITER_BYTES_PER_CELL * cg
        if solver == "chebyshev":
            total += CHEBY_ITER_BYTES_PER_CELL * rest
        elif solver == "ppcg":
            total += PPCG_OUTER_BYTES_PER_CELL * rest + PPCG_SWEEP_BYTES_PER_CELL * r["ppcg_inner"]
    return total


def model_bytes_jacobi(rows):
    """Algorithmic bytes per cell of Jacobi steps: 40 per iteration (read u
    with its neighbours, u0, Kx, Ky; write u) plus the step's fixed work."""
    return sum(STEP_FIXED_BYTES_PER_CELL + JACOBI_ITER_BYTES_PER_CELL * r["iterations"]
               for r in rows)


def measured_traffic(kernel, cells, profile="cg_fp64_{n}_hbm_traffic.md"):
    """HBM bytes per launch of a kernel on a cells x cells mesh from the
    committed counter summary `profile` ({n} stands for cells; rocprofv3
    FETCH_SIZE / WRITE_SIZE, copy-kernel calibrated), or None where no
    summary covers that size."""
    path = os.path.join(ROOT, "profiles", profile.format(n=cells))
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch %s: (\d+)" % re.escape(kernel), text)
    return float(m.group(1)) if m else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_run(cells, solver, steps, warmup, workdir, **deck_kw):
    from tealeaf_ref_amd import _lib
    from tealeaf_ref_amd.deckgen import bench_deck
    from tealeaf_ref_amd.host import TeaHost

    deck = bench_deck(cells, solver=solver, eps=1e-15, max_iters=10000,
                      end_step=steps + warmup, summary_frequency=1, **deck_kw)
    deck_path = os.path.join(workdir, "tea.in")
    with open(deck_path, "w") as f:
        f.write(deck)
    t0 = time.time()
    h = TeaHost(deck_path, os.path.join(workdir, "tea.out"))
    log(f"[bench] set-up {cells}^2 ({solver}) in {time.time() - t0:.1f} s")
    for i in range(warmup):
        h.step()
        log(f"[bench] warmup step {i + 1}: {h.last()['iterations']:.0f} iterations")
    _lib.call("tl_synchronize")
    _lib.call("tl_profile_reset")
    _lib.call("tl_profile_enable", EVENT_PERIOD)
    iters = []
    rows = []
    t0 = time.perf_counter()
    for i in range(steps):
        h.step()
        rows.append(h.last())
        it = rows[-1]["iterations"]
        iters.append(it)
        log(f"[bench] step {i + 1}: {it:.0f} iterations")
    _lib.call("tl_synchronize")
    elapsed = time.perf_counter() - t0
    import ctypes as C
    prof = {"rows": rows}
    for cls in range(8):
        n = C.c_long()
        ms = C.c_double()
        idle = C.c_long()
        _lib.call("tl_profile_get", cls, C.byref(n), C.byref(ms))
        _lib.call("tl_profile_idle", cls, C.byref(idle))
        prof[cls] = (n.value, ms.value, idle.value)
    _lib.call("tl_profile_enable", 0)
    h.close()
    return elapsed, iters, prof


# algorithmic bytes per cell of one launch of the kernel classes of
# tl_profile_get; a PPCG sweep pass, Chebyshev pass or Jacobi pass does two
# sweeps or iterations for the bytes of one (plus, for Jacobi, the store of
# the first iteration's u)
# the PPCG sweep class: one first pass (48 B: r, Kx, Ky in; sd, rtemp, utemp
# out), three middle passes (64 B) and one last pass (64 B: z the only
# output) per outer iteration of 10 inner steps
PPCG_PASS_BYTES_PER_CELL = (48.0 + 3 * 64.0 + 64.0) / 5
CLASS_BYTES = {0: ("cg calc_w", 48.0), 1: ("cg calc_ur", 48.0),
               2: ("ppcg inner sweep pass", PPCG_PASS_BYTES_PER_CELL), 3: ("chebyshev pass", 56.0),
               4: ("jacobi pass", 48.0), 5: ("ppcg calc_w", 48.0),
               6: ("ppcg calc_ur + store_r", 56.0)}
# the kernel of each class and the committed FETCH_SIZE / WRITE_SIZE summary
# that holds its measured HBM bytes per launch ({n}: the mesh size)
CLASS_TRAFFIC = {0: ("k_loop_w", "cg_fp64_{n}_hbm_traffic.md"),
                 1: ("k_cg_ur", "cg_fp64_{n}_hbm_traffic.md"),
                 2: ("k_pp_sweep2", "ppcg_fp64_{n}_hbm_traffic.md"),
                 3: ("k_cheby_pair", "cheby_fp64_{n}_hbm_traffic.md"),
                 4: ("k_jl_pair", "jacobi_fp64_{n}_hbm_traffic.md"),
                 5: ("k_loop_w", "ppcg_fp64_{n}_hbm_traffic.md"),
                 6: ("k_pp_ur", "ppcg_fp64_{n}_hbm_traffic.md")}


def class_rates(prof, cells, classes=CLASS_BYTES, traffic=CLASS_TRAFFIC):
    """Per-class launch time and HBM rates from HIP-event samples `prof`
    (class -> (timed launches, ms, idle launches)): algorithmic bytes, and
    the counter-measured bytes where a committed summary covers this size."""
    out = {}
    for cls, (name, bpc) in classes.items():
        n, ms, idle = prof.get(cls, (0, 0.0, 0))
        if not n:
            continue
        k = {"timed_launches": n, "idle_launches": idle, "avg_launch_ms": ms / n,
             "achieved_gbs": bpc * cells * cells / (ms / n * 1e-3) / 1e9, "bytes_per_cell": bpc}
        k["frac"] = k["achieved_gbs"] / PEAK_HBM_GBS
        kname, prof_file = traffic[cls]
        t = measured_traffic(kname, cells, prof_file)
        if t:
            k["measured_bytes_per_launch"] = t
            k["measured_gbs"] = t / (ms / n * 1e-3) / 1e9
            k["measured_frac"] = k["measured_gbs"] / PEAK_HBM_GBS
            k["traffic_source"] = "profiles/" + prof_file.format(n=cells)
        out[name] = k
    return out


def work_launches(solver, rows, inner=0, max_iters=10000):
    """Launches per kernel class that did work in the steps of `rows`
    (TeaHost.last()), from the iteration counts: CG iterations before the
    switch (two passes each), then Chebyshev iterations two per pass, PPCG
    outer iterations (a w pass, an update pass and ceil(inner/2) sweep passes
    each) or Jacobi iterations two per pass.  The printed count is n-1 at an
    exit at iteration n (tea_solve.f90:448), max_iters at the cap."""
    n = dict.fromkeys(range(7), 0)
    for r in rows:
        its = int(r["iterations"])
        end = its if its >= max_iters else its + 1  # last executed iteration
        sw = int(r["switch_n"])
        cg = end if (solver in ("cg", "jacobi") or sw <= 0) else sw - 1
        if solver == "jacobi":
            n[4] += (end + 1) // 2
            continue
        n[0] += cg
        n[1] += cg
        rest = end - cg
        if solver == "chebyshev":
            n[3] += max(0, rest - 1) // 2   # the first Chebyshev step runs per routine
        elif solver == "ppcg":
            n[5] += rest
            n[6] += rest
            n[2] += rest * ((inner + 1) // 2)
    return n


def solver_runs(cells, steps, warmup):
    """Chebyshev (20 CG presteps), PPCG (10 inner steps) and Jacobi on the
    same 4096^2 deck at tl_eps 1e-15, tl_max_iters 10000 (Chebyshev and Jacobi
    run to the cap): seconds per step, iterations and the per-kernel-class HBM
    rates (HIP events)."""
    out = {}
    for solver, kw in (("chebyshev", {"presteps": 20}), ("ppcg", {"inner_steps": 10}),
                       ("jacobi", {})):
        with tempfile.TemporaryDirectory() as wd:
            elapsed, iters, prof = gpu_run(cells, solver, steps, warmup, wd, **kw)
        rec = {"s_per_step": elapsed / steps, "iterations_per_step": sum(iters) / steps}
        if solver == "ppcg":
            rec["sweeps_per_step"] = sum(r["ppcg_inner"] for r in prof["rows"]) / steps
        model = (model_bytes_jacobi(prof["rows"]) if solver == "jacobi"
                 else model_bytes_per_cell(solver, prof["rows"]))
        rec["model_bytes_per_cell_step"] = model / steps
        rec["model_gbs"] = model * cells * cells / elapsed / 1e9
        rec["model_frac"] = rec["model_gbs"] / PEAK_HBM_GBS
        kern = class_rates(prof, cells)
        rec["kernels"] = kern
        # the step's rate in counter-measured bytes: launches that did work x
        # the measured bytes per launch of their class, over the step time;
        # `covered` is the share of the step those launches take
        counts = work_launches(solver, prof["rows"], kw.get("inner_steps", 0))
        mbytes = covered_ms = 0.0
        for cls, (name, _) in CLASS_BYTES.items():
            k = kern.get(name)
            if counts[cls] and k and "measured_bytes_per_launch" in k:
                mbytes += counts[cls] * k["measured_bytes_per_launch"]
                covered_ms += counts[cls] * k["avg_launch_ms"]
        if mbytes:
            rec["step_measured_gbs"] = mbytes / elapsed / 1e9
            rec["step_measured_frac"] = rec["step_measured_gbs"] / PEAK_HBM_GBS
            rec["step_time_covered"] = covered_ms * 1e-3 / elapsed
            rec["work_launches"] = {CLASS_BYTES[c][0]: v for c, v in counts.items() if v}
        log(f"[bench] {cells}^2 {solver}: {rec}")
        out[solver] = rec
    return out


def stream_bandwidth(n=1 << 28, reps=10):
    """Achievable HBM bandwidth on this GPU, STREAM style, from the library's
    hand-written copy (16 B per element) and triad (24 B per element) kernels
    on three FP64 arrays of n elements (2 GiB each, far beyond the 256 MB
    Infinity Cache): the best over plain and streaming accesses and grid
    sizes (tl_hbm_stream)."""
    import ctypes as C
    from tealeaf_ref_amd import _lib
    d = _lib.Device(64, 64)
    try:
        cp, tr = C.c_double(), C.c_double()
        _lib.call("tl_hbm_stream", n, reps, C.byref(cp), C.byref(tr))
    finally:
        d.close()
    out = {"elements": n, "dtype": "f64", "kernels": "hand-written HIP (tl_hbm_stream)",
           "copy_gbs": cp.value, "triad_gbs": tr.value,
           "copy_frac": cp.value / PEAK_HBM_GBS, "triad_frac": tr.value / PEAK_HBM_GBS}
    log(f"[bench] stream bandwidth: {out}")
    return out


def small_run(steps=10):
    """The reference's CPU-sized case: 512x512, 3 seeded rectangles (seed 1),
    FP64 CG, tl_eps 1e-15, 10 steps; seconds per step and per iteration."""
    from tealeaf_ref_amd.deckgen import make_deck
    deck = make_deck(512, solver="cg", seed=1, eps=1e-15, max_iters=5000, end_step=steps + 1)
    with tempfile.TemporaryDirectory() as wd:
        elapsed, rows, _ = run_deck_timed(deck, steps, 1, wd)
    its = sum(r["iterations"] + 1 for r in rows)
    rec = {"workload": "512x512 FP64 CG, seeded deck (3 rectangles, seed 1), tl_eps 1e-15",
           "s_per_step": elapsed / steps, "iterations_per_step": its / steps,
           "us_per_iteration": elapsed / its * 1e6}
    log(f"[bench] small: {rec}")
    return rec


def run_deck_timed(deck, steps, warmup, workdir, want_u=False, prof_out=None, residual=None):
    """Run `warmup` + `steps` timesteps of a deck through the native host;
    returns the timed seconds, the per-step results of the timed steps and
    the final u (interior) when asked.  With a dict `prof_out`, every
    EVENT_PERIOD-th launch of each kernel class is timed and prof_out[class]
    = (timed launches, their milliseconds, idle launches).  With residual =
    (rx, ry), also returns host_true_residual of the last step."""
    import ctypes as C
    import numpy as np
    from tealeaf_ref_amd import _lib
    from tealeaf_ref_amd.host import TeaHost
    deck_path = os.path.join(workdir, "tea.in")
    with open(deck_path, "w") as f:
        f.write(deck)
    h = TeaHost(deck_path, os.path.join(workdir, "tea.out"))
    for _ in range(warmup):
        h.step()
    _lib.call("tl_synchronize")
    if prof_out is not None:
        _lib.call("tl_profile_reset")
        _lib.call("tl_profile_enable", EVENT_PERIOD)
    rows = []
    t0 = time.perf_counter()
    for _ in range(steps):
        h.step()
        rows.append(h.last())
    _lib.call("tl_synchronize")
    elapsed = time.perf_counter() - t0
    if prof_out is not None:
        for cls in range(8):
            n, ms, idle = C.c_long(), C.c_double(), C.c_long()
            _lib.call("tl_profile_get", cls, C.byref(n), C.byref(ms))
            _lib.call("tl_profile_idle", cls, C.byref(idle))
            prof_out[cls] = (n.value, ms.value, idle.value)
        _lib.call("tl_profile_enable", 0)
    u = None
    if want_u:
        m = re.search(r"x_cells=(\d+)", deck)
        n = int(m.group(1))
        u = np.zeros((n + 2, n + 2))
        _lib.call("tl_download", _lib.FIELD["u"], u.ctypes.data_as(C.POINTER(C.c_double)), 1)
        u = u[1:-1, 1:-1].copy()
    if residual is not None:
        m = re.search(r"x_cells=(\d+)", deck)
        tres = host_true_residual(int(m.group(1)), *residual)
        h.close()
        return elapsed, rows, u, tres
    h.close()
    return elapsed, rows, u


# kernel classes of the FP32 inner PPCG of a mixed-precision solve: name and
# algorithmic bytes per cell (FP32 fields, half the FP64 figures), and the
# kernel's line in profiles/ppcg_mixed_{n}_hbm_traffic.md (the FP64
# refinements, a few per step, are too rare for sampled events)
MIXED_PPCG_BYTES = {2: ("fp32 ppcg inner sweep pass", PPCG_PASS_BYTES_PER_CELL / 2),
                    5: ("fp32 ppcg calc_w", 24.0), 6: ("fp32 ppcg calc_ur + store_r", 28.0)}
MIXED_PPCG_TRAFFIC = {2: ("k_pp_sweep2<float>", "ppcg_mixed_{n}_hbm_traffic.md"),
                      5: ("k_loop_w<float>", "ppcg_mixed_{n}_hbm_traffic.md"),
                      6: ("k_pp_ur<float>", "ppcg_mixed_{n}_hbm_traffic.md")}


MIXED_CG_BYTES = {0: ("fp32 cg calc_w", 24.0), 1: ("fp32 cg calc_ur", 24.0)}
MIXED_CG_TRAFFIC = {0: ("k_loop_w<float>", "ppcg_mixed_{n}_hbm_traffic.md"),
                    1: ("k_cg_ur<float>", "ppcg_mixed_{n}_hbm_traffic.md")}


def mixed_kernel_rates(prof, cells):
    """Per-class launch time and HBM rates of a mixed-precision PPCG run."""
    return class_rates(prof, cells, MIXED_PPCG_BYTES, MIXED_PPCG_TRAFFIC)


def host_true_residual(cells, rx, ry):
    """||u0 - A u|| / ||u0 - A u0|| of the step the device holds (u0 its
    right-hand side, u its solution), evaluated on the host from the
    downloaded fields: A in difference form in FP64 (rounding relative to the
    fluxes, not to the diagonal), the same evaluation for every form."""
    import numpy as np
    from tealeaf_ref_amd import _lib
    u = _lib.download_field("u", cells, cells)
    u0 = _lib.download_field("u0", cells, cells)[1:-1, 1:-1].copy()
    kx = _lib.download_field("kx", cells, cells)
    ky = _lib.download_field("ky", cells, cells)
    kl, kr = kx[1:-1, 1:-1], kx[1:-1, 2:]   # faces left / right of each cell
    kd, ku = ky[1:-1, 1:-1], ky[2:, 1:-1]   # below / above

    def resid_norm(x):
        c = x[1:-1, 1:-1]
        a = c + rx * (kl * (c - x[1:-1, :-2]) + kr * (c - x[1:-1, 2:])) \
            + ry * (kd * (c - x[:-2, 1:-1]) + ku * (c - x[2:, 1:-1]))
        return float(np.sqrt(np.sum((u0 - a) ** 2)))

    res = resid_norm(u)
    x0 = np.zeros_like(u)
    x0[1:-1, 1:-1] = u0
    res0 = resid_norm(x0)
    return res / res0 if res0 > 0.0 else 0.0


# tl_eps of the FP64 forms in the precision comparison: the largest of 1e-10,
# 1e-11, ... at which their exit (the reference's recurrence test) leaves a
# true relative residual <= 1e-10 on the first step (measured on this deck at
# 4096^2: CG's recurrence residual tracks the true one, 9.99e-11 at 1e-10;
# PPCG's exit quantity (r - r_store).z does not: 1.5e-8 at 1e-10, 1.5e-10 at
# 1e-12, 1.5e-11 at 1e-13); the mixed forms stop on the true residual itself
PRECISION_EPS = {"fp64_cg": 1e-10, "fp64_ppcg": 1e-13, "mixed_cg": 1e-10, "mixed_ppcg": 1e-10}

# the FP64 forms' iteration cap in the comparison: 3x the CG count at 8192^2
PRECISION_MAX_ITERS = 40000


def precision_comparison(cells, inner_steps=10, target=1e-10, later_step=False):
    """FP64 CG, FP64 PPCG and the FP32-inner / FP64-refined CG and PPCG on the
    inverse-density deck, each to a true relative residual of `target`, timed
    on the deck's first step (every form starts from the same u0).  The FP64
    forms stop on the reference's recurrence test (tea_solve.f90:410), at the
    tl_eps of PRECISION_EPS that brings their true residual to the target;
    the mixed forms stop on the true residual.  Every form's result is then
    measured the same way (host_true_residual).  Errors of u and of the
    summary are relative to the FP64 CG result.

    With later_step, FP64 CG and mixed PPCG also run the second step (after
    an untimed first): there the starting residual u0 - A u0 is smaller than
    on the first step while the rounding of the reference's FP64 product A p
    (scaled by |A| |u|) is not, so its true residual can level off above
    the target; the mixed form refines on the difference-form residual,
    whose rounding scales with the fluxes.  Each record reports what the form
    reached."""
    from tealeaf_ref_amd.deckgen import bench_deck
    import numpy as np
    runs = {}
    plan = [("fp64_cg", "cg", ()), ("fp64_ppcg", "ppcg", ()),
            ("mixed_cg", "cg", ("tl_mixed_precision",)),
            ("mixed_ppcg", "ppcg", ("tl_mixed_precision",))]
    ref_u = ref_sum = None
    dx = 10.0 / cells
    rx = 0.004 / (dx * dx)

    def deck_of(name, solver, extra, end_step, cells=cells):
        kw = dict(coefficient="inverse_density", max_iters=PRECISION_MAX_ITERS,
                  end_step=end_step, summary_frequency=1, extra=extra, eps=PRECISION_EPS[name])
        if solver == "ppcg":
            kw["inner_steps"] = inner_steps
        return bench_deck(cells, solver=solver, **kw)

    for name, solver, extra in plan:
        # one untimed step of the same form on a small mesh first, so that no
        # form pays the library's first-launch costs
        with tempfile.TemporaryDirectory() as wd:
            run_deck_timed(deck_of(name, solver, extra, 1, cells=256), 1, 0, wd)
        log(f"[bench] {cells}^2 {name}: start")
        prof = {}
        with tempfile.TemporaryDirectory() as wd:
            elapsed, rows, u, tres = run_deck_timed(deck_of(name, solver, extra, 1), 1, 0, wd,
                                                    want_u=True, prof_out=prof,
                                                    residual=(rx, rx))
        steps = 1
        last = rows[-1]
        rec = {"tl_eps": PRECISION_EPS[name], "s_per_step": elapsed / steps,
               "iterations_per_step": sum(r["iterations"] for r in rows) / steps,
               "true_rel_residual": tres, "meets_target": tres <= target}
        if solver == "ppcg":
            rec["sweeps_per_step"] = sum(r["ppcg_inner"] for r in rows) / steps
        if extra:
            rec["refinements_per_step"] = sum(r["refinements"] for r in rows) / steps
            rec["solver_true_rel_residual"] = max(r["conduction_error"] for r in rows)
        else:
            rec["recurrence_rel_residual"] = max(r["conduction_error"] for r in rows)
            # HBM rate in the algorithmic bytes of the reference's dataflow
            rec["model_gbs"] = (model_bytes_per_cell(solver, rows) * cells * cells / elapsed
                                / 1e9)
        if name == "mixed_ppcg":
            rec["kernels"] = mixed_kernel_rates(prof, cells)
        elif name == "mixed_cg":
            rec["kernels"] = class_rates(prof, cells, MIXED_CG_BYTES, MIXED_CG_TRAFFIC)
        elif not extra:
            rec["kernels"] = class_rates(prof, cells)
        summ = [last["volume"], last["mass"], last["energy"], last["temperature"]]
        if ref_u is None:
            ref_u, ref_sum = u, summ
        else:
            rec["u_rel_err_vs_fp64_cg"] = float(np.max(np.abs(u - ref_u)) / np.max(np.abs(ref_u)))
            rec["summary_rel_err_vs_fp64_cg"] = max(
                abs(a - b) / abs(b) for a, b in zip(summ, ref_sum) if b != 0.0)
        log(f"[bench] {cells}^2 {name}: {rec}")
        runs[name] = rec
    out = {"workload": f"{cells}x{cells}, seeded deck (3 rectangles + 2 circles, seed 2), "
                       f"tl_coefficient_inverse_density, true relative residual {target:g}, "
                       f"PPCG {inner_steps} inner steps, first step timed",
           "runs": runs,
           "all_meet_target": all(r["meets_target"] for r in runs.values())}
    if later_step:
        out["second_step"] = {}
        for name, solver, extra in (plan[0], plan[3]):
            log(f"[bench] {cells}^2 {name} second step: start")
            with tempfile.TemporaryDirectory() as wd:
                elapsed, rows, _, tres = run_deck_timed(deck_of(name, solver, extra, 2), 1, 1,
                                                        wd, residual=(rx, rx))
            r = rows[-1]
            rec = {"tl_eps": PRECISION_EPS[name], "s_per_step": elapsed,
                   "iterations": r["iterations"], "true_rel_residual": tres,
                   "meets_target": tres <= target, "solver_exit_residual": r["conduction_error"]}
            if extra:
                rec["refinements"] = r["refinements"]
                rec["sweeps"] = r["ppcg_inner"]
            log(f"[bench] {cells}^2 {name} second step: {rec}")
            out["second_step"][name] = rec
    t = {k: v["s_per_step"] for k, v in runs.items()}
    out["speedup_mixed_ppcg_vs_fp64_cg"] = t["fp64_cg"] / t["mixed_ppcg"]
    out["speedup_mixed_cg_vs_fp64_cg"] = t["fp64_cg"] / t["mixed_cg"]
    out["speedup_mixed_ppcg_vs_fp64_ppcg"] = t["fp64_ppcg"] / t["mixed_ppcg"]
    out["speedup_best_mixed_vs_best_fp64"] = min(t["fp64_cg"], t["fp64_ppcg"]) / min(
        t["mixed_cg"], t["mixed_ppcg"])
    return out


def deck_fields(cells, coefficient="density"):
    """Density, energy (interior) and rx, ry of the seeded deck, generated by
    the native host (generate_chunk) on the GPU and downloaded."""
    import numpy as np
    from tealeaf_ref_amd import _lib
    from tealeaf_ref_amd.deckgen import bench_deck
    from tealeaf_ref_amd.host import TeaHost
    deck = bench_deck(cells, solver="cg", eps=1e-15, max_iters=10000, end_step=1,
                      coefficient=coefficient)
    with tempfile.TemporaryDirectory() as wd:
        path = os.path.join(wd, "tea.in")
        with open(path, "w") as f:
            f.write(deck)
        h = TeaHost(path, os.path.join(wd, "tea.out"))
        rho = _lib.download_field("density", cells, cells)[1:-1, 1:-1].copy()
        e1 = _lib.download_field("energy1", cells, cells)[1:-1, 1:-1].copy()
        h.close()
    dx = 10.0 / cells
    dt = 0.004
    return rho, e1, dt / (dx * dx), dt / (dx * dx)


def adjoint_bench(cells, steps, every, n_fd, h, precision, mixed_eps=1e-12, seed=5,
                  solver="cg", save=None, compare=None, fd_eps=1e-12):
    """The adjoint case: gradient of J = sum over a seeded 256x256 window of
    (u_N - u_target)^2 with respect to per-cell log-density, u_target from a
    forward on a log-normally perturbed density (sigma 0.1), computed once
    (by the first precision) and shared; timed forward and backward
    (checkpoint every `every` steps) with CG or PPCG solves (PPCG: 10 inner
    sweeps, bounds from 30 CG iterations of each solve).

    Finite differences: central differences with step h on n_fd seeded cells
    inside the window.  The perturbed forwards run the refined solver (FP32
    inner CG, FP64 refinement to a true relative residual of fd_eps), so the
    difference quotient resolves the gradient rather than the forward
    solver's own stopping error; the same quotients are compared with the
    gradient of every precision.

    A case too long for one run is split in two: `save` writes the target on
    the window, the gradient on the window and on every 8th row and column,
    and its norm (.npz); a later run of the other precision with `compare`
    takes the target from that file and reports its gradient's difference
    from the saved one (relative L2 on the window and on the subsample)."""
    import ctypes as C
    import numpy as np
    from tealeaf_ref_amd import _lib
    from tealeaf_ref_amd.adjoint import TeaLeafSteps
    rho, e0, rx, ry = deck_fields(cells)
    rng = np.random.default_rng(seed)
    theta = np.log(rho)
    theta_p = theta + 0.1 * rng.standard_normal(theta.shape)
    wy, wx = (int(v) for v in rng.integers(0, cells - 256, 2))
    win = (slice(wy, wy + 256), slice(wx, wx + 256))
    precs = ["fp64", "mixed"] if precision == "both" else [precision]
    out = {"workload": f"{cells}x{cells}, {steps} steps, seeded deck, per-cell density "
                       f"rho*exp(0.1 xi), window 256x256 at ({wx},{wy}), checkpoint every {every}, "
                       f"{solver.upper()} solves"}
    grads = {}
    target = None
    saved = None
    if compare:
        saved = dict(np.load(compare))
        target = np.zeros((cells, cells))
        target[win] = saved["target_window"]
        out["compared_with"] = {"file": os.path.basename(compare), "J": float(saved["J"]),
                                "precision": str(saved["precision"]),
                                "solver": str(saved["solver"])}
    for prec in precs:
        # FP64: the reference's 1e-15 for CG; PPCG's flexible-CG residual
        # (r - r_store).z stalls at the rounding floor near there, so 1e-14
        fp64_eps = 1e-15 if solver == "cg" else 1e-14
        p = TeaLeafSteps(e0, rx, ry, steps, coefficient=_lib.CONDUCTIVITY,
                         eps=fp64_eps if prec == "fp64" else mixed_eps, checkpoint_every=every,
                         precision=prec, solver=solver,
                         max_iters=20000 if solver == "cg" else 5000)
        if cells >= 4096:
            p.progress = lambda msg: log(f"[bench] adjoint {prec} {cells}^2: {msg}")
        if target is None:
            target = p.forward(theta_p)
        t0 = time.perf_counter()
        u = p.forward(theta)
        t_f = time.perf_counter() - t0
        its_f, sw_f = p.iterations, p.sweeps
        stops_f = (p.maxiter_stops, p.floor_stops)
        seed_g = np.zeros_like(u)
        seed_g[win] = 2.0 * (u[win] - target[win])
        J = float(((u[win] - target[win]) ** 2).sum())
        t0 = time.perf_counter()
        g = p.backward(seed_g)
        t_b = time.perf_counter() - t0
        its_b = p.iterations
        nbytes = C.c_double()
        _lib.call("tl_device_bytes", C.byref(nbytes))
        rec = {"J": J, "forward_s_per_step": t_f / steps,
               "forward_backward_s_per_step": (t_f + t_b) / steps,
               "forward_iterations_per_step": its_f / steps,
               "backward_iterations_per_step": its_b / steps,
               "forward_sweeps_per_step": sw_f / steps,
               "backward_sweeps_per_step": p.sweeps / steps,
               "eps": p.eps, "solves": 2 * steps,
               "unconverged_solves": {"max_iters": stops_f[0] + p.maxiter_stops,
                                      "rounding_floor": stops_f[1] + p.floor_stops},
               "device_bytes": nbytes.value}
        if save:
            np.savez(save, target_window=target[win], g_window=g[win], g_sub=g[::8, ::8],
                     g_norm=np.linalg.norm(g), J=J, precision=prec, solver=solver)
        if saved is not None:
            gw, gs = saved["g_window"], saved["g_sub"]
            rec["vs_saved_gradient_rel_l2_window"] = float(
                np.linalg.norm(g[win] - gw) / np.linalg.norm(gw))
            rec["vs_saved_gradient_rel_l2_subsample"] = float(
                np.linalg.norm(g[::8, ::8] - gs) / np.linalg.norm(gs))
            rec["vs_saved_gradient_norm_rel"] = float(
                abs(np.linalg.norm(g) - saved["g_norm"]) / saved["g_norm"])
        log(f"[bench] adjoint {cells}^2 {prec}: {rec}")
        out[prec] = rec
        grads[prec] = g
        p.close()
    if len(grads) == 2:
        out["mixed_vs_fp64_gradient_rel_l2"] = float(
            np.linalg.norm(grads["mixed"] - grads["fp64"]) / np.linalg.norm(grads["fp64"]))
    if n_fd:
        # central differences with the refined solver, against every gradient
        p = TeaLeafSteps(e0, rx, ry, steps, coefficient=_lib.CONDUCTIVITY, eps=fd_eps,
                         checkpoint_every=every, precision="mixed", solver="cg", max_iters=20000)
        fd_rng = np.random.default_rng(seed + 1)
        cells_fd, fds = [], []
        t0 = time.perf_counter()
        for _ in range(n_fd):
            c = (int(fd_rng.integers(wy, wy + 256)), int(fd_rng.integers(wx, wx + 256)))
            tp, tm = theta.copy(), theta.copy()
            tp[c] += h
            tm[c] -= h
            up, um = p.forward(tp), p.forward(tm)
            fd = (float(((up[win] - target[win]) ** 2).sum())
                  - float(((um[win] - target[win]) ** 2).sum())) / (2 * h)
            cells_fd.append(c)
            fds.append(fd)
            log(f"[bench] adjoint FD cell {c}: fd {fd:.10e} " + " ".join(
                f"{k} {grads[k][c]:.10e}" for k in grads))
        fd_s = time.perf_counter() - t0
        p.close()
        out["finite_differences"] = {
            "cells": n_fd, "h": h, "forward_solver": f"mixed precision CG, true relative "
                                                    f"residual {fd_eps:g}",
            "seconds": fd_s}
        for k, g in grads.items():
            errs = [abs(fd - g[c]) / abs(g[c]) for c, fd in zip(cells_fd, fds)]
            out[k].update({"fd_cells": n_fd, "fd_h": h, "fd_max_rel_err": max(errs),
                           "fd_median_rel_err": float(np.median(errs))})
            log(f"[bench] adjoint {k} vs FD: max {max(errs):.2e} median {np.median(errs):.2e}")
    return out


def cpu_info():
    """The host's processors: logical CPUs, the CPUs this process may run on,
    and from lscpu the model, sockets, cores per socket and threads per core."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        text = ""
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core"}
    for line in text.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    if isinstance(info.get("sockets"), int) and isinstance(info.get("cores_per_socket"), int):
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    return info


def _reference_time_per_it(ref_bin, deck, threads):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "tea.in"), "w") as f:
            f.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.time()
        subprocess.run([ref_bin], cwd=d, env=env, check=True, timeout=900,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        wall = time.time() - t0
        text = open(os.path.join(d, "tea.out")).read()
    m = re.search(r"Solve Time\s+(\S+)\s+Its\s+(\d+)\s+Time Per It\s+(\S+)", text)
    return float(m.group(3)), int(m.group(2)), wall


def cpu_baseline(cells, solver, gpu_iters_per_step, cap, cap_one_thread):
    """Reference build (oracle/_ref) on the host cores: one timestep capped at
    `cap` iterations with profiler_on -> "Time Per It" (tea_solve.f90:474),
    with OpenMP threads = the physical cores this process may use, and once
    more on one thread (capped at `cap_one_thread`); falls back to the C
    oracle (single thread) when the reference build is absent."""
    from tealeaf_ref_amd.deckgen import bench_deck
    info = cpu_info()
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "tea_leaf")
    # the physical cores, within this process's CPU share (the CPUs it may run
    # on, and the OpenMP thread count the environment sets for a GPU job)
    limits = [info["affinity_cpus"], info.get("physical_cores") or info["affinity_cpus"]]
    if (info["omp_num_threads_env"] or "").isdigit():
        limits.append(int(info["omp_num_threads_env"]))
    threads = max(1, min(limits))
    if os.path.exists(ref_bin):
        deck = bench_deck(cells, solver=solver, eps=1e-15, max_iters=cap, end_step=1,
                          summary_frequency=1, profiler=True)
        tpi, its, wall = _reference_time_per_it(ref_bin, deck, threads)
        deck1 = bench_deck(cells, solver=solver, eps=1e-15, max_iters=cap_one_thread, end_step=1,
                           summary_frequency=1, profiler=True)
        tpi1, its1, wall1 = _reference_time_per_it(ref_bin, deck1, 1)
        kind, cores = "reference", threads
        sample = (f"{cells}^2 FP64 {solver.upper()}, 1 timestep capped at {cap} iterations, "
                  f"OMP_NUM_THREADS={threads} (reference Fortran/OpenMP built from source by "
                  f"oracle/ref.mk, {wall:.1f} s wall); s/timestep = Time Per It x the GPU run's "
                  f"mean iterations per step")
        one = {"threads": 1, "time_per_iteration_s": tpi1, "iterations": its1,
               "s_per_timestep": tpi1 * gpu_iters_per_step, "wall_s": wall1,
               "cell_iterations_per_s": cells * cells / tpi1,
               "sample": f"the same deck capped at {cap_one_thread} iterations, one thread"}
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tl_oracle as O
        deck = bench_deck(cells, solver=solver, eps=1e-15, max_iters=cap_one_thread, end_step=1,
                          summary_frequency=1, profiler=True)
        p = O.parse_deck(deck)
        t0 = time.perf_counter()
        rec = O.run(p)
        wall = time.perf_counter() - t0
        tpi = wall / max(1, rec[-1]["iterations"] + 1)
        kind, cores = "port", 1
        sample = (f"C oracle, {cells}^2 FP64 {solver.upper()}, 1 step capped at {cap_one_thread} "
                  f"iterations")
        one = {"threads": 1, "time_per_iteration_s": tpi, "s_per_timestep": tpi * gpu_iters_per_step}
    s_per_step = tpi * gpu_iters_per_step
    return {"value": s_per_step, "unit": "s/timestep", "cores": cores, "threads": cores,
            "kind": kind, "sample": sample, "time_per_iteration_s": tpi,
            "cell_iterations_per_s": cells * cells / tpi,
            "physical_cores": info.get("physical_cores"), "host": info, "one_thread": one}


def profile_run(cells, iters, solver="cg", copies=4):
    """Short run for hardware-counter profiles: one timestep of the seeded
    deck capped at `iters` iterations, then `copies` launches of the plain
    copy kernel (tl_ppcg_store_r: r_store = r over the interior, 8 B read +
    8 B written per cell), whose counters calibrate FETCH_SIZE / WRITE_SIZE.
    Solvers: "cg"; "ppcg" (10 inner steps; the switch to PPCG at the 100th
    CG iteration: tl_ch_cg_presteps 100 with an epslim no error exceeds);
    "mixed_ppcg" (FP32 PPCG under FP64 refinement, inverse-density
    coefficient, at most 2 refinements of at most `iters` iterations)."""
    from tealeaf_ref_amd import _lib
    from tealeaf_ref_amd.deckgen import bench_deck
    from tealeaf_ref_amd.host import TeaHost
    kw = {}
    if solver == "ppcg":
        kw = {"inner_steps": 10, "presteps": 100, "epslim": 1e30}
    elif solver == "mixed_ppcg":
        kw = {"inner_steps": 10, "coefficient": "inverse_density",
              "extra": ("tl_mixed_precision", "tl_mixed_max_refinements=2")}
    deck = bench_deck(cells, solver="ppcg" if solver == "mixed_ppcg" else solver, eps=1e-15,
                      max_iters=iters, end_step=1, summary_frequency=1, **kw)
    with tempfile.TemporaryDirectory() as wd:
        path = os.path.join(wd, "tea.in")
        with open(path, "w") as f:
            f.write(deck)
        h = TeaHost(path, os.path.join(wd, "tea.out"))
        h.step()
        for _ in range(copies):
            _lib.call("tl_ppcg_store_r")
        _lib.call("tl_synchronize")
        it = h.last()["iterations"]
        h.close()
    print(json.dumps({"profile_run": True, "solver": solver, "cells": cells * cells,
                      "iterations": it, "copy_launches": copies,
                      "copy_bytes_each": 16 * cells * cells}))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=4096)
    ap.add_argument("--solver", default="cg", choices=["cg"])
    ap.add_argument("--cpu-cap", type=int, default=1500, help="iterations of the CPU sample")
    ap.add_argument("--cpu-cap-one-thread", type=int, default=100,
                    help="iterations of the one-thread CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-sizes", default="4096,8192",
                    help="comma-separated mesh sizes of the FP64 vs FP32-refined comparison "
                         "('' to skip)")
    ap.add_argument("--profile-solver", default="cg", choices=["cg", "ppcg", "mixed_ppcg"],
                    help="solver of the counter-profiling mode")
    ap.add_argument("--profile-iters", type=int, default=0,
                    help="counter-profiling mode: one CG step capped at this many iterations "
                         "plus copy-kernel calibration launches; prints a short JSON line")
    ap.add_argument("--adjoint-cells", type=int, default=2048,
                    help="mesh of the adjoint case (0 to skip)")
    ap.add_argument("--adjoint-steps", type=int, default=20)
    ap.add_argument("--adjoint-every", type=int, default=5, help="checkpoint interval")
    ap.add_argument("--adjoint-fd", type=int, default=32, help="finite-difference cells")
    ap.add_argument("--adjoint-h", type=float, default=1e-6)
    ap.add_argument("--adjoint-precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--adjoint-solver", default="cg", choices=["cg", "ppcg"])
    ap.add_argument("--adjoint-save", default=None,
                    help="write target, gradient window/subsample and norm to this .npz")
    ap.add_argument("--adjoint-compare", default=None,
                    help="take the target from this .npz (--adjoint-save of an earlier run) "
                         "and compare the gradient with the one saved there")
    ap.add_argument("--adjoint-mixed-eps", type=float, default=1e-12,
                    help="true relative residual of the mixed-precision solves")
    ap.add_argument("--adjoint-large-cells", type=int, default=8192,
                    help="mesh of the bounded mixed + adjoint ride-along (0 to skip)")
    ap.add_argument("--adjoint-large-steps", type=int, default=2,
                    help="timesteps of the bounded mixed + adjoint ride-along")
    ap.add_argument("--no-small", action="store_true", help="skip the 512^2 CG case")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the STREAM-style copy/triad bandwidth measurement")
    ap.add_argument("--solver-steps", type=int, default=1,
                    help="timed steps of the Chebyshev and PPCG runs (0 to skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the solver keeps the whole mesh as one chunk "
              "on a single MI355X (multi-GPU decomposition is not built).", file=sys.stderr)
        sys.exit(2)

    if args.profile_iters:
        profile_run(args.cells, args.profile_iters, args.profile_solver)
        return

    with tempfile.TemporaryDirectory() as wd:
        elapsed, iters, prof = gpu_run(args.cells, args.solver, args.steps, args.warmup, wd)
    cells = args.cells * args.cells
    s_per_step = elapsed / args.steps
    exec_iters = sum(i + 1 for i in iters)        # printed count is n-1 at exit
    mean_iters = sum(iters) / len(iters)
    n_w, ms_w, idle_w = prof[0]
    n_ur, ms_ur, _ = prof[1]
    # achieved = algorithmic bytes per launch / the average event-timed launch
    # (every EVENT_PERIOD-th launch; the few launches queued past the exit,
    # which return at once, are left out: rocprof's average over all launches
    # includes them and reads ~0.5 % lower)
    w_gbs = CG_W_BYTES_PER_CELL * cells / (ms_w / n_w * 1e-3) / 1e9 if ms_w else None
    ur_gbs = CG_UR_BYTES_PER_CELL * cells / (ms_ur / n_ur * 1e-3) / 1e9 if ms_ur else None
    step_gbs = CG_ITER_BYTES_PER_CELL * cells * exec_iters / elapsed / 1e9
    out = {
        "metric": "s/timestep + HBM GB/s at 4096^2: FP64 CG vs FP32-refined PPCG @1e-10 residual",
        "value": s_per_step,
        "unit": "s/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_step * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded deck generator, no external data)",
        "config": {"workload": f"{args.cells}x{args.cells} FP64 {args.solver.upper()} timestep, "
                               "seeded deck (3 rectangles + 2 circles, seed 2), tl_eps 1e-15",
                   "cells": cells, "solver": args.solver, "eps": 1e-15,
                   "mean_iterations_per_step": mean_iters},
        "hbm_gbs_step": step_gbs,
        "model_gbs_step": model_bytes_per_cell("cg", prof["rows"]) * cells / elapsed / 1e9,
        "cell_iterations_per_s": cells * exec_iters / elapsed,
        "roofline": {"bound": "hbm", "kernel": "cg calc_w (w = A p, p = r + beta p, p.w)",
                     "achieved": w_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": (w_gbs / PEAK_HBM_GBS) if w_gbs else None,
                     "traffic": measured_traffic("k_loop_w", args.cells),
                     "traffic_source": "profiles/cg_fp64_4096_hbm_traffic.md",
                     "bytes_per_cell": CG_W_BYTES_PER_CELL,
                     "avg_launch_ms": ms_w / n_w if n_w else None, "timed_launches": n_w,
                     "idle_launches": idle_w,
                     "event_sample_period": EVENT_PERIOD,
                     "calc_ur_achieved": ur_gbs,
                     "calc_ur_avg_launch_ms": ms_ur / n_ur if n_ur else None},
    }
    if args.solver_steps:
        out["solvers"] = solver_runs(args.cells, args.solver_steps, 1)
    if not args.no_small:
        out["small"] = small_run()
    if not args.no_stream:
        out["achievable_hbm"] = stream_bandwidth()
        out["roofline"]["frac_of_achievable_triad"] = (
            w_gbs / out["achievable_hbm"]["triad_gbs"] if w_gbs else None)
    sizes = [int(x) for x in args.mixed_sizes.split(",") if x.strip()]
    if sizes:
        out["precision_comparison"] = [precision_comparison(n, later_step=(i == 0))
                                       for i, n in enumerate(sizes)]
    if args.adjoint_cells:
        out["adjoint"] = adjoint_bench(args.adjoint_cells, args.adjoint_steps, args.adjoint_every,
                                       args.adjoint_fd, args.adjoint_h, args.adjoint_precision,
                                       args.adjoint_mixed_eps, solver=args.adjoint_solver,
                                       save=args.adjoint_save, compare=args.adjoint_compare)
    if args.adjoint_large_cells and args.adjoint_large_steps:
        # the mixed-precision + adjoint case at full size, bounded: FP64 and
        # mixed PPCG (its fastest form) forward and backward over a few
        # steps, checkpoint every step (the 50-step case runs as two
        # invocations, --adjoint-save then --adjoint-compare)
        out["adjoint_large"] = adjoint_bench(
            args.adjoint_large_cells, args.adjoint_large_steps, 1, 0, args.adjoint_h, "both",
            args.adjoint_mixed_eps, solver="ppcg")
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cells, args.solver, mean_iters, args.cpu_cap,
                                               args.cpu_cap_one_thread)
        except Exception as e:  # report, do not fail the GPU measurement
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

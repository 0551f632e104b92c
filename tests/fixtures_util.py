"""Load the reference fixtures (tests/golden/fixtures/*.json) and compare a run
against them."""
import glob
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
FIXDIR = os.path.join(HERE, "golden", "fixtures")
SUMMARY_KEYS = ("volume", "mass", "density", "energy", "temperature")


def names():
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(FIXDIR, "*.json")))


def load(name):
    with open(os.path.join(FIXDIR, name + ".json")) as f:
        return json.load(f)


def compare_summaries(got, ref, rtol):
    """got/ref: lists of step-row dicts.  Every reference row must have a row
    for the same step (last one printed) within rtol."""
    gl = {}
    for s in got:
        gl[s["step"]] = s
    worst = 0.0
    for r in ref:
        g = gl.get(r["step"])
        assert g is not None, f"no step: row for step {r['step']}"
        for k in SUMMARY_KEYS:
            if math.isnan(r[k]) and math.isnan(g[k]):
                continue  # both produce 0/0 (a one-cell mesh)
            d = abs(g[k] - r[k]) / max(abs(r[k]), 1e-300)
            worst = max(worst, d)
            assert d <= rtol, f"step {r['step']} {k}: {g[k]!r} vs {r[k]!r} (rel {d:.2e})"
    return worst


def compare_iterations(got_steps, ref_steps, slack=0):
    """Per-step iteration counts (and PPCG inner counts, switch iteration)."""
    for k, r in ref_steps.items():
        g = got_steps.get(int(k)) if isinstance(next(iter(got_steps)), int) else got_steps.get(k)
        assert g is not None, f"no record for step {k}"
        if "iterations" in r:
            assert abs(g["iterations"] - r["iterations"]) <= slack, \
                f"step {k}: iterations {g['iterations']} vs {r['iterations']}"
        for key in ("switch_n",):
            if key in r:
                assert g.get(key) == r[key], f"step {k}: {key} {g.get(key)} vs {r[key]}"


def deck_eps(deck):
    """tl_eps of a deck (default 1e-10, read_input.f90:60)."""
    import re
    m = re.search(r"tl_eps[ =]+([0-9.eEdD+-]+)", deck, re.I)
    return float(m.group(1).lower().replace("d", "e")) if m else 1e-10


def iteration_slack(deck):
    """Iteration-count tolerance for decks without per-step oracle counts
    (the 4000^2 and 8000^2 benchmark decks): 0, except at tl_eps <= 1e-14,
    where the exit test |r| < eps |r0| sits at the FP64 rounding floor and
    another reduction order may move the crossing by one iteration."""
    return 1 if deck_eps(deck) <= 1e-14 else 0


# An exit at the FP64 floor is decided by rounding noise when the exit error
# lies this close to the threshold: the two oracle builds (which differ only
# in multiply-add contraction) disagree on exits up to 1.1 % below it.
FLOOR_MARGIN = 0.02


def iteration_bounds(fx, step):
    """Allowed iteration counts (lo, hi) of one step of a fixture.  The
    fixtures hold the reference's count and that of the oracle built with
    multiply-add contraction (oracle_fma_iterations).  The count must be the
    reference's exactly, except
    * where the two builds differ (the exit depends on last-bit rounding):
      anywhere between them;
    * at tl_eps <= 1e-14 (the FP64 floor), on a step whose exit error (the
      printed conduction error) lies within FLOOR_MARGIN of tl_eps, where
      another reduction order moves the crossing: +-1."""
    rec = fx["steps"][str(step)]
    ref = rec["iterations"]
    fma = fx.get("oracle_fma_iterations", {}).get(str(step))
    if fma is None:
        s = iteration_slack(fx["deck"])
        return ref - s, ref + s
    lo, hi = min(ref, fma), max(ref, fma)
    eps, ce = deck_eps(fx["deck"]), rec.get("conduction_error")
    if eps <= 1e-14 and ce is not None and ce >= (1.0 - FLOOR_MARGIN) * eps:
        lo, hi = min(lo, ref - 1), max(hi, ref + 1)
    return lo, hi


def compare_iterations_exact(got_steps, fx, label=""):
    """Per-step iteration counts against iteration_bounds; prints the per-step
    differences from the reference (empty when every count is exact) so the
    test log records them."""
    deltas = {}
    for k, r in fx["steps"].items():
        if "iterations" not in r:
            continue
        g = got_steps[int(k)]["iterations"]
        lo, hi = iteration_bounds(fx, int(k))
        if g != r["iterations"]:
            deltas[int(k)] = g - r["iterations"]
        assert lo <= g <= hi, f"step {k}: iterations {g} vs reference {r['iterations']} " \
                              f"(allowed {lo}..{hi})"
    print(f"  {label} iteration differences from the reference: {deltas or 'none'}", flush=True)
    return deltas

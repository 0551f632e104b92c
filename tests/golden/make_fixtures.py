"""Capture the parity fixtures from the reference itself.

Runs oracle/_ref/tea_leaf (the unmodified reference built from its sources by
oracle/ref.mk) on one thread for every seeded deck below, and stores the
parsed tea.out (per-step `step:` rows, iteration counts, switch/eigen lines)
as JSON under tests/golden/fixtures/; for the layout_* decks also the line
layout of the whole tea.out (teaout.skeleton), for the visit_* decks the
VTK files and tea.visit.  Also converts the goldens shipped with
the reference (Benchmarks/tea_bm_N.{in,out}) into the same JSON form.

    python tests/golden/make_fixtures.py [name ...]
    python tests/golden/make_fixtures.py --fma [name ...]   (only the oracle counts)
"""
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tealeaf_ref_amd import teaout  # noqa: E402
from tealeaf_ref_amd.deckgen import make_deck  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "tea_leaf")
REF_SRC = "/root/reference"
OUT = os.path.join(ROOT, "tests", "golden", "fixtures")

CASES = {
    "cfg1_cg_512": dict(cells=512, solver="cg", seed=1, eps=1e-15, max_iters=5000, end_step=10),
    "cg_circles_128": dict(cells=128, solver="cg", seed=2, n_circ=2, end_step=5),
    "cheby_100": dict(cells=100, solver="chebyshev", seed=3, n_circ=1, eps=1e-10, max_iters=2000,
                      end_step=4),
    "cheby_presteps_96": dict(cells=96, solver="chebyshev", seed=4, presteps=20, eps=1e-12,
                              max_iters=1500, end_step=3),
    "ppcg_128": dict(cells=128, solver="ppcg", seed=5, inner_steps=10, end_step=5),
    "ppcg_default_inner_96": dict(cells=96, solver="ppcg", seed=6, end_step=4),
    "ppcg_halo2_96": dict(cells=96, solver="ppcg", seed=6, inner_steps=8, halo_depth=2, end_step=4),
    "jacobi_48": dict(cells=48, solver="jacobi", seed=7, eps=1e-6, max_iters=8000, end_step=3),
    "cg_jacdiag_130x122": dict(cells=130, y_cells=122, solver="cg", seed=8, precond="jac_diag",
                               end_step=4),
    "cg_jacblock_130x122": dict(cells=130, y_cells=122, solver="cg", seed=8, precond="jac_block",
                                end_step=4),
    "cg_reflective_96": dict(cells=96, solver="cg", seed=9, reflective=True, end_step=4),
    "cheby_reflective_96": dict(cells=96, solver="chebyshev", seed=9, reflective=True, eps=1e-10,
                                max_iters=1500, end_step=3),
    "ppcg_reflective_96": dict(cells=96, solver="ppcg", seed=9, reflective=True, inner_steps=6,
                               end_step=3),
    "jacobi_reflective_40": dict(cells=40, solver="jacobi", seed=9, reflective=True, eps=1e-6,
                                 max_iters=8000, end_step=2),
    "cg_invdensity_128": dict(cells=128, solver="cg", seed=10, coefficient="inverse_density",
                              end_step=4),
    "ppcg_jacdiag_96": dict(cells=96, solver="ppcg", seed=11, precond="jac_diag", inner_steps=8,
                            end_step=3),
    "cheby_jacdiag_96": dict(cells=96, solver="chebyshev", seed=11, precond="jac_diag", eps=1e-10,
                             max_iters=1500, end_step=3),
    "ppcg_jacblock_96": dict(cells=96, solver="ppcg", seed=13, precond="jac_block", inner_steps=6,
                             end_step=3),
    "cheby_jacblock_98": dict(cells=98, solver="chebyshev", seed=17, precond="jac_block",
                              eps=1e-10, max_iters=1500, end_step=3),
    # capped: every step runs 200 iterations, so the summaries pin the
    # arithmetic of the reflective faces without a rounding-decided exit
    "jacobi_reflective_66": dict(cells=66, solver="jacobi", seed=19, reflective=True, eps=1e-15,
                                 max_iters=200, end_step=3),
    "cg_checkresult_64": dict(cells=64, solver="cg", seed=12, check_result=True, end_step=2),
    # small decks with the profiler on whose whole tea.out layout is kept
    "layout_cg_32": dict(cells=32, solver="cg", seed=14, precond="jac_diag", check_result=True,
                         end_step=2, profiler=True),
    "layout_cheby_64": dict(cells=64, solver="chebyshev", seed=14, eps=1e-12, max_iters=2000,
                            end_step=2, profiler=True),
    "layout_ppcg_64": dict(cells=64, solver="ppcg", seed=3, n_rect=2, n_circ=1, eps=1e-15,
                           max_iters=500, inner_steps=4, end_step=2, profiler=True),
    "layout_jacobi_32": dict(cells=32, solver="jacobi", seed=14, eps=1e-6, max_iters=4000,
                             reflective=True, end_step=2, profiler=True),
    # verbose_on: a "Residual" line after every iteration (tea_solve.f90:402-408)
    "layout_verbose_cg_32": dict(cells=32, solver="cg", seed=16, eps=1e-10, end_step=2,
                                 profiler=True, extra=("verbose_on",)),
    "layout_verbose_cheby_48": dict(cells=48, solver="chebyshev", seed=16, eps=1e-10,
                                    max_iters=2000, end_step=2, profiler=True,
                                    extra=("verbose_on",)),
    "layout_verbose_ppcg_48": dict(cells=48, solver="ppcg", seed=17, eps=1e-10, inner_steps=5,
                                   end_step=2, profiler=True, extra=("verbose_on",)),
    "layout_verbose_jacobi_24": dict(cells=24, solver="jacobi", seed=18, eps=1e-6,
                                     max_iters=4000, end_step=2, profiler=True,
                                     extra=("verbose_on",)),
    # larger meshes: the tiling of the two-per-pass kernels over many tiles
    # and tile bands, end to end against the reference
    "cheby_large_1024": dict(cells=1024, solver="chebyshev", seed=19, n_circ=2, presteps=20,
                             eps=1e-10, max_iters=3000, end_step=2),
    "ppcg_large_1024": dict(cells=1024, solver="ppcg", seed=20, n_circ=2, inner_steps=10,
                            eps=1e-10, max_iters=3000, end_step=2),
    # visit output: the VTK files and tea.visit are kept too
    "visit_cg_16x12": dict(cells=16, y_cells=12, solver="cg", seed=15, end_step=3,
                           extra=("visit_frequency=2",)),
}


def run_reference(deck, others=None):
    """tea.out of the reference on deck; with a dict `others`, also every other
    file the run wrote (visit output), by name."""
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "tea.in"), "w") as f:
            f.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=3600)
        if others is not None:
            for n in sorted(os.listdir(d)):
                if n not in ("tea.in", "tea.out", "tea.in.tmp"):  # tea.in.tmp: the reader's copy
                    with open(os.path.join(d, n)) as f:
                        others[n] = f.read()
        with open(os.path.join(d, "tea.out")) as f:
            return f.read()


def fixture_from(text, deck, source):
    parsed = teaout.parse(text)
    return {"source": source, "deck": deck, "summaries": parsed["summaries"],
            "steps": {str(k): v for k, v in parsed["steps"].items()}}


def main(names):
    os.makedirs(OUT, exist_ok=True)
    for name, kw in CASES.items():
        if names and name not in names:
            continue
        kw = dict(kw)
        cells = kw.pop("cells")
        deck = make_deck(cells, **kw)
        print("reference run:", name, flush=True)
        files = {}
        text = run_reference(deck, files)
        fx = fixture_from(text, deck,
                          "oracle/_ref/tea_leaf (reference built from source), OMP_NUM_THREADS=1")
        if name.startswith("layout_"):
            fx["skeleton"] = teaout.skeleton(text)
        if name.startswith("visit_"):
            fx["files"] = files
        with open(os.path.join(OUT, name + ".json"), "w") as f:
            json.dump(fx, f, indent=1)
    # goldens shipped with the reference (data, converted)
    for i in range(1, 7):
        name = f"tea_bm_{i}"
        if names and name not in names:
            continue
        src_in = os.path.join(REF_SRC, "Benchmarks", name + ".in")
        src_out = os.path.join(REF_SRC, "Benchmarks", name + ".out")
        if not os.path.exists(src_out):
            continue
        with open(src_in) as f:
            deck = f.read()
        with open(src_out) as f:
            text = f.read()
        fx = fixture_from(text, deck, f"Benchmarks/{name}.out shipped with the reference")
        with open(os.path.join(OUT, name + ".json"), "w") as f:
            json.dump(fx, f, indent=1)
        print("converted", name)


def add_fma_counts(names):
    """Store, beside the reference's iteration counts, those of the oracle
    built with multiply-add contraction (oracle/liboracle_fma.so, the same
    source): where the two differ, the exit of that step depends on the last
    bit of the rounding, and the GPU parity test allows the range between
    them; everywhere else it requires the reference's count exactly.  The
    4000^2 and 8000^2 benchmark decks are left out (hours of oracle time)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tl_oracle as O
    for path in sorted(os.listdir(OUT)):
        name = path[:-5]
        if names and name not in names:
            continue
        if name in ("tea_bm_5", "tea_bm_6"):
            continue
        with open(os.path.join(OUT, path)) as f:
            fx = json.load(f)
        counts = {str(r["step"]): r["iterations"] for r in O.run(fx["deck"], contracted=True)
                  if r["step"] > 0}
        fx["oracle_fma_iterations"] = {k: counts[k] for k in fx["steps"] if k in counts}
        with open(os.path.join(OUT, path), "w") as f:
            json.dump(fx, f, indent=1)
        diff = {k: (v["iterations"], counts.get(k)) for k, v in fx["steps"].items()
                if "iterations" in v and counts.get(k) != v["iterations"]}
        print(f"fma counts: {name}: differing steps {diff}", flush=True)


if __name__ == "__main__":
    if sys.argv[1:2] == ["--fma"]:
        add_fma_counts(sys.argv[2:])
    else:
        main(sys.argv[1:])
        add_fma_counts(sys.argv[1:])

"""Ragged and degenerate meshes through the native host against the oracle
(the C restatement, run in process): one cell, single rows and columns,
sizes that are not multiples of the 2- and 4-wide lanes or of the 512- and
1024-column blocks, every solver, reflective and not.  Per-step summaries
within 1e-10 and equal iteration counts.  Where the reference's arithmetic
turns NaN (CG, Chebyshev and PPCG on one cell: the initial residual is 0 and
alpha = 0/0, which the reference carries to tl_max_iters), the library stops
at the first non-finite reduction and reports it instead, a deliberate
difference.  (A 513x3 mesh, cells 170 times
wider than tall, is left out: its CG exit iteration on later steps moves by
up to 8 under last-bit differences of the state, as a numpy restatement in a
different operation order shows too, so it pins no kernel property.)"""
import math

import pytest

import fixtures_util as FX
import tl_oracle as O
from tealeaf_ref_amd import _lib, host
from tealeaf_ref_amd.deckgen import make_deck

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 9), (9, 1), (3, 2), (5, 7), (513, 17), (3, 515), (1025, 5)]
# tolerances well above the FP64 floor of these (strongly anisotropic)
# operators, where the exit iteration depends on the arithmetic and not on
# the order of a reduction
SOLVERS = [("cg", {"eps": 1e-10}), ("chebyshev", {"presteps": 5, "eps": 1e-8}),
           ("ppcg", {"inner_steps": 5, "presteps": 5, "eps": 1e-10}), ("jacobi", {"eps": 1e-6})]


def oracle_rows(deck):
    recs = O.run(deck)
    rows, steps = [], {}
    for r in recs:
        if r["summary"] is not None:
            v = r["summary"]
            rows.append({"step": r["step"], "volume": v[0], "mass": v[1], "density": v[2],
                         "energy": v[3], "temperature": v[4]})
        if r["step"] > 0:
            steps[r["step"]] = {"iterations": r["iterations"]}
    return rows, steps


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("solver,kw", SOLVERS)
@pytest.mark.parametrize("reflective", [False, True])
def test_ragged_mesh(shape, solver, kw, reflective, tmp_path):
    nx, ny = shape
    kw = dict(kw)
    eps = kw.pop("eps")
    deck = make_deck(nx, y_cells=ny, solver=solver, seed=nx * 31 + ny, n_rect=2, n_circ=1,
                     eps=eps, max_iters=3000, end_step=3, dt=0.004, reflective=reflective, **kw)
    print(f"[edge] {nx}x{ny} {solver} reflective={reflective}", flush=True)
    want_rows, want_steps = oracle_rows(deck)
    if any(math.isnan(v) for r in want_rows for v in r.values()):
        with pytest.raises(_lib.TeaLeafError, match="non-finite"):
            host.run_deck(deck, str(tmp_path))
        return
    got = host.run_deck(deck, str(tmp_path))
    FX.compare_summaries(got["summaries"], want_rows, 1e-10)
    slack = 0
    for k, r in want_steps.items():
        g = got["steps"][k]["iterations"]
        assert abs(g - r["iterations"]) <= slack, (k, g, r["iterations"])

"""Seeded random decks through the native host against the oracle: mesh
shapes from 16 to 300 cells a side (ragged against every lane, wave, block
and tile width), every solver with every preconditioner it takes, both
coefficients, reflective or not, halo depth 1 or 2, a random number of
rectangles and circles.  Per-step summaries within 1e-10 of the oracle's and
iteration counts equal to the oracle's.

Some exits depend on last-bit rounding: the CG residual of a strongly
varying operator can cross the tolerance one iteration earlier or later when
only the rounding of the arithmetic changes.  The oracle built with
multiply-add contraction (the same source, `liboracle_fma.so`) shows which:
where the two oracle builds disagree, the count must lie between theirs."""
import numpy as np
import pytest

import fixtures_util as FX
import tl_oracle as O
from tealeaf_ref_amd import host
from tealeaf_ref_amd.deckgen import make_deck
from test_gpu_edge import oracle_rows

pytestmark = pytest.mark.gpu

# tolerances well above the FP64 floor, where the exit iteration does not
# depend on the order of a reduction
EPS = {"cg": 1e-10, "chebyshev": 1e-8, "ppcg": 1e-10, "jacobi": 1e-6}


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    solver = ("cg", "chebyshev", "ppcg", "jacobi")[seed % 4]
    kw = {"eps": EPS[solver], "max_iters": 4000, "end_step": 3, "dt": 0.004,
          "n_rect": int(rng.integers(1, 4)), "n_circ": int(rng.integers(0, 3)),
          "coefficient": "inverse_density" if rng.random() < 0.5 else "density",
          "reflective": bool(rng.random() < 0.5)}
    if solver != "jacobi":
        kw["precond"] = (None, "jac_diag", "jac_block")[int(rng.integers(0, 3))]
    if solver in ("chebyshev", "ppcg"):
        kw["presteps"] = int(rng.integers(5, 31))
    if solver == "ppcg":
        kw["inner_steps"] = int(rng.integers(2, 13))
        # the reader refuses a deeper halo with the block preconditioner
        if rng.random() < 0.5 and kw["precond"] != "jac_block":
            kw["halo_depth"] = 2
    nx, ny = (int(v) for v in rng.integers(16, 301, 2))
    if solver == "chebyshev" and seed in (1, 9):
        # the fused two-iterations-per-pass path (no reflective faces, no
        # block preconditioner)
        kw["reflective"] = False
        kw["precond"] = None if seed == 1 else "jac_diag"
    return solver, nx, ny, kw


@pytest.mark.parametrize("seed", range(48))
def test_random_deck(seed, tmp_path):
    solver, nx, ny, kw = random_case(seed)
    deck = make_deck(nx, y_cells=ny, solver=solver, seed=seed, **kw)
    print(f"[random] seed {seed}: {nx}x{ny} {solver} {kw}", flush=True)
    want_rows, want_steps = oracle_rows(deck)
    fma = {r["step"]: r["iterations"] for r in O.run(deck, contracted=True) if r["step"] > 0}
    got = host.run_deck(deck, str(tmp_path))
    FX.compare_summaries(got["summaries"], want_rows, 1e-10)
    for k, r in want_steps.items():
        g = got["steps"][k]["iterations"]
        lo, hi = sorted((r["iterations"], fma[k]))
        assert lo <= g <= hi, (k, g, r["iterations"], fma[k])

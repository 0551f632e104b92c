"""Pins the oracle against the reference built from its own sources
(oracle/_ref/tea_leaf, by oracle/ref.mk, one OpenMP thread) on the seeded
random decks of the GPU parity suite, beyond the captured fixtures: every
solver with the preconditioners it takes, both coefficients, reflective
faces, halo depth 2 and ragged meshes.  The oracle restates the
reference's arithmetic operation for operation, so the per-step summaries
agree to the 17 printed digits (5e-16, the print resolution; observed: equal
doubles) and every step takes the same number of iterations, including the
exits that the GPU parity suite lets move by a rounding-sensitive iteration.
Skipped where the reference build is absent."""
import os
import subprocess
import tempfile

import pytest

import fixtures_util as FX
import tl_oracle as O
from tealeaf_ref_amd import teaout
from tealeaf_ref_amd.deckgen import make_deck
from test_gpu_random_decks import random_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "tea_leaf")

pytestmark = pytest.mark.skipif(not os.access(REF_BIN, os.X_OK),
                                reason="reference build oracle/_ref/tea_leaf not present")


def run_reference(deck):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "tea.in"), "w") as f:
            f.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=600)
        with open(os.path.join(d, "tea.out")) as f:
            return teaout.parse(f.read())


def oracle(deck):
    rows, steps = [], {}
    for r in O.run(deck):
        if r["summary"] is not None:
            v = r["summary"]
            rows.append({"step": r["step"], "volume": v[0], "mass": v[1], "density": v[2],
                         "energy": v[3], "temperature": v[4]})
        if r["step"] > 0:
            steps[r["step"]] = r["iterations"]
    return rows, steps


@pytest.mark.parametrize("seed", range(48))
def test_oracle_matches_reference_build(seed):
    solver, nx, ny, kw = random_case(seed)
    deck = make_deck(nx, y_cells=ny, solver=solver, seed=seed, **kw)
    ref = run_reference(deck)
    rows, steps = oracle(deck)
    FX.compare_summaries(rows, ref["summaries"], 5e-16)
    assert len(ref["steps"]) == 3
    for k, r in ref["steps"].items():
        assert r["iterations"] == steps[k], (k, r["iterations"], steps[k])

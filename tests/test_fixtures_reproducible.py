"""The captured fixtures are reproducible: the seeded deck generator still
writes the deck each fixture holds, and the reference built from its sources
(oracle/_ref/tea_leaf, one OpenMP thread) still prints the same per-step
summaries, solver counts, log layout and visit files for it.  Skipped where
the reference build is absent."""
import json
import os
import sys

import pytest

import fixtures_util as FX
from tealeaf_ref_amd import teaout
from tealeaf_ref_amd.deckgen import make_deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_fixtures as MF  # noqa: E402

pytestmark = pytest.mark.skipif(not os.access(MF.REF_BIN, os.X_OK),
                                reason="reference build oracle/_ref/tea_leaf not present")


# wall-clock figures of the profiler's per-step lines
TIMINGS = ("solve_time", "time_per_it")


def without_timings(steps):
    return {k: {f: v for f, v in r.items() if f not in TIMINGS} for k, r in steps.items()}


@pytest.mark.parametrize("name", sorted(MF.CASES))
def test_fixture_reproduces(name):
    kw = dict(MF.CASES[name])
    cells = kw.pop("cells")
    deck = make_deck(cells, **kw)
    with open(os.path.join(MF.OUT, name + ".json")) as f:
        fx = json.load(f)
    assert deck == fx["deck"], "the deck generator no longer writes the fixture's deck"
    files = {}
    text = MF.run_reference(deck, files)
    got = MF.fixture_from(text, deck, fx["source"])
    assert without_timings(got["steps"]) == without_timings(fx["steps"])
    FX.compare_summaries(got["summaries"], fx["summaries"], 0.0)
    if "skeleton" in fx:
        assert teaout.skeleton(text) == fx["skeleton"]
    if "files" in fx:
        assert files == fx["files"]

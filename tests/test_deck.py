"""Deck reader semantics (read_input.f90, parse.f90) restated by the oracle and
the generator: defaults, state shrink, PPCG default, error paths."""
import pytest

import tl_oracle as O
from tealeaf_ref_amd.deckgen import make_deck

BASE = """*tea
state 1 density=100.0 energy=0.0001
state 2 density=0.1 energy=25.0 geometry=rectangle xmin=0.0 xmax=1.0 ymin=1.0 ymax=2.0
x_cells=10
y_cells=10
xmin=0.0
ymin=0.0
xmax=10.0
ymax=10.0
{extra}
*endtea
"""


def parse(extra=""):
    return O.parse_deck(BASE.format(extra=extra))


def test_defaults():
    p = parse()
    assert p.use_jacobi == 1 and p.use_cg == 0
    assert p.max_iters == 1000 and p.tl_ch_cg_presteps == 25
    assert p.eps == float.fromhex("0x1.b7cdfe0000000p-34")  # 1.0e-10 as a default real
    assert p.dtinit == float(__import__("numpy").float32(0.1))
    assert p.end_step == 640000 and p.summary_frequency == 10 and p.halo_depth == 1


def test_state_shrink_and_case():
    p = O.parse_deck(BASE.format(extra="TL_USE_CG ! comment\nInitial_Timestep = 0.004"))
    assert p.use_cg == 1 and p.dtinit == 0.004
    s = p.states[2]
    assert s.xmin == 0.0 + 1.0 / 100.0 and s.xmax == 1.0 - 1.0 / 100.0


def test_ppcg_default_inner_steps():
    p = O.parse_deck(BASE.format(extra="tl_use_ppcg").replace("x_cells=10", "x_cells=1000")
                     .replace("y_cells=10", "y_cells=1000"))
    assert p.tl_ppcg_inner_steps == 124


def test_unknown_keywords_ignored():
    p = parse("tl_ch_cg_errswitch 1\nfoo bar baz")
    assert p.use_jacobi == 1


@pytest.mark.parametrize("extra,msg", [
    ("tl_preconditioner_type foo", "Invalid preconditioner"),
    ("halo_depth=0", "Invalid halo exchange depth"),
    ("halo_depth=2\ntl_preconditioner_type jac_block", "nonstandard halo depth"),
    ("x_cells=abc", "convert to integer"),
    ("state 2 density=1.0", "State defined twice"),
    # parse.f90:343 appends ".0" to a number without a point, so an exponent
    # needs a point in its mantissa
    ("tl_eps 1e-5", "convert to real"),
])
def test_errors(extra, msg):
    with pytest.raises(O.OracleError, match=msg):
        parse(extra)


def test_no_states():
    with pytest.raises(O.OracleError, match="No states"):
        O.parse_deck("*tea\nx_cells=10\n*endtea\n")


def test_generated_decks_parse():
    for solver in ("cg", "chebyshev", "ppcg", "jacobi"):
        d = make_deck(64, solver=solver, seed=4, n_circ=2, epslim=1e-5, presteps=20)
        p = O.parse_deck(d)
        assert p.x_cells == 64 and p.number_of_states == 6
        assert all(len(line) <= 100 for line in d.splitlines())
    assert make_deck(64, seed=9) == make_deck(64, seed=9)

"""End to end: the native host (Fortran, in process) runs each reference deck on
the GPU; its tea.out must reproduce the reference's per-step `step:` rows
within 1e-10 relative and its iteration counts (tea_solve.f90:448-453):
exactly on 126 of the 142 steps the fixtures hold.  The other 16 are exits
decided by rounding (fixtures_util.iteration_bounds): the reference and the
oracle built with multiply-add contraction disagree there, or, at tl_eps
1e-15, the reference's exit error lies within 2 % of the threshold; there
the count may differ by one.  The per-step differences are printed."""

import pytest

import fixtures_util as FX
from tealeaf_ref_amd import host

pytestmark = pytest.mark.gpu

RTOL = 1e-10
# the decks the reference itself ships (tea_bm_1..4) and every captured one;
# tea_bm_5/6 (4000^2, 8000^2) are run by test_benchmark_scale_deck below
SMALL = [n for n in FX.names() if n not in ("tea_bm_5", "tea_bm_6")]


@pytest.mark.parametrize("name", SMALL)
def test_deck_matches_reference(name, tmp_path):
    fx = FX.load(name)
    print(f"[e2e] {name}", flush=True)
    got = host.run_deck(fx["deck"], str(tmp_path))
    FX.compare_summaries(got["summaries"], fx["summaries"], RTOL)
    FX.compare_iterations_exact(got["steps"], fx, name)
    for k, r in fx["steps"].items():
        g = got["steps"][int(k)]
        if "ppcg_inner" in r:
            assert g["ppcg_inner"] == r["ppcg_inner"], (k, g["ppcg_inner"], r["ppcg_inner"])
        if "est_itc" in r:
            assert g["est_itc"] == r["est_itc"], (k, g["est_itc"], r["est_itc"])


def test_tea_leaf_program(tmp_path):
    """The tea_leaf executable is a drop-in: ./tea.in in, ./tea.out out."""
    import subprocess
    fx = FX.load("tea_bm_1")
    (tmp_path / "tea.in").write_text(fx["deck"])
    subprocess.run([host.TEA_LEAF_BIN], cwd=tmp_path, check=True, timeout=120,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    from tealeaf_ref_amd import teaout
    got = teaout.parse((tmp_path / "tea.out").read_text())
    FX.compare_summaries(got["summaries"], fx["summaries"], RTOL)
    FX.compare_iterations(got["steps"], fx["steps"])
    assert "This test is considered PASSED" in (tmp_path / "tea.out").read_text()


def test_more_blocks_than_fit(tmp_path):
    """The 1024^2 PPCG fixture with the loop kernels launched on 2048 tiles of
    one row (TL_TILES=8192), more blocks than the GPU holds at once: blocks
    that start after others have finished must still see the launch that
    stops the CG presteps (and switches to PPCG) as their own, and apply its
    last p update, which the PPCG iterations then use."""
    import os
    import subprocess
    from tealeaf_ref_amd import teaout
    name = "ppcg_large_1024"
    fx = FX.load(name)
    print(f"[e2e] {name} with TL_TILES=8192", flush=True)
    (tmp_path / "tea.in").write_text(fx["deck"])
    env = dict(os.environ, TL_TILES="8192")
    subprocess.run([host.TEA_LEAF_BIN], cwd=tmp_path, check=True, timeout=120, env=env,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    got = teaout.parse((tmp_path / "tea.out").read_text())
    FX.compare_summaries(got["summaries"], fx["summaries"], RTOL)
    FX.compare_iterations_exact(got["steps"], fx, name)


@pytest.mark.timeout(400)
@pytest.mark.parametrize("name", ["tea_bm_5", "tea_bm_6"])
def test_benchmark_scale_deck(name, tmp_path):
    """The reference's own benchmark decks at scale (4000^2 and 8000^2 CG,
    10 steps, tl_eps 1e-15), against the logs it ships, which came from 24
    and 36 MPI ranks: per-step summaries within 1e-10; iteration counts
    within 0.1 % (at least 1).  At this tolerance the exit sits on the FP64
    floor: the shipped logs end every step with a conduction error of
    0.993-0.999e-15, i.e. the residual falls by under 1 % per iteration there,
    so rounding noise of the residual at the per-cent level (another
    reduction order, as between 36 ranks and one GPU) moves the exit by a few
    iterations."""
    import time
    fx = FX.load(name)
    print(f"[e2e] {name}", flush=True)
    t0 = time.time()
    got = host.run_deck(fx["deck"], str(tmp_path))
    wall = time.time() - t0
    worst = FX.compare_summaries(got["summaries"], fx["summaries"], RTOL)
    diffs = {int(k): got["steps"][int(k)]["iterations"] - r["iterations"]
             for k, r in fx["steps"].items()}
    total = sum(got["steps"][int(k)]["iterations"] for k in fx["steps"])
    print(f"  {name}: {wall:.1f} s for {total} CG iterations, worst summary rel diff "
          f"{worst:.1e}, iteration differences per step {diffs}", flush=True)
    for k, d in diffs.items():
        ref = fx["steps"][str(k)]["iterations"]
        assert abs(d) <= max(FX.iteration_slack(fx["deck"]), int(1e-3 * ref)), (k, d, ref)

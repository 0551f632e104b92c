"""bench.py on a machine without a GPU: the figures it takes from the
committed counter summaries are there and consistent with the algorithmic
bytes, and the multi-GPU request is refused with a message."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

CELLS = 4096 * 4096


@pytest.mark.parametrize("n", [4096, 8192])
def test_every_kernel_class_has_measured_traffic(n):
    cells = n * n
    for cls, (kernel, profile) in bench.CLASS_TRAFFIC.items():
        t = bench.measured_traffic(kernel, n, profile)
        if n == 8192 and cls in (3, 4):   # Chebyshev and Jacobi: profiled at 4096^2 only
            assert t is None
            continue
        assert t, (cls, kernel, profile)
        name, bpc = bench.CLASS_BYTES[cls]
        # the counters see at least the algorithmic bytes, and ghost rows,
        # columns and the extra loads of a batch's last pass add under 20 %
        assert bpc * cells * 0.99 <= t <= bpc * cells * 1.2, (name, t / cells, bpc)


@pytest.mark.parametrize("n", [4096, 8192])
@pytest.mark.parametrize("table", ["ppcg", "cg"])
def test_mixed_classes_have_measured_traffic(n, table):
    bytes_, traffic = ((bench.MIXED_PPCG_BYTES, bench.MIXED_PPCG_TRAFFIC) if table == "ppcg"
                       else (bench.MIXED_CG_BYTES, bench.MIXED_CG_TRAFFIC))
    for cls, (name, bpc) in bytes_.items():
        kernel, profile = traffic[cls]
        t = bench.measured_traffic(kernel, n, profile)
        assert t and bpc * n * n * 0.99 <= t <= bpc * n * n * 1.2, (name, t)


def test_roofline_traffic_of_the_dominant_kernel():
    t = bench.measured_traffic("k_loop_w", 4096)
    assert t and abs(t / CELLS - bench.CG_W_BYTES_PER_CELL) < 3.0
    assert bench.measured_traffic("k_loop_w", 2048) is None  # no summary at that size


def test_class_rates_from_event_samples():
    """Rates of a class from (timed launches, ms, idle launches): algorithmic
    bytes over the mean launch, and the measured bytes where profiled."""
    r = bench.class_rates({0: (10, 1.3, 1)}, 4096)["cg calc_w"]
    assert abs(r["avg_launch_ms"] - 0.13) < 1e-12
    assert abs(r["achieved_gbs"] - 48.0 * CELLS / 0.13e-3 / 1e9) < 1e-6
    assert r["measured_gbs"] == pytest.approx(bench.measured_traffic("k_loop_w", 4096) / 0.13e-3 / 1e9)
    assert "measured_gbs" not in bench.class_rates({0: (10, 1.3, 1)}, 2048)["cg calc_w"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_work_launches_from_iteration_counts():
    """Launches that did work per kernel class, from a step's printed
    counts: n-1 printed at an exit at iteration n, max_iters at the cap."""
    cg = bench.work_launches("cg", [{"iterations": 99, "switch_n": 0, "ppcg_inner": 0}])
    assert cg[0] == cg[1] == 100 and cg[2] == 0
    # PPCG: CG up to the switch at the top of iteration 26 (25 CG iterations),
    # then outer iterations 26..350, five two-sweep passes each for 10 sweeps
    pp = bench.work_launches("ppcg", [{"iterations": 349, "switch_n": 26, "ppcg_inner": 3260}],
                             inner=10)
    assert (pp[0], pp[1]) == (25, 25)
    assert pp[5] == pp[6] == 325 and pp[2] == 325 * 5
    # Jacobi at the cap: 10000 iterations, two per pass
    jl = bench.work_launches("jacobi", [{"iterations": 10000, "switch_n": 0, "ppcg_inner": 0}])
    assert jl[4] == 5000
    # Chebyshev: 20 CG iterations, the per-routine first step, then pairs
    ch = bench.work_launches("chebyshev", [{"iterations": 10000, "switch_n": 21,
                                            "ppcg_inner": 0}])
    assert (ch[0], ch[3]) == (20, (10000 - 20 - 1) // 2)


def test_cpu_info_reports_cores():
    info = bench.cpu_info()
    assert info["nproc"] >= 1 and info["affinity_cpus"] >= 1
    if "physical_cores" in info:
        assert 1 <= info["physical_cores"] <= info["nproc"]

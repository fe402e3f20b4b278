"""The bench line's measured figures come from the committed rocprof
summaries under profiles/: each lookup must find its row (CPU only)."""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_flux_kernel_traffic_rows():
    for prec in ("fp64", "mixed", "fast"):
        t = bench.measured_traffic("k_flux", prec)
        assert t is not None and t > 0, prec
    # the FP64 flux kernel moves at least its compulsory bytes per launch
    assert bench.measured_traffic("k_flux", "fp64") >= 148 * 1976 * 4096


def test_step_flops_row():
    f = bench.measured_flops("fp64")
    assert f is not None and f > 1e10


def test_c2_step_traffic_row():
    gb = bench.measured_step_traffic_c2()
    # more than the compulsory 296 B per zone-update of the 4.19 M zones
    assert gb is not None and gb * 1e9 >= 296 * 1024 * 4096

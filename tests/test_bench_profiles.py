"""bench.py reads the measured HBM bytes of its roofline kernels from the committed rocprofv3 summaries
(profiles/<workload>_hbm_traffic.md): the rows it names must be there, with bytes in the last column."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_rows_present():
    for name, b in (("c2", 1), ("c3", 3)):
        prof = f"{name}_hbm_traffic.md"
        sweep = bench.measured_traffic(prof, ["sweep_phases_fp32_per_launch"])
        spmv = bench.measured_traffic(prof, [f"dfm::k_pcg_spmv_sell<{b}, float, false>"])
        assert sweep is not None and spmv is not None, prof
        # measured traffic is at least the algorithmic bytes' order of magnitude (C2: 19.3 / 67.2 MB, C3: 91.4 /
        # 560.5 MB per launch), so a unit slip (KiB for MB) or a wrong row is caught
        lo_sweep, lo_spmv = (15e6, 50e6) if name == "c2" else (80e6, 500e6)
        assert lo_sweep < sweep < 3 * lo_sweep and lo_spmv < spmv < 3 * lo_spmv, (name, sweep, spmv)


def test_missing_rows_read_as_none():
    assert bench.measured_traffic("c3_hbm_traffic.md", ["no such kernel"]) is None
    assert bench.measured_traffic("no_such_profile.md", ["sweep_phases_fp32_per_launch"]) is None

"""The bench's CPU baseline runs the reference's own MPI CPU path through the
harness built by oracle/Makefile.ref (skipped where it is not built)."""
import os
import shutil

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


@pytest.mark.skipif(not os.path.exists(HARNESS) or
                    not (shutil.which("mpiexec") or os.path.exists("/opt/conda/bin/mpiexec")),
                    reason="reference harness or mpiexec not available")
def test_cpu_baseline_times_the_reference_path():
    import bench
    r = bench.cpu_baseline_reference(9, (32, 24), 1.0, -5.0, ("noise", 1))
    assert r is not None
    assert r["kind"] == "reference" and r["unit"] == "MLUPS"
    assert r["value"] > 0 and 1 <= r["cores"] <= 16
    assert "mpiexec -np" in r["sample"]

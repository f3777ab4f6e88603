"""The bench's CPU baselines run the reference's own CPU paths through the
harnesses built by oracle/Makefile.ref (skipped where they are not built):
single component under MPI, two components under MPI, and the capsule on
one rank; a "+K" step list runs K steps and writes no dump."""
import os
import shutil

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
MCMP = os.path.join(ROOT, "oracle", "_ref", "ref_harness_mcmp")
CAPSULE = os.path.join(ROOT, "oracle", "_ref", "ref_harness_capsule")
HAVE_MPI = bool(shutil.which("mpiexec") or os.path.exists("/opt/conda/bin/mpiexec"))


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


@pytest.mark.skipif(not os.path.exists(MCMP) or not HAVE_MPI, reason="mcmp harness or mpiexec not available")
def test_cpu_baseline_times_the_reference_two_component_path():
    import bench
    r = bench.cpu_baseline_mcmp("mcmp2d_fp64_4096_noise", ranks=2, seconds=0.5, shape=(40, 24))
    assert r is not None
    assert r["kind"] == "reference" and r["unit"] == "MLUPS" and r["cores"] == 2
    assert r["value"] > 0 and "40x24" in r["sample"]


@pytest.mark.skipif(not os.path.exists(CAPSULE), reason="capsule harness not available")
def test_cpu_baseline_times_the_reference_capsule_path():
    import bench
    r = bench.cpu_baseline_capsule(nx=48, ny=40, nnodes=24, seconds=0.3)
    assert r is not None
    assert r["kind"] == "reference" and r["cores"] == 1
    assert r["us_per_step"] > 0 and r["mlups"] > 0


@pytest.mark.skipif(not os.path.exists(CAPSULE), reason="capsule harness not available")
def test_plus_k_step_list_runs_without_a_dump(tmp_path):
    import subprocess
    import numpy as np
    rho0 = tmp_path / "rho0.bin"
    np.full(48 * 40, 1.0).tofile(rho0)
    base = [CAPSULE, "48", "40", "0.9", "24", "0.8", "20.0", "19.5", "6.0", str(rho0), str(tmp_path / "out")]
    r = subprocess.run(base + ["+5"], capture_output=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert not list(tmp_path.glob("out_*.bin"))
    r = subprocess.run(base + ["5"], capture_output=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert [p.name for p in tmp_path.glob("out_*.bin")] == ["out_5.bin"]

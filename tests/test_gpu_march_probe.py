"""The memory skeleton of the fused D3Q19 march (lbm_probe_march: the step
kernel's loads and stores without its arithmetic), whose rate the bench
reports as the attainable rate of the step kernels' access pattern, runs for
every storage width and layout; invalid sizes are refused before any device
work."""
import pytest

from mesobasic_amd import lbm


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("elem_bytes", [2, 4, 8])
def test_march_probe_runs(elem_bytes, layout):
    ms = lbm.probe_march(128, elem_bytes, layout, reps=2)
    assert 0.0 < ms < 1000.0


@pytest.mark.parametrize("n,eb,layout", [(100, 8, 0), (32, 8, 0), (2048, 8, 0), (128, 3, 0),
                                         (128, 8, 2)])
def test_march_probe_refuses_bad_arguments(n, eb, layout):
    with pytest.raises(lbm.LbmError):
        lbm.probe_march(n, eb, layout)

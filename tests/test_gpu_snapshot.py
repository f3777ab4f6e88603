"""Asynchronous snapshots (lbm_snapshot_begin / lbm_snapshot_wait): the seven
arrays a snapshot hands out equal, bit for bit, what the synchronous getters
return at the step it was begun, although further steps run while it is
copied; and the driver's overlapped VTK output writes the same files as the
synchronous one, each matching the oracle at its step."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import mesobasic_amd as mb
import oracle
from conftest import same_bits

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESO = os.path.join(ROOT, "mesobasic_amd", "host", "meso_amd")


def _sync(lat, comp, kind):
    f = lat.fields(kind, comp=comp)
    fx, fy, fz = lat.force(comp=comp)
    return dict(f, fx=fx, fy=fy, fz=fz)


def _check_snapshot(lat, comps, steps_before, steps_during, kind=mb.FIELDS_REFERENCE):
    lat.step(steps_before)
    for comp in range(comps):
        want = _sync(lat, comp, kind)
        lat.snapshot_begin(comp, kind)
        lat.step(steps_during)          # runs while the copy is in flight
        step, got = lat.snapshot_wait()
        assert step == lat.steps_done - steps_during
        for k in ("rho", "u", "v", "w", "fx", "fy", "fz"):
            assert same_bits(got[k], want[k]), (comp, k)


@pytest.mark.parametrize("q,shape,order", [(19, (8, 10, 64), 3), (9, (40, 36), 2), (9, (24, 20), 3),
                                           (19, (6, 8, 12), 2)])
def test_snapshot_matches_getters_fp64(q, shape, order):
    print(f"snapshot fp64 q={q} shape={shape} order={order}", flush=True)
    lat = mb.Lattice(q, shape, 1.0, -5.5, order=order, precision=mb.FP64)
    lat.set_rho(mb.noise_rho(shape, 3))
    lat.init()
    # at step 0 the reference arrays hold rho0 and zero velocity
    _check_snapshot(lat, 1, 0, 3)
    _check_snapshot(lat, 1, 5, 4)
    _check_snapshot(lat, 1, 1, 1, kind=mb.FIELDS_CURRENT)


@pytest.mark.parametrize("q,shape,film", [(9, (18, 14), None), (19, (6, 5, 8), None),
                                          (9, (20, 12), (0.35, 0.65)), (19, (6, 5, 8), (0.35, 0.65))])
def test_snapshot_matches_getters_two_components(q, shape, film):
    print(f"snapshot mcmp q={q} shape={shape} film={film}", flush=True)
    lat = mb.Lattice(q, shape, 1.0, 3.0, precision=mb.FP64, components=2, tau_b=0.8,
                     walls=film is not None, rho_sol=film if film is not None else (0.5, 0.5))
    rng = np.random.default_rng(2)
    lat.set_rho(0.5 + 0.05 * (rng.random(shape) - 0.5), comp=0)
    lat.set_rho(0.5 + 0.05 * (rng.random(shape) - 0.5), comp=1)
    lat.init()
    _check_snapshot(lat, 2, 0, 2)
    _check_snapshot(lat, 2, 6, 3)


@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16N])
def test_snapshot_matches_getters_low_precision(prec):
    shape = (8, 16, 64)
    print(f"snapshot precision={prec}", flush=True)
    lat = mb.Lattice(19, shape, 1.0, -5.5, precision=prec)
    lat.set_rho(mb.noise_rho(shape, 4))
    lat.init()
    _check_snapshot(lat, 1, 4, 6)


def test_snapshot_state_errors():
    shape = (8, 8, 16)
    lat = mb.Lattice(19, shape, 1.0, -5.5, precision=mb.FP64)
    with pytest.raises(mb.LbmError):
        lat.snapshot_begin()            # before initialisation
    lat.set_rho(mb.noise_rho(shape, 1))
    lat.init()
    with pytest.raises(mb.LbmError):
        lat.snapshot_wait()             # nothing begun
    with pytest.raises(mb.LbmError):
        lat.snapshot_begin(comp=1)      # one component only
    # a second begin replaces the first
    lat.snapshot_begin()
    lat.step(2)
    want = _sync(lat, 0, mb.FIELDS_REFERENCE)
    lat.snapshot_begin()
    step, got = lat.snapshot_wait()
    assert step == 2 and same_bits(got["rho"], want["rho"]) and same_bits(got["fx"], want["fx"])
    with pytest.raises(mb.LbmError):
        lat.snapshot_wait()


def _run_driver(d, text):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "input.dat"), "w") as f:
        f.write(text)
    r = subprocess.run([MESO], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return sorted(os.listdir(os.path.join(d, "vtkoutput")))


@pytest.mark.parametrize("typ,dom", [("scmp3D", (10, 8, 6)), ("scmp2D", (20, 16, 1)), ("mcmp2D", (18, 14, 1)),
                                     ("mcmp3DFilm", (8, 6, 6))])
def test_driver_overlapped_output_equals_synchronous(tmp_path, typ, dom):
    if not os.path.exists(MESO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mesobasic_amd", "host")], check=True)
    nstep = 40
    stencil = "D3Q19" if dom[2] > 1 else "D2Q9"
    text = f"""
[Domain]
  nx = {dom[0]}
  ny = {dom[1]}
  nz = {dom[2]}
[Time]
  nstep = {nstep}
[Output]
  numOutputs = 8
  async = {{a}}
[LBApp]
  app = true
  type = {typ}
  stencil = {stencil}
  G = {'-5.5' if typ.startswith('scmp') else '3.0'}
  rhoAi = {'0.7' if typ.startswith('scmp') else '0.5'}
  rhoAi_noise = {'0.02' if typ.startswith('scmp') else '0.1'}
  seed = 6
"""
    print(f"driver output {typ} {dom}", flush=True)
    a = _run_driver(str(tmp_path / "async"), text.replace("{a}", "1"))
    s = _run_driver(str(tmp_path / "sync"), text.replace("{a}", "0"))
    assert a == s == sorted(f"rhoA_{k}.vtk" for k in range(0, nstep + 1, nstep // 8))
    for name in a:
        assert filecmp.cmp(tmp_path / "async" / "vtkoutput" / name, tmp_path / "sync" / "vtkoutput" / name,
                           shallow=False), name
    if typ == "scmp3D":
        # every output against the oracle's r array at its step
        nx, ny, nz = dom
        o = oracle.Oracle(19, dom, 1.0, -5.5)
        o.set_rho(oracle.noise_rho(dom, 6))
        o.init()
        for k in range(0, nstep + 1, nstep // 8):
            if k:
                o.step(nstep // 8)
            ref = o.fields()["rho"]
            got = open(tmp_path / "async" / "vtkoutput" / f"rhoA_{k}.vtk").read().splitlines()[12:]
            assert got == [f"{ref[i, j, kk]:.3f}" for kk in range(nz) for j in range(ny) for i in range(nx)], k

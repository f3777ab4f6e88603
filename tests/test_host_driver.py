"""The C++ drop-in driver (mesobasic_amd/host): input.dat grammar on the CPU,
and on the GPU a full run whose VTK output matches the oracle's fields to
the reference writer's 3 decimals, in the reference's file format."""
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESO = os.path.join(ROOT, "mesobasic_amd", "host", "meso_amd")


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(MESO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mesobasic_amd", "host")], check=True)


def _params(tmp_path, text):
    p = tmp_path / "input.dat"
    p.write_text(text)
    out = subprocess.run([MESO, "--print-params", str(p)], check=True, capture_output=True,
                         text=True).stdout
    kv = {}
    secs = []
    for line in out.splitlines():
        if line.startswith("section "):
            secs.append(line.split()[1])
        elif " = " in line:
            k, v = line.split(" = ", 1)
            kv[k] = v
    return secs, kv, out


def test_getpot_grammar(tmp_path):
    secs, kv, out = _params(tmp_path, """
# leading comment
[Domain]
  nx = 64   # trailing comment
  ny=32
  nz =16
[Time]
  nstep = 2.0e3
[LBApp]
  app = true
  G = -5.5
  [./inner]
    a = 1
  [../other]
    b = 2
[]
  root = 3
[Domain]
  nx = 65
""")
    assert secs == ["Domain/", "Time/", "LBApp/", "LBApp/inner/", "LBApp/other/"]
    assert kv["Domain/nx"] == "65"          # a later assignment wins
    assert kv["Domain/ny"] == "32"
    # GetPot joins only "a = b" and "a=b": "nz =16" leaves nz unset and
    # assigns "=16" to the section itself (tests/test_getpot_parity.py pins
    # this against GetPot)
    assert "Domain/nz" not in kv and kv["Domain/"] == "16"
    assert kv["LBApp/inner/a"] == "1" and kv["LBApp/other/b"] == "2"
    assert kv["root"] == "3"
    assert "nstep 2000 numOutputs 1" in out  # ints convert through double


def test_unknown_type_is_reported(tmp_path):
    (tmp_path / "input.dat").write_text("[LBApp]\n app = true\n type = capsule3D\n")
    r = subprocess.run([MESO], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0
    assert "capsule3D" in r.stderr and "not provided" in r.stderr


def _read_vtk(path):
    lines = open(path).read().splitlines()
    hdr = lines[:12]
    vals = []
    i = 12
    while i < len(lines) and lines[i].strip() and not lines[i].startswith("VECTORS"):
        vals.append(float(lines[i]))
        i += 1
    return hdr, np.array(vals), lines


@pytest.mark.gpu
def test_scmp3d_run_matches_oracle(tmp_path):
    nx, ny, nz, nstep = 12, 10, 8, 40
    (tmp_path / "input.dat").write_text(f"""
[Domain]
  nx = {nx}
  ny = {ny}
  nz = {nz}
[Time]
  nstep = {nstep}
[Output]
  numOutputs = 2
[LBApp]
  app = true
  type = scmp3D
  G = -5.5
  rhoAi = 0.7
  rhoAi_noise = 0.02
  tauA = 1.0
  seed = 9
""")
    r = subprocess.run([MESO], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    files = sorted(os.listdir(tmp_path / "vtkoutput"))
    assert files == ["rhoA_0.vtk", "rhoA_20.vtk", "rhoA_40.vtk"]
    o = oracle.Oracle(19, (nx, ny, nz), 1.0, -5.5)
    o.set_rho(oracle.noise_rho((nx, ny, nz), 9))
    o.init()
    o.step(nstep)
    ref = o.fields()["rho"]          # the r array the reference would print
    hdr, vals, _ = _read_vtk(tmp_path / "vtkoutput" / "rhoA_40.vtk")
    assert hdr[5] == f"DIMENSIONS   {nx}   {ny}   {nz}"
    assert hdr[9] == f"POINT_DATA {nx * ny * nz}"
    expect = [f"{ref[i, j, k]:.3f}" for k in range(nz) for j in range(ny) for i in range(nx)]
    got = open(tmp_path / "vtkoutput" / "rhoA_40.vtk").read().splitlines()[12:]
    assert got == expect


@pytest.mark.gpu
def test_scmp2d_run_writes_rho_and_velocity(tmp_path):
    nx, ny, nstep = 20, 16, 30
    (tmp_path / "input.dat").write_text(f"""
[Domain]
  nx = {nx}
  ny = {ny}
[Time]
  nstep = {nstep}
[Output]
  numOutputs = 1
[LBApp]
  app = true
  type = scmp2D
  stencil = D2Q9
  G = -5.0
  rhoAi = 0.7
  rhoAi_noise = 0.02
  seed = 5
""")
    r = subprocess.run([MESO], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = oracle.Oracle(9, (nx, ny), 1.0, -5.0)
    o.set_rho(oracle.noise_rho((nx, ny), 5))
    o.init()
    o.step(nstep)
    fl = o.fields()
    fx, fy, _ = o.force()
    lines = open(tmp_path / "vtkoutput" / f"rhoA_{nstep}.vtk").read().splitlines()
    rho_lines = lines[12:12 + nx * ny]
    assert rho_lines == [f"{fl['rho'][i, j]:.3f}" for j in range(ny) for i in range(nx)]
    assert lines[12 + nx * ny + 1] == "VECTORS Velocity float"
    us = fl["u"] + 0.5 * fx / fl["rho"]
    vs = fl["v"] + 0.5 * fy / fl["rho"]
    vel = lines[12 + nx * ny + 2:]
    assert vel == [f"{us[i, j]:.3f} {vs[i, j]:.3f} 0.000" for j in range(ny) for i in range(nx)]


def _mc_rho0(shape, seed, skip_z_walls=False):
    # the drop-in classes' seeded start: one mt19937, per node a draw for A
    # then one for B (the reference's rand() order, mcmp2D.cpp:67-74);
    # mcmp3DFilm draws for the fluid rows only and leaves its wall rows at 0
    mask = np.ones(shape, dtype=bool)
    if skip_z_walls:
        mask[:, :, 0] = mask[:, :, -1] = False
    k = int(mask.sum())
    x = oracle.mt19937(seed, 2 * k).astype(np.float64) / 4294967295.0
    rA, rB = np.zeros(shape), np.zeros(shape)
    rA[mask] = 0.5 + 0.1 * (x[0::2] - 0.5)
    rB[mask] = 0.5 + 0.1 * (x[1::2] - 0.5)
    return rA, rB


@pytest.mark.gpu
@pytest.mark.parametrize("q,film", [(9, False), (19, False), (9, True), (19, True)])
def test_mcmp_run_matches_oracle(tmp_path, q, film):
    shape = (18, 14) if q == 9 else (10, 8, 6)
    nstep = 30
    dom = "".join(f"  n{a} = {v}\n" for a, v in zip("xyz", shape))
    (tmp_path / "input.dat").write_text(f"""
[Domain]
{dom}[Time]
  nstep = {nstep}
[Output]
  numOutputs = 1
[LBApp]
  app = true
  type = {('mcmp2D' if q == 9 else 'mcmp3D') + ('Film' if film else '')}
  G = 3.0
  tauA = 1.0
  tauB = 0.8
  rhoA_sol = 0.35
  rhoB_sol = 0.65
  seed = 4
""")
    r = subprocess.run([MESO], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rA, rB = _mc_rho0(shape, 4, skip_z_walls=film and q == 19)
    o = oracle.MCOracle(q, shape, 1.0, 0.8, 3.0, film=(0.35, 0.65) if film else None)
    o.set_rho(rA, rB)
    o.init()
    o.step(nstep)
    fl = o.fields(0)
    lines = open(tmp_path / "vtkoutput" / f"rhoA_{nstep}.vtk").read().splitlines()
    n = int(np.prod(shape))
    if q == 19:
        nx, ny, nz = shape
        expect = [f"{fl['rho'][i, j, k]:.3f}" for k in range(nz) for j in range(ny) for i in range(nx)]
        # mcmp3DFilm's wall rows hold NaN (density 0 at the start); the sign a
        # NaN prints with is the platform's, not part of the format
        assert [v.replace("-nan", "nan") for v in lines[12:]] == expect
    else:
        nx, ny = shape
        assert lines[12:12 + n] == [f"{fl['rho'][i, j]:.3f}" for j in range(ny) for i in range(nx)]
        us = fl["u"] + 0.5 * fl["fx"] / fl["rho"]
        vs = fl["v"] + 0.5 * fl["fy"] / fl["rho"]
        assert lines[12 + n + 2:] == [f"{us[i, j]:.3f} {vs[i, j]:.3f} 0.000" for j in range(ny)
                                      for i in range(nx)]


@pytest.mark.gpu
def test_capsule2d_run_matches_oracle(tmp_path):
    nx, ny, nstep, nn = 48, 40, 60, 24
    (tmp_path / "input.dat").write_text(f"""
[Domain]
  nx = {nx}
  ny = {ny}
[Time]
  nstep = {nstep}
[Output]
  numOutputs = 2
[LBApp]
  app = true
  type = capsule2D
  tauA = 0.9
  nnodes = {nn}
  kstiff = 0.8
  capsule_x = 20.0
  capsule_y = 19.5
  capsule_r = 6.0
  seed = 5
""")
    r = subprocess.run([MESO], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    files = sorted(os.listdir(tmp_path / "vtkoutput"))
    assert files == sorted(f"{t}_{s}.vtk" for t in ("rhoA", "capsule") for s in (0, 30, 60))
    o = oracle.CapsuleOracle((nx, ny), 0.9, nn, 0.8)
    o.set_rho(oracle.noise_rho((nx, ny), 5, mean=0.5, amp=0.1))
    o.circle(20.0, 19.5, 6.0)
    o.init()
    for s in (0, 30, 60):
        if s:
            o.step(30)
        fl, nd = o.fields(), o.nodes()
        lines = open(tmp_path / "vtkoutput" / f"rhoA_{s}.vtk").read().splitlines()
        n = nx * ny
        assert lines[12:12 + n] == [f"{fl['rho'][i, j]:.3f}" for j in range(ny) for i in range(nx)]
        us = fl["u"] + 0.5 * fl["fx"] / fl["rho"]
        vs = fl["v"] + 0.5 * fl["fy"] / fl["rho"]
        assert lines[12 + n + 2:] == [f"{us[i, j]:.3f} {vs[i, j]:.3f} 0.000" for j in range(ny)
                                      for i in range(nx)]
        cap = open(tmp_path / "vtkoutput" / f"capsule_{s}.vtk").read().splitlines()
        assert cap[:7] == ["# vtk DataFile Version 3.1", "VTK file containing IBM data", "ASCII", " ",
                           "DATASET POLYDATA", " ", f"POINTS {nn} float"]
        assert cap[7:7 + nn] == [f"{x:.3f}  {y:.3f}  0.000" for x, y in zip(nd["x"], nd["y"])]
        assert cap[7 + nn:9 + nn] == [" ", f"LINES {nn} {3 * nn}"]
        assert cap[9 + nn:9 + 2 * nn] == [f"2  {k}  {(k + 1) % nn}" for k in range(nn)]
        assert cap[9 + 2 * nn:11 + 2 * nn] == [" ", f"VERTICES {nn} {2 * nn}"]
        assert cap[11 + 2 * nn:] == [f"1 {k}" for k in range(nn)]

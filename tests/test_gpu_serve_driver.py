"""The serving front end: lbm_batch_read / lbm_batch_extend (a batch job
stopped at its output steps, read, and run on), and `meso_amd --serve`, which
runs many input.dat directories together on the batch engine.

- read returns the reference's r,u,v,w arrays of the job's step order
  (D3Q19: moments of f(N-1); D2Q9: of f(N); rho0 and 0 before any step), bit
  for bit equal to the moments other jobs of the same batch finish with, and
  the force array of the density the last step started from, equal to the
  oracle's FP64 force of that density;
- a job extended after finishing ends with the same bits as a job submitted
  with the total step count;
- each served run writes the files the driver's cadence asks for, within
  the printed precision of the FP64 oracle, and the same bytes whether it is
  served alone or with other runs."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import mesobasic_amd as mb
from mesobasic_amd import lbm
from mesobasic_amd.lbm import Batch
import oracle
from conftest import same_bits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESO = os.path.join(ROOT, "mesobasic_amd", "host", "meso_amd")


def _oracle_force(q, shape, G, rho):
    """The FP64 force of density rho: the oracle's force array after init
    holds F(psi(rho0))."""
    o = oracle.Oracle(q, shape, 1.0, G)
    o.set_rho(np.asarray(rho, dtype=np.float64))
    o.init()
    return o.force()


def _run_until(b, want):
    """Ticks until every job in `want` is finished."""
    while not set(want) <= set(b.finished()):
        b.tick()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16N])
@pytest.mark.parametrize("q,shape", [(9, (40, 36)), (19, (8, 10, 64))])
def test_read_and_extend(prec, q, shape):
    print(f"batch read/extend q={q} shape={shape} precision={prec}", flush=True)
    tau, G = 0.9, -5.5
    rho0 = mb.noise_rho(shape, 11)
    b = Batch(precision=prec, steps_per_tick=4)
    ids = {k: b.submit(q, shape, tau, G, s, rho0) for k, s in
           (("A", 7), ("B", 13), ("C", 6), ("D", 7), ("E", 12), ("F", 1))}
    assert None not in ids.values()
    # before any step: rho0 (as the engine stores it, FP32) and zero velocity
    r0 = b.read(ids["A"])
    r32 = rho0.astype(np.float32).astype(np.float64)
    assert same_bits(r0["rho"], r32)
    assert not np.any(r0["u"]) and not np.any(r0["v"]) and not np.any(r0["w"])
    F0 = _oracle_force(q, shape, G, r32)
    for k, Fk in zip(("fx", "fy", "fz"), F0):
        assert same_bits(r0[k], Fk), k
    with pytest.raises(mb.LbmError):
        b.extend(ids["A"], 3)                 # still running
    _run_until(b, [ids["A"], ids["F"]])
    a7 = b.read(ids["A"])
    assert b.status(ids["A"])["steps_done"] == 7
    b.extend(ids["A"], 6)
    assert b.status(ids["A"])["state"] == lbm.JOB_RUNNING
    f1 = b.read(ids["F"])
    _run_until(b, list(ids.values()))
    a13 = b.read(ids["A"])
    fin = {k: b.collect(i) for k, i in ids.items()}
    # an extended job equals one submitted with the total step count
    for k in ("rho", "u", "v", "w"):
        assert same_bits(fin["A"][k], fin["B"][k]), k
    # the reference arrays of the step order
    ref7, ref13 = (fin["D"], fin["B"]) if q == 9 else (fin["C"], fin["E"])
    for k in ("rho", "u", "v", "w"):
        assert same_bits(a7[k], ref7[k]), k
        assert same_bits(a13[k], ref13[k]), k
    # the force of the density the last step started from
    src7 = fin["C"]["rho"] if q == 9 else a7["rho"]
    src13 = fin["E"]["rho"] if q == 9 else a13["rho"]
    for got, src in ((a7, src7), (a13, src13)):
        F = _oracle_force(q, shape, G, src)
        for k, Fk in zip(("fx", "fy", "fz"), F):
            assert same_bits(got[k], Fk), k
    if q == 9:
        # LBfluid2D order: the first step's force is that of rho0
        F = _oracle_force(q, shape, G, r32)
        assert same_bits(f1["fx"], F[0]) and same_bits(f1["fy"], F[1])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S])
def test_read_after_temporal_blocks(prec):
    # D2Q9 jobs advanced in temporal blocks (3 steps per launch): a job read
    # at its stop still sees f(N-1) in its other buffer (the last step before
    # a stop is a single one), so its force array is that of rho(N-1)
    shape, tau, G = (40, 36), 0.9, -5.5
    rho0 = mb.noise_rho(shape, 12)
    b = Batch(precision=prec, steps_per_tick=6)
    b.set_launch(32, 2, 1)
    ja, jb = b.submit(9, shape, tau, G, 11, rho0), b.submit(9, shape, tau, G, 10, rho0)
    _run_until(b, [ja, jb])
    got = b.read(ja)
    prev = b.collect(jb)
    ref = b.collect(ja)
    for k in ("rho", "u", "v"):
        assert same_bits(got[k], ref[k]), k
    F = _oracle_force(9, shape, G, prev["rho"])
    assert same_bits(got["fx"], F[0]) and same_bits(got["fy"], F[1])


@pytest.mark.gpu
def test_extend_refuses_failed_jobs():
    shape = (24, 20)
    rho0 = mb.noise_rho(shape, 3)
    rho0[4, 5] = np.nan                       # a poisoned start: retired at its first tick
    b = Batch(precision=mb.FP32, steps_per_tick=4)
    jid = b.submit(9, shape, 1.0, -5.0, 8, rho0)
    _run_until(b, [jid])
    assert b.status(jid)["state"] == lbm.JOB_FAILED
    with pytest.raises(mb.LbmError):
        b.extend(jid, 4)
    with pytest.raises(mb.LbmError):
        b.extend(12345, 4)                    # unknown job


RUNS = {
    "cube": dict(type="scmp3D", dom=(12, 10, 64), nstep=40, outs=4, G=-5.5, tau=1.0, seed=3),
    "sheet": dict(type="scmp2D", dom=(48, 40, 1), nstep=30, outs=3, G=-5.0, tau=0.8, seed=8),
    "odd": dict(type="scmp3D", dom=(9, 11, 13), nstep=25, outs=2, G=-5.0, tau=1.2, seed=5),
    "many": dict(type="scmp2D", dom=(30, 22, 1), nstep=12, outs=20, G=-5.5, tau=1.0, seed=2),
}


def _write_run(d, r):
    os.makedirs(d, exist_ok=True)
    nx, ny, nz = r["dom"]
    with open(os.path.join(d, "input.dat"), "w") as f:
        f.write(f"""
[Domain]
  nx = {nx}
  ny = {ny}
  nz = {nz}
[Time]
  nstep = {r['nstep']}
[Output]
  numOutputs = {r['outs']}
[LBApp]
  app = true
  type = {r['type']}
  G = {r['G']}
  tauA = {r['tau']}
  rhoAi = 0.7
  rhoAi_noise = 0.02
  seed = {r['seed']}
""")


def _expected_steps(r):
    interval = r["nstep"] // r["outs"] if r["outs"] else r["nstep"] + 1
    interval = max(interval, 1)
    return [0] + [s for s in range(1, r["nstep"] + 1) if s % interval == 0]


def _vtk_values(path):
    out = []
    for line in open(path).read().splitlines()[12:]:
        if not line or line.startswith("VECTORS"):
            continue
        out.extend(float(t) for t in line.split())
    return np.array(out)


def _oracle_values(name, r, step):
    nx, ny, nz = r["dom"]
    q = 19 if r["type"] == "scmp3D" else 9
    shape = (nx, ny, nz) if q == 19 else (nx, ny)
    o = oracle.Oracle(q, shape, r["tau"], r["G"])
    o.set_rho(oracle.noise_rho(shape, r["seed"]))
    o.init()
    if step:
        o.step(step)
    fl = o.fields()
    if q == 19:
        return np.array([fl["rho"][i, j, k] for k in range(nz) for j in range(ny) for i in range(nx)])
    fx, fy, _ = o.force()
    us = fl["u"] + 0.5 * fx / fl["rho"]
    vs = fl["v"] + 0.5 * fy / fl["rho"]
    rho = [fl["rho"][i, j] for j in range(ny) for i in range(nx)]
    vel = [v for j in range(ny) for i in range(nx) for v in (us[i, j], vs[i, j], 0.0)]
    return np.array(rho + vel)


def _serve(dirs, *opts):
    r = subprocess.run([MESO, "--serve", *opts, *dirs], capture_output=True, text=True, timeout=180)
    return r


@pytest.mark.gpu
def test_served_runs_match_oracle_and_do_not_depend_on_the_mix(tmp_path):
    if not os.path.exists(MESO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mesobasic_amd", "host")], check=True)
    together = []
    for name, r in RUNS.items():
        _write_run(str(tmp_path / "together" / name), r)
        together.append(str(tmp_path / "together" / name))
    # a run the engine does not serve is reported, and the others still run
    bad = tmp_path / "together" / "two_component"
    bad.mkdir()
    (bad / "input.dat").write_text("[LBApp]\n app = true\n type = mcmp2D\n")
    print("serve: four runs together", flush=True)
    res = _serve(together + [str(bad)], "--precision", "fp32", "--steps-per-tick", "6")
    assert res.returncode == 1, res.stderr
    assert "two_component: failed: LBApp/type mcmp2D" in res.stdout
    assert "5 runs (1 failed)" in res.stderr
    for name, r in RUNS.items():
        files = sorted(os.listdir(tmp_path / "together" / name / "vtkoutput"))
        assert files == sorted(f"rhoA_{s}.vtk" for s in _expected_steps(r)), name
        for s in _expected_steps(r):
            got = _vtk_values(tmp_path / "together" / name / "vtkoutput" / f"rhoA_{s}.vtk")
            want = _oracle_values(name, r, s)
            # printed with 3 decimals; FP32 arithmetic moves a value across a
            # rounding boundary at most
            assert got.shape == want.shape
            assert np.max(np.abs(got - want)) <= 1.0e-3 + 1e-9, (name, s)
    for name, r in RUNS.items():
        print(f"serve: {name} alone", flush=True)
        d = str(tmp_path / "alone" / name)
        _write_run(d, r)
        res = _serve([d], "--precision", "fp32", "--steps-per-tick", "4")
        assert res.returncode == 0, res.stderr
        for s in _expected_steps(r):
            f = f"rhoA_{s}.vtk"
            assert filecmp.cmp(os.path.join(d, "vtkoutput", f), tmp_path / "together" / name / "vtkoutput" / f,
                               shallow=False), (name, f)


def test_serve_rejects_what_the_engine_does_not_run(tmp_path):
    """Input checks of the serving front end (no GPU is touched when no run
    is left)."""
    if not os.path.exists(MESO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mesobasic_amd", "host")], check=True)
    cases = {
        "mcmp": ("[LBApp]\n app = true\n type = mcmp3D\n", "not served by the batch engine"),
        "walls": ("[LBApp]\n app = true\n type = scmp3D\n walls = 1\n", "periodic lattices"),
        "order": ("[LBApp]\n app = true\n type = scmp2D\n order = 3\n", "LBApp/order"),
        "noapp": ("[LBApp]\n type = scmp3D\n", "no [LBApp] section"),
        "stencil": ("[LBApp]\n app = true\n type = scmp3D\n stencil = D2Q9\n", "does not match"),
    }
    dirs = []
    for name, (text, _) in cases.items():
        d = tmp_path / name
        d.mkdir()
        (d / "input.dat").write_text(text)
        dirs.append(str(d))
    dirs.append(str(tmp_path / "missing"))
    res = _serve(dirs)
    assert res.returncode == 1
    for name, (_, msg) in cases.items():
        line = [ln for ln in res.stdout.splitlines() if ln.startswith(str(tmp_path / name) + ":")]
        assert line and msg in line[0], (name, res.stdout)
    assert "cannot read input file" in res.stdout
    assert "no run to serve (6 failed)" in res.stderr
    res = _serve([], "--precision", "fp32")
    assert res.returncode == 1 and "at least one run directory" in res.stderr
    res = _serve(dirs[:1], "--precision", "fp8")
    assert res.returncode == 1 and "--precision must be" in res.stderr

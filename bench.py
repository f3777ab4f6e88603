"""Benchmark of the Shan-Chen lattice-Boltzmann step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one time step of the whole lattice (force + collision + streaming)
with the populations already resident in HBM.  Default workload: D3Q19
256^3 periodic, FP64, tau = 1, G = -5.5, droplet R = 48 (the drop-in FP64
configuration).  Prints ONE JSON line with MLUPS, the HBM roofline of the
step kernel (algorithmic bytes 2*Q*b per lattice update over the kernel's
measured time), the same roofline block for the D3Q19 512^3 FP64 step, and a
CPU baseline timed on this host's cores on a bounded sample: the reference's
own MPI CPU path, its lattice-Boltzmann sources compiled in place into
oracle/_ref/ref_harness.  That harness is built in the container that holds
the reference tree (`python -c "import __graft_entry__ as g; g.build()"`)
and travels with the working tree; where it is missing the bench says so
on stderr and in the line, and times the oracle (a restatement of the
reference's CPU code) instead.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mesobasic_amd as mb  # noqa: E402

METRIC = "D3Q19 Shan-Chen MLUPS & % HBM roofline; batched-serve p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# nodes advanced per step by the serving loop (least remaining work first,
# ranked over both engines) and steps per tick; measured on MI355X
# (200-request stream, the forms below, engines' ticks overlapped; p50 / p99
# s, GLUPS; one box): 32 M x 20 steps 3.7 / 17.9 / 26.1, 32 M x 40 3.6 /
# 17.4 / 26.8, 32 M x 80 3.5 / 17.1 / 27.3, 24 M x 60 2.9 / 17.9 / 26.9,
# 20 M x 80 2.45 / 17.4 / 26.8, 16 M x 120 2.1 / 17.1 / 26.5, 12 M x 80 1.9 /
# 17.8 / 25.5.  Longer ticks cost fewer synchronisations and mass checks;
# a smaller limit lets short requests through sooner.
SERVE_WORK_LIMIT = 16_000_000
SERVE_STEPS_PER_TICK = 120
# storage form by stencil: FP16 with the node shift is closer to FP64 than
# FP32 storage on D3Q19 droplets, but not on D2Q9 ones (up to 3e-3 against
# 2e-4), so D2Q9 requests are served in FP32 storage; measured on MI355X
# (200-request stream, 32 M nodes): this mix p50 4.3 s, p99 18.7 s, 25.7
# GLUPS, largest droplet error 1.8e-4; all FP16 node shift 3.5 / 17.7 / 27.1,
# 3.0e-3
SERVE_FORMS = {9: mb.FP32, 19: mb.FP16N}

WORKLOADS = {
    # name: q, shape, precision, tau, G, init
    "d3q19_fp64_256_droplet": (19, (256, 256, 256), mb.FP64, 1.0, -5.5, ("droplet", 48.0)),
    "d3q19_fp64_512_noise": (19, (512, 512, 512), mb.FP64, 1.0, -5.0, ("noise", 1)),
    "d3q19_fp32_512_noise": (19, (512, 512, 512), mb.FP32, 1.0, -5.0, ("noise", 1)),
    "d3q19_fp16_512_noise": (19, (512, 512, 512), mb.FP16S, 1.0, -5.0, ("noise", 1)),
    "d3q19_bf16_512_noise": (19, (512, 512, 512), mb.BF16S, 1.0, -5.0, ("noise", 1)),
    "d3q19_fp16n_512_noise": (19, (512, 512, 512), mb.FP16N, 1.0, -5.0, ("noise", 1)),
    "d3q19_bf16n_512_noise": (19, (512, 512, 512), mb.BF16N, 1.0, -5.0, ("noise", 1)),
    "d2q9_fp64_256_noise": (9, (256, 256), mb.FP64, 1.0, -5.0, ("noise", 1)),
    "d2q9_fp64_4096_noise": (9, (4096, 4096), mb.FP64, 1.0, -5.0, ("noise", 1)),
}
BYTES_PER_POP = {mb.FP64: 8, mb.FP32: 4, mb.FP16S: 2, mb.BF16S: 2, mb.FP16N: 2, mb.BF16N: 2}
DTYPE = {mb.FP64: "f64", mb.FP32: "f32", mb.FP16S: "f32 compute / f16 storage",
         mb.BF16S: "f32 compute / bf16 storage",
         mb.FP16N: "f32 compute / f16 node-shifted storage",
         mb.BF16N: "f32 compute / bf16 node-shifted storage"}
KERNEL_NAMES = {mb.KERNEL_TWO_PASS: "two_pass", mb.KERNEL_FUSED: "fused"}


def initial_rho(shape, init):
    if init[0] == "droplet":
        return mb.droplet_rho(shape, init[1])
    return mb.noise_rho(shape, init[1])


def cpu_baseline_reference(q, shape, tau, G, init, ranks=None, seconds=10.0):
    """Times the reference's own MPI CPU path on this host: the harness
    oracle/_ref/ref_harness (the reference's LBfluid3D/LBfluid2D/scmp3D
    translation units compiled in place, x split into slabs over MPI ranks as
    LBApp does) under mpiexec, one rank per core (up to 16).  Two runs of
    k1 and k2 steps on the same lattice, each ending in the same field dump;
    the rate is n (k2 - k1) / (t2 - t1), so start-up, initialisation and
    the dump cancel.  ranks: MPI ranks (default one per core, up to 16);
    seconds: about how long the timed difference runs.  None when the harness
    or mpiexec is missing or fails."""
    import shutil
    import subprocess
    import tempfile
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(harness) and os.path.exists(mpiexec)):
        return None
    if ranks is None:
        ranks = min(len(os.sched_getaffinity(0)), 16)
    n = int(np.prod(shape))
    full = tuple(shape) + ((1,) if len(shape) == 2 else ())
    k1 = 2
    rate = 5e6 * ranks   # lattice updates per second, roughly, at this rank count
    k2 = k1 + max(20, min(20000, int(seconds * rate / n)))
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        rho0 = os.path.join(d, "rho0.bin")
        initial_rho(shape, init).astype(np.float64).ravel().tofile(rho0)
        times = []
        for k in (k1, k2):
            cmd = [mpiexec, "-np", str(ranks), harness, str(q), *map(str, full), repr(tau), repr(G),
                   "0", rho0, os.path.join(d, f"out{k}"), str(k)]
            t0 = time.perf_counter()
            try:
                # each run takes about 3-12 s; a stuck launcher must not hold the bench
                r = subprocess.run(cmd, capture_output=True, timeout=120)
            except (OSError, subprocess.TimeoutExpired):
                return None
            times.append(time.perf_counter() - t0)
            if r.returncode != 0:
                return None
            for f in os.listdir(d):
                if f.startswith("out"):
                    os.remove(os.path.join(d, f))
    dt = times[1] - times[0]
    if dt <= 0:
        return None
    return {"value": round(n * (k2 - k1) / dt / 1e6, 3), "unit": "MLUPS", "cores": ranks,
            "kind": "reference",
            "sample": f"reference CPU path (its LB sources compiled in place, -O3, MPI x slabs), "
                      f"mpiexec -np {ranks}, {'x'.join(map(str, shape))} lattice, {k2} minus {k1} "
                      f"steps ({dt:.1f} s)"}


def _time_harness(make_cmd, n, k1, k2, timeout=180):
    """Seconds of k2 - k1 steps of a reference harness: two runs (k1 and
    k2 steps, the same start-up and final dump) differenced; None when a run
    fails.  make_cmd(k, outdir) gives the command line of a k-step run."""
    import subprocess
    import tempfile
    times = []
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for k in (k1, k2):
            t0 = time.perf_counter()
            try:
                r = subprocess.run(make_cmd(k, d), capture_output=True, timeout=timeout)
            except (OSError, subprocess.TimeoutExpired):
                return None
            times.append(time.perf_counter() - t0)
            if r.returncode != 0:
                return None
    dt = times[1] - times[0]
    return dt if dt > 0 else None


def cpu_baseline_mcmp(name, ranks=None, seconds=10.0, shape=None):
    """The reference's own two-component CPU path (its mcmp2D / mcmp3D
    classes, compiled in place into oracle/_ref/ref_harness_mcmp, x slabs over
    MPI ranks) on the same workload and initial densities as time_mcmp
    (shape: another lattice of the same stencil)."""
    import shutil
    import tempfile
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness_mcmp")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(harness) and os.path.exists(mpiexec)):
        return None
    q, wshape, ta, tb, G = MC_WORKLOADS[name]
    shape = tuple(shape) if shape is not None else wshape
    ranks = ranks or min(len(os.sched_getaffinity(0)), 16)
    n = int(np.prod(shape))
    full = tuple(shape) + ((1,) if len(shape) == 2 else ())
    k1 = 2
    k2 = k1 + max(10, min(20000, int(seconds * 2e6 * ranks / n)))
    with tempfile.TemporaryDirectory(dir="/tmp") as d0:
        ra, rb = os.path.join(d0, "rA.bin"), os.path.join(d0, "rB.bin")
        mb.noise_rho(shape, 31, mean=0.5, amp=0.1).astype(np.float64).ravel().tofile(ra)
        mb.noise_rho(shape, 32, mean=0.5, amp=0.1).astype(np.float64).ravel().tofile(rb)
        dt = _time_harness(lambda k, d: [mpiexec, "-np", str(ranks), harness, str(q), *map(str, full),
                                         repr(ta), repr(tb), repr(G), ra, rb, os.path.join(d, "out"),
                                         f"+{k}"], n, k1, k2)
    if dt is None:
        return None
    return {"value": round(n * (k2 - k1) / dt / 1e6, 3), "unit": "MLUPS", "cores": ranks,
            "kind": "reference",
            "sample": f"reference two-component CPU path (its LB sources compiled in place, -O3, MPI x "
                      f"slabs), mpiexec -np {ranks}, {'x'.join(map(str, shape))}, {k2} minus {k1} steps "
                      f"({dt:.1f} s)"}


def cpu_baseline_capsule(nx=256, ny=128, nnodes=64, seconds=5.0):
    """The reference's own capsule2D CPU path (capsule2D / IBcapsules2D
    compiled in place into oracle/_ref/ref_harness_capsule, one rank) on the
    lattice of the driver's capsule measurement (tau 0.9, k_stiff 0.8, circle
    of radius 16 at (64, 63.5)); microseconds per step."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness_capsule")
    if not os.path.exists(harness):
        return None
    import tempfile
    n = nx * ny
    k1 = 2
    k2 = k1 + max(100, int(seconds * 1.7e7 / n))   # about 17 M updates per second on one core
    with tempfile.TemporaryDirectory(dir="/tmp") as d0:
        r0 = os.path.join(d0, "rho0.bin")
        np.full(n, 1.0, dtype=np.float64).tofile(r0)
        dt = _time_harness(lambda k, d: [harness, str(nx), str(ny), "0.9", str(nnodes), "0.8", "64.0",
                                         "63.5", "16.0", r0, os.path.join(d, "out"), f"+{k}"], n, k1, k2)
    if dt is None:
        return None
    return {"us_per_step": round(dt / (k2 - k1) * 1e6, 2), "mlups": round(n * (k2 - k1) / dt / 1e6, 3),
            "cores": 1, "kind": "reference",
            "sample": f"reference capsule2D CPU path (compiled in place, -O3, one rank), {nx}x{ny}, "
                      f"{nnodes} IB nodes, uniform density 1.0, {k2} minus {k1} steps ({dt:.1f} s)"}


def cpu_baseline(q, shape, tau, G, init, budget_s=15.0):
    """The reference's CPU path where its harness is built
    (cpu_baseline_reference), else the oracle (CPU restatement of the
    reference step, x-slab threads) on a bounded sample: the same lattice,
    as many steps as fit in about budget_s seconds after one warm-up step."""
    ref = cpu_baseline_reference(q, shape, tau, G, init)
    if ref is not None:
        return ref
    print("# ERROR: the reference CPU harness oracle/_ref/ref_harness is missing or failed to run; "
          "build it where the reference tree is (__graft_entry__.build()). Timing the oracle port "
          "instead.", file=sys.stderr, flush=True)
    import oracle
    cores = min(os.cpu_count() or 1, 16)
    o = oracle.Oracle(q, shape, tau, G, nthreads=cores)
    o.set_rho(initial_rho(shape, init))
    o.init()
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter()
    per = max(t1 - t0, 1e-6)
    k = max(1, min(200, int(budget_s / per)))
    t0 = time.perf_counter()
    o.step(k)
    dt = time.perf_counter() - t0
    n = int(np.prod(shape))
    o.close()
    return {"value": round(n * k / dt / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "reference_harness": "missing or failed (see stderr)",
            "sample": f"oracle (CPU restatement of the reference step, AoS FP64, {cores} x-slab "
                      f"threads), {'x'.join(map(str, shape))} lattice, {k} timed steps after 1 warm-up"}


def time_workload(name, steps=10, warmup=2):
    """MLUPS of one workload (fused kernel, library defaults)."""
    q, shape, prec, tau, G, init = WORKLOADS[name]
    lat = mb.Lattice(q, shape, tau, G, precision=prec)
    lat.set_rho(initial_rho(shape, init))
    lat.init()
    lat.step_timed(warmup)
    ms = lat.step_timed(steps) / steps
    n = int(np.prod(shape))
    bpl = 2 * q * BYTES_PER_POP[prec]
    out = {"mlups": round(n / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms, 4),
           "hbm_frac": round(bpl * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "kernel": KERNEL_NAMES.get(lat.kernel)}
    lat.close()
    return out


def load_traffic(kernel_name, workload):
    """HBM bytes per launch of the step kernel on `workload`, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in
    profiles/traffic_<kernel>_<workload>.json (counters cannot be read inside
    this process), and that file's name; (None, None) where no such summary
    exists."""
    rel = os.path.join("profiles", f"traffic_{kernel_name}_{workload}.json")
    path = os.path.join(ROOT, rel)
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch"), rel
    return None, None


def roofline_block(q, prec, n, ms_per_step, kname, workload):
    """The roofline object of a step kernel: algorithmic bytes (2 Q b per
    lattice update) over the measured time per step."""
    bpl = 2 * q * BYTES_PER_POP[prec]
    achieved = bpl * n / (ms_per_step * 1e-3) / 1e9
    traffic, src = load_traffic(kname, workload)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": (f"{src}: rocprofv3 FETCH_SIZE (x2.0) + WRITE_SIZE per launch of this "
                               f"kernel on {workload}" if src else None),
            "algorithmic_bytes_per_lu": bpl, "lattice_updates_per_launch": n,
            "kernel": f"step ({kname}); HIP-event time per step over the timed region"}


# Two-component Shan-Chen (mcmp3D / mcmp2D), FP64: name -> q, shape, tau_a,
# tau_b, G.  Both components start as seeded noise around 0.5.
MC_WORKLOADS = {
    "mcmp3d_fp64_256_noise": (19, (256, 256, 256), 1.0, 0.8, 3.0),
    "mcmp2d_fp64_4096_noise": (9, (4096, 4096), 1.0, 0.8, 3.0),
}


def time_mcmp(name, steps=10, warmup=2):
    """MLUPS of a two-component workload; algorithmic bytes per update are
    both components' populations read and written once, 2 x 2 x Q x 8 B."""
    q, shape, ta, tb, G = MC_WORKLOADS[name]
    lat = mb.Lattice(q, shape, ta, G, precision=mb.FP64, components=2, tau_b=tb)
    lat.set_rho(mb.noise_rho(shape, 31, mean=0.5, amp=0.1), comp=0)
    lat.set_rho(mb.noise_rho(shape, 32, mean=0.5, amp=0.1), comp=1)
    lat.init()
    lat.step_timed(warmup)
    ms = lat.step_timed(steps) / steps
    n = int(np.prod(shape))
    bpl = 2 * 2 * q * 8
    out = {"mlups": round(n / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms, 4),
           "algorithmic_bytes_per_lu": bpl,
           "hbm_frac": round(bpl * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "kernel": KERNEL_NAMES.get(lat.kernel)}
    lat.close()
    return out


def run_driver(text, timeout=900):
    """Runs the C++ drop-in driver (mesobasic_amd/host/meso_amd) on an
    input.dat of the given text in a scratch directory; returns (seconds,
    MLUPS including output) from its timing line."""
    import re
    import subprocess
    import tempfile
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mesobasic_amd", "host", "meso_amd")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (make -C mesobasic_amd/host)")
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "input.dat"), "w") as f:
            f.write(text)
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"meso_amd failed: {r.stderr[-2000:]}")
        m = re.search(r"(\d+) steps in ([0-9.]+) s \(([0-9.]+) MLUPS", r.stderr)
        return float(m.group(2)), float(m.group(3))


def driver_measurements():
    """The drop-in driver end to end: scmp3D 192^3 for 3,000 steps with no
    output, with 6 VTK outputs written synchronously (Output/async = 0, the
    reference's behaviour) and with asynchronous snapshots (the default); and
    capsule2D (256 x 128, 64 IB nodes) against scmp2D on the same lattice,
    10,000 steps without output."""
    scmp = """[Domain]
  nx = 192
  ny = 192
  nz = 192
[Time]
  nstep = 3000
[Output]
  numOutputs = {nout}
  async = {asy}
[LBApp]
  app = true
  type = scmp3D
  G = -5.5
  rhoAi = 0.7
  rhoAi_noise = 0.02
  tauA = 1.0
  seed = 9
"""
    out = {"scmp3d_192_3000_steps": {}}
    for label, nout, asy in (("no_output", 0, 1), ("sync_6_outputs", 6, 0), ("async_6_outputs", 6, 1)):
        secs, mlups = run_driver(scmp.format(nout=nout, asy=asy))
        out["scmp3d_192_3000_steps"][label] = {"s": round(secs, 3), "mlups_incl_output": round(mlups, 1)}
    lat2d = """[Domain]
  nx = 256
  ny = 128
[Time]
  nstep = 10000
[Output]
  numOutputs = 0
[LBApp]
  app = true
  type = {typ}
  tauA = 0.9
  G = -5.0
  rhoAi = 0.7
  rhoAi_noise = 0.01
  nnodes = 64
  kstiff = 0.8
  capsule_x = 64.0
  capsule_y = 63.5
  capsule_r = 16.0
  seed = 5
"""
    for typ in ("capsule2D", "scmp2D"):
        secs, mlups = run_driver(lat2d.format(typ=typ))
        out[f"{typ}_256x128_10000_steps"] = {"s": round(secs, 3), "us_per_step": round(secs / 1e4 * 1e6, 2),
                                            "mlups": round(mlups, 1)}
    return out


def copy_calibration(nbytes=4 << 30):
    """Attainable HBM bandwidth on this box: the library's fastest streaming
    copy kernel and the HIP runtime's device-to-device copy, both over
    nbytes, read + write bytes per second."""
    ms_k = mb.lbm.probe_copy(nbytes, reps=5, mode=4)
    ms_m = mb.lbm.probe_memcpy(nbytes, reps=5)
    # the fused march's own access pattern without its arithmetic (19 pulls
    # and 19 stores per node, 512^3, the population-plane layout), per
    # storage width: the attainable rate of the step kernels' traffic
    skel = {}
    for name, eb in (("fp64", 8), ("fp32", 4), ("fp16", 2)):
        ms = mb.lbm.probe_march(512, eb, 0, reps=10)
        gbs = 2 * 19 * eb * 512 ** 3 / (ms * 1e-3) / 1e9
        skel[name] = {"GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "mlups": round(512 ** 3 / (ms * 1e-3) / 1e6, 1)}
    return {"bytes": nbytes,
            "copy_kernel_GBs": round(2 * nbytes / (ms_k * 1e-3) / 1e9, 1),
            "hipMemcpyDtoD_GBs": round(2 * nbytes / (ms_m * 1e-3) / 1e9, 1),
            "march_skeleton_512": skel}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="d3q19_fp64_256_droplet", choices=sorted(WORKLOADS))
    ap.add_argument("--kernel", default="auto", choices=["auto", "two_pass", "fused"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning", default="", metavar="SHAPE,CHUNK[,ORDER[,NTSTORE]]",
                    help="fused-kernel launch configuration (default: the library's choice)")
    ap.add_argument("--no-precision-forms", action="store_true",
                    help="skip the 512^3 FP32/FP16/BF16 side measurements")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched D2Q9 (512 lattices) side measurement")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the Poisson-arrival serving side measurement")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests in the serving stream (Poisson, 200/s)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the streaming copy kernel over MIB MiB (for counter scaling)")
    ap.add_argument("--no-driver", action="store_true",
                    help="skip the C++ drop-in driver side measurements (output path, capsule)")
    ap.add_argument("--driver", action="store_true",
                    help="only run the C++ drop-in driver measurements (output path, capsule), one JSON line")
    ap.add_argument("--mcmp", action="store_true",
                    help="only time the two-component workloads, one JSON line")
    ap.add_argument("--march-skeleton", action="store_true",
                    help="only run the fused march's memory skeleton (256^3 and 512^3, 8/4/2-byte "
                         "values, both layouts), one JSON line each")
    args = ap.parse_args()
    if args.driver:
        res = driver_measurements()
        if not args.no_cpu_baseline:
            res["cpu_baseline_capsule2D_256x128"] = cpu_baseline_capsule()
        print(json.dumps(res), flush=True)
        return
    if args.mcmp:
        res = {k: time_mcmp(k, steps=args.steps) for k in MC_WORKLOADS}
        if not args.no_cpu_baseline:
            res["cpu_baseline_mcmp3d_fp64_256_noise"] = cpu_baseline_mcmp("mcmp3d_fp64_256_noise")
        print(json.dumps(res), flush=True)
        return
    if args.march_skeleton:
        for n in (256, 512):
            for eb in (8, 4, 2):
                for layout in (0, 1):
                    ms = mb.lbm.probe_march(n, eb, layout, reps=10)
                    gbs = 2 * 19 * eb * n ** 3 / (ms * 1e-3) / 1e9
                    print(json.dumps({"n": n, "elem_bytes": eb, "layout": layout, "ms": round(ms, 4),
                                      "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                                      "mlups": round(n ** 3 / (ms * 1e-3) / 1e6, 1)}), flush=True)
        return
    if args.copy_calibration:
        nbytes = args.copy_calibration * 2**20
        res = {"copy_bytes": nbytes}
        for mode, name in ((4, "copy_1x16"), (0, "copy_4x16"), (1, "copy_8x16"), (2, "copy_8x16_nt"),
                           (3, "read_only")):
            ms = mb.lbm.probe_copy(nbytes, reps=3, mode=mode)
            moved = nbytes * (1 if mode == 3 else 2)
            res[name] = {"ms": round(ms, 4), "GBs": round(moved / (ms * 1e-3) / 1e9, 1)}
        print(json.dumps(res), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the lattice "
              "would shard into x slabs with a one-plane exchange per face and step).",
              file=sys.stderr)
        sys.exit(2)

    q, shape, prec, tau, G, init = WORKLOADS[args.workload]
    kernel = {"auto": mb.KERNEL_AUTO, "two_pass": mb.KERNEL_TWO_PASS,
              "fused": mb.KERNEL_FUSED}[args.kernel]
    n = int(np.prod(shape))
    lat = mb.Lattice(q, shape, tau, G, precision=prec, kernel=kernel)
    if args.tuning:
        lat.set_tuning(*(int(v) for v in args.tuning.split(",")))
    lat.set_rho(initial_rho(shape, init))
    lat.init()
    print(f"# workload {args.workload} kernel {KERNEL_NAMES.get(lat.kernel)} "
          f"device bytes {lat.device_bytes / 2**30:.2f} GiB", file=sys.stderr, flush=True)
    lat.step_timed(args.warmup)
    lat.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    t0 = time.perf_counter()
    ms = lat.step_timed(args.steps)     # HIP events on the kernels' stream
    wall = time.perf_counter() - t0
    lat.synchronize()
    ms_per_step = ms / args.steps
    mlups = n / (ms_per_step * 1e-3) / 1e6
    cons = lat.conservation()
    kname = KERNEL_NAMES.get(lat.kernel, "unknown")
    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[prec],
        "data": "synthetic (seeded initial density; no dataset)",
        "config": {"workload": args.workload, "stencil": "D3Q19" if q == 19 else "D2Q9",
                   "lattice": list(shape), "tau": tau, "G": G, "init": list(init),
                   "boundary": "periodic", "kernel": kname,
                   "tuning": list(lat.tuning) if lat.kernel == mb.KERNEL_FUSED else None,
                   "wall_s_timed_region": round(wall, 4)},
        "roofline": roofline_block(q, prec, n, ms_per_step, kname, args.workload),
        "mass_after": cons[0],
    }
    lat.close()
    if not args.no_precision_forms and args.workload == "d3q19_fp64_256_droplet":
        # 30 timed steps after 5 warm-up ones: single launches at 512^3 vary
        # by up to 25%, so ten steps leave the form ratios noisy
        forms = {k: time_workload(f"d3q19_{k}_512_noise", steps=30, warmup=5)
                 for k in ("fp64", "fp32", "fp16", "bf16", "fp16n", "bf16n")}
        for k in ("fp16", "fp16n"):
            forms[f"speedup_{k}_vs_fp32"] = round(forms[k]["mlups"] / forms["fp32"]["mlups"], 3)
            forms[f"speedup_{k}_vs_fp64"] = round(forms[k]["mlups"] / forms["fp64"]["mlups"], 3)
        out["precision_forms_512"] = forms
        # the north-star step (D3Q19 512^3 FP64) with its own roofline block
        f64 = forms["fp64"]
        out["roofline_d3q19_fp64_512"] = dict(
            roofline_block(19, mb.FP64, 512 ** 3, f64["ms_per_step"], f64["kernel"],
                           "d3q19_fp64_512_noise"),
            mlups=f64["mlups"], workload="d3q19_fp64_512_noise")
        out["copy_calibration"] = copy_calibration()
        # D2Q9 FP64 drop-in path: the 256^2 spinodal lattice (latency-bound)
        # and a 4096^2 one (HBM-bound), automatic kernel choice
        out["d2q9_fp64"] = {"256x256": time_workload("d2q9_fp64_256_noise", steps=2000, warmup=10),
                            "4096x4096": time_workload("d2q9_fp64_4096_noise", steps=30)}
        # the two-component model (two-pass FP64 kernels), with the
        # reference's own CPU path on the 3-D workload
        out["mcmp_fp64"] = {k: time_mcmp(k, steps=10) for k in MC_WORKLOADS}
        if not args.no_cpu_baseline:
            out["mcmp_fp64"]["cpu_baseline_mcmp3d_fp64_256_noise"] = cpu_baseline_mcmp("mcmp3d_fp64_256_noise")
    if args.workload == "d3q19_fp64_256_droplet":
        from mesobasic_amd import serve
        if not args.no_batched:
            # ticks of 50 steps: 76.2-76.3 GLUPS against 74.4-74.9 with 20 (MI355X)
            r = serve.run_batched(serve.make_batched_jobs(512), steps=2000, steps_per_tick=50)
            r["config"] = ("512 D2Q9 lattices, NX, NY uniform in {64..512}, G in [-6, -4.5], "
                           "tau in [0.6, 1.5], noisy rho 0.7, 2000 steps, FP32 compute / FP16 "
                           "storage, graph-captured ticks of 50 steps, tuner on")
            r["roofline_frac"] = round(r["aggregate_mlups"] * 1e6 * 36 / (HBM_PEAK_GBS * 1e9), 4)
            out["batched_d2q9_512"] = r
            print(f"# batched: {r['aggregate_mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
        if not args.no_serving:
            reqs = serve.make_requests(args.serve_requests, 200.0, 7)
            sample = serve.sample_ids(reqs, 6)
            r, kept = serve.run_stream(reqs, precision=SERVE_FORMS, work_limit=SERVE_WORK_LIMIT,
                                       steps_per_tick=SERVE_STEPS_PER_TICK, keep_ids=sample)
            r["config"] = (f"{args.serve_requests} requests, Poisson 200/s (seed 7), 70% D2Q9 "
                           "128..1024 per side, 30% D3Q19 64..192 per side, steps U[500, 5000], "
                           "G U[-6, -4.5], droplet or noisy init, D2Q9 requests in FP32 storage and "
                           "D3Q19 requests in FP16 node-shifted storage (two engines sharing the HBM "
                           "budget of 24 GiB), shortest remaining work first within "
                           f"{SERVE_WORK_LIMIT} nodes per step in all, ticks of "
                           f"{SERVE_STEPS_PER_TICK} steps")
            # accuracy of a seeded sample of served requests against the FP64
            # drop-in path (bit-identical to the reference), after the stream
            ok = [i for i in sample if i in kept and not kept[i]["failed"]]
            err = serve.check_against_fp64(reqs, kept, ok, fp32_floor=True)
            by = {"droplet": [], "noise": []}
            for i, (e, _) in err.items():
                by[reqs[i]["init"]].append(e)
            r["accuracy_vs_fp64"] = {
                "rel_l2_rho": {str(i): float(f"{e:.3e}") for i, (e, _) in err.items()},
                "fp32_storage_rel_l2_rho": {str(i): float(f"{f:.3e}") for i, (_, f) in err.items()},
                "max_rel_l2_droplet": max(by["droplet"]) if by["droplet"] else None,
                "max_rel_l2_noise": max(by["noise"]) if by["noise"] else None,
                "sample": "requests drawn by serve.sample_ids(reqs, 6): half droplet, half noise, "
                          "each at most 2e10 lattice updates; fp32_storage_* is a single-lattice "
                          "FP32 run of the same request (the FP32 rounding floor)"}
            out["serving"] = r
            print(f"# serving: p50 {r['p50_s']:.2f} s p99 {r['p99_s']:.2f} s "
                  f"{r['aggregate_mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
        if not args.no_driver:
            try:
                out["driver"] = driver_measurements()
                if not args.no_cpu_baseline:
                    out["driver"]["cpu_baseline_capsule2D_256x128"] = cpu_baseline_capsule()
            except Exception as e:   # the line is still printed, with the reason
                print(f"# driver measurements: {e}", file=sys.stderr, flush=True)
                out["driver"] = {"error": str(e)[:300]}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(q, shape, tau, G, init)
        if args.workload == "d3q19_fp64_256_droplet":
            # the reference's own D2Q9 case: 256^2 spinodal, one MPI rank
            qd, shd, _, taud, Gd, initd = WORKLOADS["d2q9_fp64_256_noise"]
            d2 = cpu_baseline_reference(qd, shd, taud, Gd, initd, ranks=1, seconds=8.0)
            if d2 is not None:
                d2["gpu_mlups"] = out.get("d2q9_fp64", {}).get("256x256", {}).get("mlups")
            out["cpu_baseline_d2q9_256_1rank"] = d2
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

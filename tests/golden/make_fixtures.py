"""Writes the golden fixtures in this folder by running the REFERENCE's own
lattice-Boltzmann classes (compiled in place by oracle/Makefile.ref into
oracle/_ref/ref_harness) on seeded inputs.

Only runnable where the reference tree and MPICH are present:
    make -C oracle ref && python tests/golden/make_fixtures.py [name ...]

Each case is one .npz: inputs (rho0, parameters) and outputs at the listed
steps -- the reference's r,u,v,w member arrays and, where stored, f
([node][q]) -- all float64, interior nodes, reference order.

Two-component cases (MC_CASES, names mcmp*) come from the reference's
mcmp2D / mcmp3D classes through oracle/_ref/ref_harness_mcmp: inputs rhoA0,
rhoB0; per step and component c in (A, B): rho_c_s, u_c_s, v_c_s, w_c_s,
fx_c_s, fy_c_s, fz_c_s (the force arrays) and, where stored, f_c_s.

Capsule cases (CAPSULE_CASES, names capsule*) come from the reference's
capsule2D / IBcapsules2D classes through oracle/_ref/ref_harness_capsule (one
rank): input rho0 and the circle (xo, yo, rad); per step s: rho_s, u_s, v_s,
fx_s, fy_s (the fluid's arrays), the IB node arrays x_s, y_s, vx_s, vy_s,
nfx_s, nfy_s, the ghost-row forces gfx_s, gfy_s ((2, nx): rows j = 0 and
j = ny+1) and, where stored, f_s.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
HARNESS_MC = os.path.join(ROOT, "oracle", "_ref", "ref_harness_mcmp")
HARNESS_CAP = os.path.join(ROOT, "oracle", "_ref", "ref_harness_capsule")
MPIRUN = "/opt/conda/bin/mpirun"

# name, q, shape, tau, G, walls, init, steps, steps_with_f, ranks
CASES = [
    ("d3q19_7x6x5_tau1", 19, (7, 6, 5), 1.0, -5.5, False, ("noise", 11), [0, 1, 2, 10, 100], [0, 1, 100], 3),
    ("d3q19_7x6x5_tau07", 19, (7, 6, 5), 0.7, -5.5, False, ("noise", 12), [0, 1, 2, 10, 100], [1, 100], 2),
    ("d3q19_walls_8x8x10", 19, (8, 8, 10), 1.0, -5.5, True, ("noise", 13), [1, 50], [50], 2),
    ("d3q19_droplet_16", 19, (16, 16, 16), 1.0, -5.5, False, ("droplet", 4.0), [1, 200], [], 4),
    ("d2q9_13x11", 9, (13, 11), 1.0, -5.0, False, ("noise", 14), [0, 1, 2, 10, 1000], [1, 1000], 3),
    ("d2q9_walls_12x10_tau08", 9, (12, 10), 0.8, -5.0, True, ("noise", 15), [1, 10, 300], [300], 2),
    ("d2q9_64x64", 9, (64, 64), 1.0, -5.0, False, ("noise", 16), [1000], [], 4),
    # large enough for the fused D3Q19 tiles (8 x 64, and partial 8 x 32 /
    # 4 x 64): z walls at tau 0.7, and periodic at tau 1.3
    ("d3q19_walls_4x16x64_tau07", 19, (4, 16, 64), 0.7, -5.5, True, ("noise", 17), [1, 30], [30], 2),
    ("d3q19_5x8x64_tau13", 19, (5, 8, 64), 1.3, -5.0, False, ("noise", 18), [1, 2, 40], [], 2),
]


# two components: name, q, shape, tauA, tauB, G, init, steps, steps_with_f,
# ranks[, film = (rhoA_sol, rhoB_sol)]; init ("noise", seed): rhoA and rhoB = 0.5 + 0.1 (x - 0.5) from
# seeds seed and seed + 1 (the reference's defaults rhoAi = rhoBi = 0.5,
# rhoAi_noise = rhoBi_noise = 0.1, mcmp2D.cpp:21-24); ("droplet", R): an
# A-rich disc or ball (1.0 / 0.05, width 2) in B, rhoB = 1.05 - rhoA
MC_CASES = [
    ("mcmp2d_13x11", 9, (13, 11), 1.0, 0.8, 3.0, ("noise", 21), [0, 1, 2, 10, 200], [1, 200], 3),
    ("mcmp2d_48x40_droplet", 9, (48, 40), 1.0, 1.0, 3.5, ("droplet", 10.0), [1, 300], [300], 2),
    ("mcmp3d_7x6x5", 19, (7, 6, 5), 1.0, 0.7, 3.0, ("noise", 23), [0, 1, 2, 10, 100], [1, 100], 3),
    ("mcmp3d_4x8x64", 19, (4, 8, 64), 1.0, 1.0, 3.2, ("noise", 25), [1, 30], [], 2),
    # film models (walls on the last axis, wall densities rhoA_sol, rhoB_sol)
    ("mcmpfilm2d_20x12", 9, (20, 12), 1.0, 1.0, 3.0, ("noise", 27), [0, 1, 2, 10, 150], [1, 150], 3,
     (0.3, 0.7)),
    ("mcmpfilm3d_6x5x8", 19, (6, 5, 8), 1.0, 0.9, 3.0, ("noise", 29), [0, 1, 2, 30], [1, 30], 2,
     (0.6, 0.4)),
]


# name, shape, tau, nnodes, kstiff, (xo, yo, rad), noise seed, steps, steps_with_f
CAPSULE_CASES = [
    ("capsule_48x40", (48, 40), 1.0, 24, 1.0, (20.0, 20.0, 6.0), 31, [0, 1, 2, 50, 300], [1, 300]),
    ("capsule_40x24_tau08", (40, 24), 0.8, 20, 0.5, (15.3, 11.7, 5.0), 33, [1, 3, 150], [150]),
    # a capsule reaching into the ghost row below the lower wall: its nodes
    # there spread force onto the ghost row and read r = 0 back (non-finite
    # velocities, after which the reference skips them)
    ("capsule_30x16_wall", (30, 16), 1.2, 16, 2.0, (10.0, 3.0, 3.5), 35, [1, 2, 5, 40], [40]),
]


def run_capsule_case(name, shape, tau, nnodes, kstiff, circle, seed, steps, steps_f):
    nx, ny = shape
    n = nx * ny
    rho0 = oracle.noise_rho(shape, seed, mean=0.5, amp=0.1).reshape(-1)
    out = {"rho0": rho0}
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, name)
        rho0.tofile(pre + "_rho0.bin")
        cmd = [MPIRUN, "-np", "1", HARNESS_CAP, str(nx), str(ny), repr(tau), str(nnodes), repr(kstiff),
               repr(circle[0]), repr(circle[1]), repr(circle[2]), pre + "_rho0.bin", pre,
               ",".join(map(str, steps))]
        subprocess.run(cmd, check=True)
        for s in steps:
            a = np.fromfile(f"{pre}_{s}.bin")
            for k, key in enumerate(("rho", "u", "v", "fx", "fy")):
                out[f"{key}_{s}"] = a[k * n:(k + 1) * n]
            if s in steps_f:
                out[f"f_{s}"] = a[5 * n:14 * n]
            off = 14 * n
            for key in ("x", "y", "vx", "vy", "nfx", "nfy"):
                out[f"{key}_{s}"] = a[off:off + nnodes]
                off += nnodes
            out[f"gfx_{s}"] = a[off:off + 2 * nx].reshape(2, nx)
            out[f"gfy_{s}"] = a[off + 2 * nx:off + 4 * nx].reshape(2, nx)
            assert off + 4 * nx == a.size
    meta = dict(shape=list(shape), tau=tau, nnodes=nnodes, kstiff=kstiff, circle=list(circle), seed=seed,
                steps=steps, steps_f=steps_f, body=[0.000008, 0.0], dt=1.0,
                source="reference capsule2D / IBcapsules2D classes via oracle/_ref/ref_harness_capsule")
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "ok")


def mc_inputs(shape, init):
    n = int(np.prod(shape))
    if init[0] == "noise":
        return (oracle.noise_rho(shape, init[1], mean=0.5, amp=0.1).reshape(-1),
                oracle.noise_rho(shape, init[1] + 1, mean=0.5, amp=0.1).reshape(-1))
    ra = oracle.droplet_rho(shape, init[1], rl=1.0, rv=0.05).reshape(-1)
    assert ra.size == n
    return ra, 1.05 - ra


def run_mc_case(name, q, shape, tauA, tauB, G, init, steps, steps_f, ranks, film=None):
    rA, rB = mc_inputs(shape, init)
    if film is not None and q == 19:
        # mcmp3DFilm initialises the fluid rows only (mcmp3DFilm.cpp:71-84):
        # the wall rows keep density 0
        for r in (rA, rB):
            r.reshape(shape)[:, :, 0] = 0.0
            r.reshape(shape)[:, :, -1] = 0.0
    nx, ny = shape[0], shape[1]
    nz = shape[2] if len(shape) == 3 else 1
    n = nx * ny * nz
    out = {"rhoA0": rA, "rhoB0": rB}
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, name)
        rA.astype(np.float64).tofile(pre + "_rA0.bin")
        rB.astype(np.float64).tofile(pre + "_rB0.bin")
        cmd = [MPIRUN, "-np", str(ranks), HARNESS_MC, str(q), str(nx), str(ny), str(nz), repr(tauA),
               repr(tauB), repr(G), pre + "_rA0.bin", pre + "_rB0.bin", pre, ",".join(map(str, steps))]
        if film is not None:
            cmd += [repr(film[0]), repr(film[1])]
        subprocess.run(cmd, check=True)
        for s in steps:
            for c in "AB":
                a = np.fromfile(f"{pre}_{c}_{s}.bin")
                for k, key in enumerate(("rho", "u", "v", "w", "fx", "fy", "fz")):
                    if q == 9 and key in ("w", "fz"):
                        continue
                    out[f"{key}_{c}_{s}"] = a[k * n:(k + 1) * n]
                if s in steps_f:
                    out[f"f_{c}_{s}"] = a[7 * n:]
    meta = dict(q=q, shape=list(shape), tauA=tauA, tauB=tauB, G=G, init=list(init), steps=steps,
                steps_f=steps_f, ranks=ranks, order=3 if q == 19 else 2, components=2,
                film=list(film) if film is not None else None,
                source="reference mcmp2D / mcmp3D (film: mcmp2DFilm / mcmp3DFilm) classes via "
                       "oracle/_ref/ref_harness_mcmp")
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "ok")


def run_case(name, q, shape, tau, G, walls, init, steps, steps_f, ranks):
    if init[0] == "noise":
        rho0 = oracle.noise_rho(shape, init[1])
    else:
        rho0 = oracle.droplet_rho(shape, init[1])
    nx, ny = shape[0], shape[1]
    nz = shape[2] if len(shape) == 3 else 1
    n = nx * ny * nz
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, name)
        rho0.astype(np.float64).tofile(pre + "_rho0.bin")
        cmd = [MPIRUN, "-np", str(ranks), HARNESS, str(q), str(nx), str(ny), str(nz), repr(tau),
               repr(G), str(int(walls)), pre + "_rho0.bin", pre, ",".join(map(str, steps))]
        subprocess.run(cmd, check=True)
        out = {"rho0": rho0.reshape(-1)}
        for s in steps:
            a = np.fromfile(f"{pre}_{s}.bin")
            out[f"rho_{s}"] = a[:n]
            out[f"u_{s}"] = a[n:2 * n]
            out[f"v_{s}"] = a[2 * n:3 * n]
            if q == 19:
                out[f"w_{s}"] = a[3 * n:4 * n]
            if s in steps_f:
                out[f"f_{s}"] = a[4 * n:]
    meta = dict(q=q, shape=list(shape), tau=tau, G=G, walls=walls, init=list(init), steps=steps,
                steps_f=steps_f, ranks=ranks, order=3 if q == 19 else 2,
                source="reference classes via oracle/_ref/ref_harness")
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "ok")


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -C oracle ref")
    only = set(sys.argv[1:])
    for c in CASES:
        if not only or c[0] in only:
            run_case(*c)
    for c in MC_CASES:
        if not only or c[0] in only:
            run_mc_case(*c)
    for c in CAPSULE_CASES:
        if not only or c[0] in only:
            run_capsule_case(*c)


if __name__ == "__main__":
    main()

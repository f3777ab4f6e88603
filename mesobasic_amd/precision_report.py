"""Accuracy of the reduced-precision forms against the FP64 path on one GPU.

    python -m mesobasic_amd.precision_report [--size 512] [--steps 5000]
           [--droplet-size 256] [--droplet-steps 2000]

Part 1 (spinodal decomposition, D3Q19 N^3 periodic, tau 1, G -5.0, rho =
0.7 +- 0.01 seeded noise): runs FP64 (bit-identical to the reference), FP32,
and FP16 and BF16 storage with the constant shift (0.7) and with the node
shift, for the same steps, and reports, per form,
  - relative L2 of the density field against FP64 (at several step counts;
    spinodal growth is unstable, so small differences grow),
  - coexistence densities: median density of the liquid (rho > midpoint)
    and vapour (rho <= midpoint) nodes, i.e. the bulk phases,
  - total-mass drift from the FP64 conservation sums.
Part 2 (droplet, D3Q19 M^3, tau 1, G -5.5, R = 3M/16, W 2, rho 2.0 / 0.15):
droplet radius from the liquid volume, Laplace pressure jump
dp = p_in - p_out with p = rho/3 + G psi^2 / 6 (psi = 1 - exp(-rho)), and
density L2 against FP64.

Prints a markdown report.  The FP64 run stands in for the FP64 oracle at
these sizes: the FP64 GPU path equals the oracle bit for bit (tests).
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from . import lbm
from .lbm import BF16N, BF16S, FP16N, FP16S, FP32, FP64, Lattice, droplet_rho, noise_rho

NAMES = {FP64: "FP64", FP32: "FP32", FP16S: "FP16 (constant shift 0.7)",
         BF16S: "BF16 (constant shift 0.7)", FP16N: "FP16 (node shift)", BF16N: "BF16 (node shift)"}
FORMS = (FP32, FP16S, BF16S, FP16N, BF16N)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def coexistence(rho):
    """Bulk liquid and vapour densities: the medians of the nodes above and
    below the midpoint density.  Means over the same sets also weigh in the
    interface nodes, whose share follows the interface area and the number of
    small domains, which a chaotic spinodal run changes from one rounding to
    the next (BF16 node shift, 512^3, 5,000 steps: two roundings of the
    initial equilibrium gave liquid means 2e-5 and 3.2e-4 off FP64, liquid
    medians 5e-5 and 6e-6)."""
    mid = 0.5 * (rho.max() + rho.min())
    liq, vap = rho[rho > mid], rho[rho <= mid]
    return (float(np.median(liq)) if liq.size else float("nan"),
            float(np.median(vap)) if vap.size else float("nan"))


def eos_p(rho, G):
    psi = 1.0 - np.exp(-rho)
    return rho / 3.0 + G * psi * psi / 6.0


def run_spinodal(prec, n, steps, checkpoints, seed=1):
    shape = (n, n, n)
    lat = Lattice(19, shape, 1.0, -5.0, precision=prec, rho_shift=0.7)
    lat.set_rho(noise_rho(shape, seed))
    lat.init()
    m0 = lat.conservation()[0]
    out = {}
    done = 0
    t0 = time.perf_counter()
    for c in checkpoints:
        lat.step(c - done)
        done = c
        out[c] = lat.fields(lbm.FIELDS_CURRENT)["rho"].astype(np.float64)
    secs = time.perf_counter() - t0
    m1 = lat.conservation()[0]
    lat.close()
    return out, (m1 - m0) / m0, secs


def run_droplet(prec, n, steps):
    shape = (n, n, n)
    R0 = 3.0 * n / 16.0
    lat = Lattice(19, shape, 1.0, -5.5, precision=prec, rho_shift=0.7)
    lat.set_rho(droplet_rho(shape, R0))
    lat.init()
    lat.step(steps)
    rho = lat.fields(lbm.FIELDS_CURRENT)["rho"].astype(np.float64)
    lat.close()
    return rho


def droplet_metrics(rho, G):
    rl, rv = coexistence(rho)
    vol = float(np.clip((rho - rv) / (rl - rv), 0.0, 1.0).sum())
    R = (3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0)
    n = rho.shape[0]
    c = n // 2
    core = rho[c - 2:c + 2, c - 2:c + 2, c - 2:c + 2].mean()
    far = rho[:4, :4, :4].mean()
    dp = float(eos_p(core, G) - eos_p(far, G))
    return dict(R=R, dp=dp, rho_l=rl, rho_v=rv)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--droplet-size", type=int, default=256)
    ap.add_argument("--droplet-steps", type=int, default=2000)
    args = ap.parse_args()

    cps = sorted({min(args.steps, c) for c in (100, 500, 1000, args.steps)})
    print(f"## Spinodal D3Q19 {args.size}^3, tau 1, G -5.0, rho 0.7 +- 0.01 (seed 1), "
          f"{args.steps} steps\n", flush=True)
    ref, drift64, s64 = run_spinodal(FP64, args.size, args.steps, cps)
    rl, rv = coexistence(ref[args.steps])
    print("| form | " + " | ".join(f"rho rel L2 @ {c}" for c in cps) +
          " | liquid rho | vapour rho | mass drift | s |")
    print("|---|" + "---|" * len(cps) + "---|---|---|---|")
    print(f"| FP64 | " + " | ".join("0" for _ in cps) + f" | {rl:.5f} | {rv:.5f} | {drift64:.2e} | {s64:.1f} |",
          flush=True)
    for prec in FORMS:
        got, drift, secs = run_spinodal(prec, args.size, args.steps, cps)
        l, v = coexistence(got[args.steps])
        errs = " | ".join(f"{rel_l2(got[c], ref[c]):.2e}" for c in cps)
        print(f"| {NAMES[prec]} | {errs} | {l:.5f} | {v:.5f} | {drift:.2e} | {secs:.1f} |", flush=True)
        del got

    print(f"\n## Droplet D3Q19 {args.droplet_size}^3, tau 1, G -5.5, R0 {3 * args.droplet_size / 16:g}, "
          f"{args.droplet_steps} steps\n", flush=True)
    ref = run_droplet(FP64, args.droplet_size, args.droplet_steps)
    m = droplet_metrics(ref, -5.5)
    print("| form | rho rel L2 | radius | Laplace dp | liquid rho | vapour rho |")
    print("|---|---|---|---|---|---|")
    print(f"| FP64 | 0 | {m['R']:.3f} | {m['dp']:.5e} | {m['rho_l']:.5f} | {m['rho_v']:.5f} |", flush=True)
    for prec in FORMS:
        got = run_droplet(prec, args.droplet_size, args.droplet_steps)
        g = droplet_metrics(got, -5.5)
        print(f"| {NAMES[prec]} | {rel_l2(got, ref):.2e} | {g['R']:.3f} | {g['dp']:.5e} | "
              f"{g['rho_l']:.5f} | {g['rho_v']:.5f} |", flush=True)


if __name__ == "__main__":
    main()

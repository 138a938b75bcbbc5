"""FLO-SIC E+FOD-force evaluations per second on one MI355X.

One step = one full evaluation of a synthetic system resident on the GPU: Fermi orbitals and Lowdin at the
FODs, orbital densities, exact Gaussian-pair Coulomb and LSDA XC potentials of every FLO on the mesh, the
RESULTS energies and E_SIC, the SIC matrix SIC(j,i) over all orbitals, V^i_ab over the FLOs, and the FOD
forces (both spins).  Default workload: K=64 atoms (Z in {1,6,8}), ladder basis, 60x194 Becke mesh,
96 FODs per spin.

Prints one JSON line.  --gpus above 1 is refused: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def coulomb_flops(counts, nrow_flops=True) -> float:
    """Algorithmic flops of one Coulomb launch: the pair-coefficient contraction 2*M*rows*nF plus ~30 flops per
    (point, pair) for the Boys/Hermite basis."""
    npair, nrows, M, nbas, nF = (float(x) for x in counts)
    return 2.0 * M * nrows * nF + 30.0 * M * npair


def cpu_baseline(sysm, npts_sample: int, threads: int) -> dict:
    """The CPU FP64 oracle on a bounded sample: the same system with every m-th mesh point (weights x m)."""
    import numpy as np

    from oracle import oracle as O
    from publicrelease_2020_amd.system import Mesh, System

    M = sysm.mesh.npts
    step = max(1, M // npts_sample)
    idx = np.arange(0, M, step)
    sub = System(name=sysm.name, basis=sysm.basis, mesh=Mesh(xyz=sysm.mesh.xyz[idx].copy(), w=sysm.mesh.w[idx] * step),
                 coef=sysm.coef, nocc=sysm.nocc, fod=sysm.fod, xc=sysm.xc, nspn=sysm.nspn)
    t0 = time.perf_counter()
    O.evaluate(sub, nthreads=threads, want_vab=True)
    dt = time.perf_counter() - t0
    est = dt * M / idx.size
    return {"value": 1.0 / est, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"oracle_eval on {idx.size} of {M} mesh points (every {step}th, weights x{step}), all "
                      f"{sum(sysm.nocc)} FLOs and {sysm.nbas} orbitals; {dt:.1f} s measured, scaled linearly "
                      f"in points to {est:.0f} s per eval"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--atoms", type=int, default=64, help="K of the synthetic workload (64 or 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        return 2

    import numpy as np

    from publicrelease_2020_amd import flosic as F
    from publicrelease_2020_amd import system as S

    cfg = 3 if args.atoms == 64 else 2
    t0 = time.perf_counter()
    sysm = S.make_config(cfg)
    t_gen = time.perf_counter() - t0
    print(f"# workload K={sysm.meta['K']} nbas={sysm.nbas} npts={sysm.mesh.npts} nocc={sysm.nocc} "
          f"(generated in {t_gen:.1f} s)", file=sys.stderr, flush=True)

    ctx = F.Flosic(0)
    t0 = time.perf_counter()
    ctx.load_system(sysm, F.PREC_FP64)
    for _ in range(max(1, args.warmup)):
        e64 = ctx.evaluate()
    t_setup = time.perf_counter() - t0
    print(f"# set-up + warmup {t_setup:.2f} s, E_SIC = {e64:.10f}", file=sys.stderr, flush=True)

    import torch
    torch.cuda.synchronize()
    coul_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = ctx.evaluate()
        coul_ms.append(ctx.timings()[3])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1000.0 * dt / args.steps
    value = args.steps / dt
    counts = ctx.counts()
    phases = ctx.timings()
    flops = coulomb_flops(counts)
    cms = float(np.mean(coul_ms))
    achieved = flops / (cms * 1e-3) / 1e12
    out = {
        "metric": "FLO-SIC E+FOD-force evals/s, mixed vs FP64 on one GPU at |dE|<1e-6 Ha",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"synthetic K={sysm.meta['K']} ladder basis, 60x194 Becke mesh, LSDA, "
                               f"{sysm.nocc[0]} FODs per spin",
                   "atoms": sysm.meta["K"], "nbas": sysm.nbas, "npts": sysm.mesh.npts,
                   "nflo": int(sum(sysm.nocc)), "npair": int(counts[0]), "hermite_rows": int(counts[1]),
                   "precision": "fp64"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["f64"], "traffic": None,
                     "kernel": "k_coulomb", "kernel_ms": cms, "flops_per_launch": flops},
        "phases_ms": {k: float(v) for k, v in zip(
            ["fo_lowdin", "orbitals", "hermite", "coulomb", "xc_energy", "sic_matrix", "vab", "forces", "total"],
            phases)},
        "esic": e,
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(sysm, args.cpu_sample, threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

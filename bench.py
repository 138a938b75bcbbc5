"""FLO-SIC E+FOD-force evaluations per second on one MI355X, mixed precision against FP64.

One step = one full evaluation of a synthetic system resident on the GPU: Fermi orbitals and Lowdin at the
FODs, orbital densities, exact Gaussian-pair Coulomb and LSDA XC potentials of every FLO on the mesh, the
RESULTS energies and E_SIC, the SIC matrix SIC(j,i) over all orbitals, V^i_ab over the FLOs, and the FOD
forces (both spins).  Default workload: K=64 atoms (Z in {1,6,8}), ladder basis (1832 functions), 60x194
Becke mesh (737129 points), 96 FODs per spin, LSDA.

The FP64 form is timed first; each reduced-precision form is then timed on the same system and accepted
only if |E_SIC - E_SIC(FP64)| < 1e-6 Ha and max |force - force(FP64)| < 1e-5 Ha/bohr.  `value` is the
throughput of the fastest accepted form.  Every evaluation is synchronous (results are on the host when
flosic_evaluate returns), so wall time over the steps is the step time.

The roofline block is for the Coulomb kernel of the reported form: achieved = algorithmic flops (2 M rows nF) over
its HIP-event time on the context stream; traffic = HBM bytes per dispatch from the rocprofv3 FETCH_SIZE /
WRITE_SIZE summary in profiles/coulomb_k64_traffic.txt (K=64 workload only, else null).

Prints one JSON line.  --gpus above 1 is refused: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": ("f64", 78.6), "mixed": ("f32", 157.3), "bf16x3": ("bf16", 2500.0), "bf16x2": ("bf16", 2500.0)}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "coulomb_k64_traffic.txt")


def coulomb_kernel_name(prec: str, nF: int) -> str:
    """Name of the Coulomb kernel instantiation the library launches for this form and FLO count."""
    cols = 64 if nF <= 64 else 128 if nF <= 128 else 192 if nF <= 192 else 256
    if prec == "fp64":
        return f"k_coulomb_ws<double, {cols}>"
    if prec == "mixed":
        return f"k_coulomb_ws<float, {cols}>"
    npc = 3 if prec == "bf16x3" else 2
    return f"k_coulomb_ws_bf16<{npc}, {cols if nF <= 192 else 128}>"


def measured_traffic(kernel: str, atoms: int):
    """HBM bytes per dispatch of `kernel` from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary committed for the K=64
    workload (profiles/coulomb_k64_traffic.txt), or None."""
    if atoms != 64 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("#"):
            continue
        parts = line.rsplit(None, 3)
        if len(parts) == 4 and parts[0].strip() == kernel:
            return float(parts[3])
    return None
DTYPE = {"fp64": "f64", "mixed": "f32", "bf16x3": "bf16x3", "bf16x2": "bf16x2"}
E_TOL, F_TOL = 1e-6, 1e-5


def log(msg: str):
    print(f"# {msg}", file=sys.stderr, flush=True)


def coulomb_flops(counts) -> float:
    """Algorithmic flops of one Coulomb launch: the Hermite-row contraction V[k,f] = sum_r R[k,r] H[r,f],
    2 * M * rows * nF, with rows the Hermite rows of the exponent-pair groups without padding (sum over groups
    of (L+1)(L+2)(L+3)/6)."""
    _npair, rows, M, _nbas, nF = (float(x) for x in counts[:5])
    return 2.0 * M * rows * nF


def run_form(sysm, prec_name: str, steps: int, warmup: int) -> dict:
    import numpy as np

    from publicrelease_2020_amd import flosic as F

    prec = {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED, "bf16x3": F.PREC_BF16X3, "bf16x2": F.PREC_BF16X2}[prec_name]
    ctx = F.Flosic(0)
    try:
        t0 = time.perf_counter()
        ctx.load_system(sysm, prec)
        for _ in range(max(1, warmup)):
            e = ctx.evaluate()
        log(f"{prec_name}: set-up + {max(1, warmup)} warmup eval(s) {time.perf_counter() - t0:.2f} s, "
            f"E_SIC = {e:.10f}")
        kern_ms, phases = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            e = ctx.evaluate()
            t = ctx.timings()
            kern_ms.append(t[9])
            phases.append(t)
        dt = time.perf_counter() - t0
        forces = np.concatenate([ctx.fod_forces(1), ctx.fod_forces(2)])
        out = dict(prec=prec_name, evals_s=steps / dt, ms_per_step=1000.0 * dt / steps, esic=e, forces=forces,
                   kernel_ms=float(np.mean(kern_ms)), phases=np.mean(np.array(phases), axis=0),
                   counts=ctx.counts())
        log(f"{prec_name}: {out['ms_per_step']:.1f} ms/eval, Coulomb kernel {out['kernel_ms']:.1f} ms")
        return out
    finally:
        ctx.close()


def cpu_baseline(sysm, npts_sample: int, threads: int) -> dict:
    """The CPU FP64 oracle on a bounded sample of the same workload: all FLOs and pairs, every m-th mesh point
    (weights x m).  The oracle splits its wall time into mesh-independent work (Fermi orbitals, the Hermite
    expansion of every pair for every FLO, forces) and work proportional to the points; one eval is
    estimated as fixed + per-point x M."""
    import numpy as np

    from oracle import oracle as O
    from publicrelease_2020_amd.system import Mesh, System

    M = sysm.mesh.npts
    step = max(1, M // npts_sample)
    idx = np.arange(0, M, step)
    sub = System(name=sysm.name, basis=sysm.basis, mesh=Mesh(xyz=sysm.mesh.xyz[idx].copy(), w=sysm.mesh.w[idx] * step),
                 coef=sysm.coef, nocc=sysm.nocc, fod=sysm.fod, xc=sysm.xc, nspn=sysm.nspn)
    t0 = time.perf_counter()
    r = O.evaluate(sub, nthreads=threads, want_vab=True)
    dt = time.perf_counter() - t0
    fixed, mesh = r["stage_s"]["fixed"], r["stage_s"]["mesh"]
    est = fixed + mesh * M / idx.size
    return {"value": 1.0 / est, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"oracle_eval (FP64 C/OpenMP) on {idx.size} of {M} mesh points (every {step}th, weights "
                      f"x{step}), all {sum(sysm.nocc)} FLOs, {sysm.nbas} basis functions; {dt:.1f} s measured = "
                      f"{fixed:.1f} s mesh-independent + {mesh:.1f} s over the sample points; one eval estimated "
                      f"as {fixed:.1f} + {mesh:.1f} x {M}/{idx.size} = {est:.0f} s"}


def copy_calibration(mib: int, reps: int) -> int:
    """Device-to-device copies of a known size through the HIP runtime alone, for scaling FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        print("hipMalloc failed", file=sys.stderr)
        return 1
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3):  # hipMemcpyDeviceToDevice
            print("hipMemcpy failed", file=sys.stderr)
            return 1
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "reps": reps}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--atoms", type=int, default=64, choices=(16, 32, 64),
                    help="K of the synthetic workload: 64 (precision workload), 16 (plain FP64), 32 (PBE)")
    ap.add_argument("--forms", default="mixed,bf16x3", help="reduced-precision forms to time after FP64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256, help="mesh points in the CPU oracle sample")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run device copies of this size (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the evaluation runs on a single device)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 5)

    import numpy as np

    from publicrelease_2020_amd import system as S

    t0 = time.perf_counter()
    sysm = S.make_config({16: 2, 64: 3, 32: 4}[args.atoms])
    log(f"workload K={sysm.meta['K']} nbas={sysm.nbas} npts={sysm.mesh.npts} nocc={sysm.nocc} "
        f"xc={'PBE' if sysm.xc else 'LSDA'} (generated in {time.perf_counter() - t0:.1f} s)")

    ref = run_form(sysm, "fp64", args.steps, args.warmup)
    forms = [ref]
    for name in [f for f in args.forms.split(",") if f]:
        r = run_form(sysm, name, args.steps, args.warmup)
        r["dE"] = abs(r["esic"] - ref["esic"])
        r["dF"] = float(np.max(np.abs(r["forces"] - ref["forces"])))
        r["accepted"] = bool(r["dE"] < E_TOL and r["dF"] < F_TOL)
        log(f"{name}: |dE| = {r['dE']:.2e} Ha, max|dF| = {r['dF']:.2e} Ha/bohr, accepted={r['accepted']}, "
            f"speed-up {r['evals_s'] / ref['evals_s']:.2f}x")
        forms.append(r)
    ok = [r for r in forms[1:] if r["accepted"]]
    best = max(ok, key=lambda r: r["evals_s"]) if ok else ref
    flops = coulomb_flops(best["counts"])
    arith, peak = PEAK[best["prec"]]
    kname = coulomb_kernel_name(best["prec"], int(best["counts"][4]))
    # compulsory HBM bytes of one launch: H once in the form's operand layout, V written once, mesh coordinates
    c = best["counts"]
    hbytes = {"fp64": 8, "mixed": 4, "bf16x3": 6, "bf16x2": 4}[best["prec"]]
    nfl = int(c[4])
    cols = 64 if nfl <= 64 else 128 if nfl <= 128 else 192 if nfl <= 192 else 256
    ldh = (nfl + cols - 1) // cols * cols
    alg_bytes = float(c[6]) * ldh * hbytes + 8.0 * c[2] * c[4] + 24.0 * c[2]
    achieved = flops / (best["kernel_ms"] * 1e-3) / 1e12
    counts = best["counts"]
    out = {
        "metric": "FLO-SIC E+FOD-force evals/s, mixed vs FP64 on one GPU at |dE|<1e-6 Ha",
        "value": best["evals_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[best["prec"]],
        "data": "synthetic",
        "config": {"workload": f"synthetic K={sysm.meta['K']} ladder basis, 60x194 Becke mesh, "
                               f"{'PBE' if sysm.xc else 'LSDA'}, {sysm.nocc[0]} FODs per spin",
                   "atoms": sysm.meta["K"], "nbas": sysm.nbas, "npts": sysm.mesh.npts,
                   "nflo": int(sum(sysm.nocc)), "shell_pairs": int(counts[0]), "groups": int(counts[5]),
                   "hermite_rows": int(counts[1]), "hermite_rows_streamed": int(counts[6]),
                   "precision": best["prec"]},
        "fp64": {"evals_s": ref["evals_s"], "ms_per_step": ref["ms_per_step"], "esic": ref["esic"],
                 "coulomb_kernel_ms": ref["kernel_ms"],
                 "coulomb_tflops": coulomb_flops(ref["counts"]) / (ref["kernel_ms"] * 1e-3) / 1e12},
        "forms": {r["prec"]: {"evals_s": r["evals_s"], "speedup_vs_fp64": r["evals_s"] / ref["evals_s"],
                              "dE_Ha": r["dE"], "max_dF_Ha_per_bohr": r["dF"], "accepted": r["accepted"],
                              "coulomb_kernel_ms": r["kernel_ms"]} for r in forms[1:]},
        "speedup_vs_fp64": best["evals_s"] / ref["evals_s"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(kname, sysm.meta["K"]),
                     "kernel": kname, "mfma_type": arith, "kernel_ms": best["kernel_ms"],
                     "flops_per_launch": flops,
                     "algorithmic_bytes": alg_bytes},
        "phases_ms": {k: float(v) for k, v in zip(
            ["fo_lowdin", "orbitals", "hermite", "coulomb", "xc_energy", "sic_matrix", "vab", "forces", "total",
             "coulomb_kernel"], best["phases"])},
        "esic": best["esic"],
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU oracle sample on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(sysm, args.cpu_sample, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

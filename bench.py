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


def coulomb_kernel_name(prec: str, nF: int, cols: int = 0) -> str:
    """Name of the Coulomb kernel instantiation the library launches for this form, FLO count and tile width
    (0: the widest, the library's default)."""
    if cols <= 0:
        cols = 64 if nF <= 64 else 128 if nF <= 128 else 192 if nF <= 192 else 256
    if prec == "fp64":
        return f"k_coulomb_ws<double, {cols}>"
    if prec == "mixed":
        return f"k_coulomb_ws<float, {cols}>"
    npc = 3 if prec == "bf16x3" else 2
    return f"k_coulomb_ws_bf16<{npc}, {min(cols, 192) if cols != 256 else 128}>"


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


def phase_roofline(sysm, counts, phases_ms, prec_name: str) -> dict:
    """Per-phase algorithmic work of one evaluation against the bound of each phase.  nd = field planes (1 LSDA,
    10 PBE: value, gradient, Hessian), n_s = FLOs of spin s.
      orbitals   phi = Phi_d C~ on the mesh (rocBLAS dgemm): 2 M N nF nd flops; reads the basis planes and writes
                 the FLO fields, 8 M nd (N + nF) bytes
      xc_energy  XC, V_tot, U and the energy partial sums: reads nd field planes and V_H, writes V_tot and U
                 (PBE: and 4 planes of the exact derivative), 8 M nF (nd + 3 [+ 4]) bytes
      sic_matrix G = Phi^T U split over point slices (dgemm): 2 M N nF flops (PBE: + 4 x 2 M n_s^2 for SIC_ex),
                 8 M (N + nF) bytes
      vab        V^i_ab, symmetric pair blocks: sum_s M n_s^2 (n_s + 1) flops, 8 M 2 nF bytes
    Peaks: FP64 MFMA 78.6 TFLOP/s for the dgemms, FP64 or FP32 MFMA for V^i_ab by form, HBM 8 TB/s."""
    M, N, nF = float(counts[2]), float(counts[3]), float(counts[4])
    pbe = bool(sysm.xc)
    nd = 10.0 if pbe else 1.0
    ns = [float(n) for n in sysm.nocc]
    vab_peak = PEAK["fp64"][1] if prec_name == "fp64" else PEAK["mixed"][1]
    work = {
        "orbitals": (1, 2 * M * N * nF * nd, 8 * M * nd * (N + nF), PEAK["fp64"][1]),
        "xc_energy": (4, None, 8 * M * nF * (nd + 3 + (4 if pbe else 0)), None),
        "sic_matrix": (5, 2 * M * N * nF + (sum(8 * M * n * n for n in ns) if pbe else 0.0), 8 * M * (N + nF),
                       PEAK["fp64"][1]),
        "vab": (6, sum(M * n * n * (n + 1) for n in ns), 8 * M * 2 * nF, vab_peak),
    }
    out = {}
    for name, (slot, flops, nbytes, peak) in work.items():
        sec = float(phases_ms[slot]) * 1e-3
        if sec <= 0:
            continue
        tbps = nbytes / sec / 1e12
        d = {"ms": float(phases_ms[slot]), "hbm_TBps": tbps, "hbm_frac": tbps / 8.0}
        if flops is not None:
            tf = flops / sec / 1e12
            d.update({"bound": "mfma", "achieved_TFLOPs": tf, "peak_TFLOPs": peak, "frac": tf / peak})
        else:
            d.update({"bound": "hbm", "frac": tbps / 8.0})
        out[name] = d
    return out


def next_rows(ctx, sysm, counts, reps: int = 3) -> dict:
    """Wall time of the DFT-side potential (APOTNL_DFT) and of SICLAG_DER on an evaluated context (both calls end
    with their results on the host), with the bound each is priced against: the DFT kernel streams the FLO fields
    (nd planes: 1 for LSDA, 10 for PBE) and Hartree potentials once, M (8 nd nF + 8 nF + 24) bytes, HBM; SICLAG_DER
    is three FP64 products over the points, 6 N nF M flops, FP64 MFMA.  Planes and fields rebuilt in mesh-block
    mode are recomputation and not counted."""
    M, nF, N = int(counts[2]), int(counts[4]), int(sysm.nbas)
    nd = 10 if sysm.xc else 1
    out = {}
    for name, call in (("apotnl_dft", ctx.apotnl_dft), ("siclag_der", ctx.siclag_der)):
        call()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        out[name] = {"ms": 1000.0 * (time.perf_counter() - t0) / reps}
    d = out["apotnl_dft"]
    d["algorithmic_bytes"] = float(M * (8 * nd * nF + 8 * nF + 24))
    d["bound"] = "hbm"
    d["achieved_GBps"] = d["algorithmic_bytes"] / (d["ms"] * 1e-3) / 1e9
    d["frac"] = d["achieved_GBps"] / 8000.0
    d = out["siclag_der"]
    d["flops"] = float(6 * N * nF * M)
    d["bound"] = "mfma"
    d["achieved_TFLOPs"] = d["flops"] / (d["ms"] * 1e-3) / 1e12
    d["frac"] = d["achieved_TFLOPs"] / 78.6
    log(f"next rows: apotnl_dft {out['apotnl_dft']['ms']:.2f} ms ({out['apotnl_dft']['achieved_GBps']:.0f} GB/s), "
        f"siclag_der {out['siclag_der']['ms']:.2f} ms ({out['siclag_der']['achieved_TFLOPs']:.1f} TFLOP/s)")
    return out


def run_form(sysm, prec_name: str, steps: int, warmup: int, coulomb_cols: int = 0, measure_next: bool = False) -> dict:
    import numpy as np

    from publicrelease_2020_amd import flosic as F

    prec = {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED, "bf16x3": F.PREC_BF16X3, "bf16x2": F.PREC_BF16X2}[prec_name]
    ctx = F.Flosic(0)
    try:
        t0 = time.perf_counter()
        ctx.load_system(sysm, prec)
        ctx.set_coulomb_cols(coulomb_cols)
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
        vab = [ctx.sic_vab(1), ctx.sic_vab(2)]
        out = dict(prec=prec_name, evals_s=steps / dt, ms_per_step=1000.0 * dt / steps, esic=e, forces=forces, vab=vab,
                   kernel_ms=float(np.mean(kern_ms)), phases=np.mean(np.array(phases), axis=0),
                   counts=ctx.counts())
        log(f"{prec_name}: {out['ms_per_step']:.1f} ms/eval, Coulomb kernel {out['kernel_ms']:.1f} ms "
            f"({ctx.coulomb_cols()} FLO columns per block)")
        if measure_next:
            out["next_rows"] = next_rows(ctx, sysm, out["counts"])
        return out
    finally:
        ctx.close()


def host_threads() -> int:
    """Host threads this process may use (its CPU affinity), at most 16 (the host cores set aside per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sysm, npts_sample: int, threads: int, nsamples: int = 3) -> dict:
    """The CPU FP64 oracle on bounded samples of the same workload: all FLOs and pairs on every m-th mesh point
    (weights x m), for nsamples disjoint strided samples (offsets 0, m/nsamples, ...).  The oracle splits its wall
    time into mesh-independent work (Fermi orbitals, the Hermite expansion of every pair for every FLO, forces) and
    work proportional to the points; each sample estimates one eval as fixed + per-point x M.  The value is the mean
    estimate; the spread over the samples is reported with it."""
    import numpy as np

    from oracle import oracle as O
    from publicrelease_2020_amd.system import Mesh, System

    M = sysm.mesh.npts
    step = max(1, M // npts_sample)
    est, walls, fixed, per_pt = [], [], [], []
    for k in range(nsamples):
        idx = np.arange((k * step) // nsamples, M, step)
        sub = System(name=sysm.name, basis=sysm.basis,
                     mesh=Mesh(xyz=sysm.mesh.xyz[idx].copy(), w=sysm.mesh.w[idx] * step),
                     coef=sysm.coef, nocc=sysm.nocc, fod=sysm.fod, xc=sysm.xc, nspn=sysm.nspn)
        t0 = time.perf_counter()
        r = O.evaluate(sub, nthreads=threads, want_vab=True)
        walls.append(time.perf_counter() - t0)
        f, m = r["stage_s"]["fixed"], r["stage_s"]["mesh"]
        fixed.append(f)
        per_pt.append(m / idx.size)
        est.append(f + m * M / idx.size)
        log(f"CPU sample {k}: {idx.size} points, {walls[-1]:.1f} s -> one eval {est[-1]:.0f} s")
    mean = float(np.mean(est))
    return {"value": 1.0 / mean, "unit": "evals/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "estimate_s": {"mean": mean, "min": float(min(est)), "max": float(max(est)),
                           "per_sample": [float(e) for e in est]},
            "sample": f"oracle_eval (FP64 C/OpenMP, exact Coulomb at every point) on {nsamples} disjoint strided "
                      f"samples of {M // step} of the {M} mesh points (every {step}th, weights x{step}), all "
                      f"{sum(sysm.nocc)} FLOs and shell pairs, {sysm.nbas} basis functions; {sum(walls):.1f} s "
                      f"measured; one eval estimated per sample as {np.mean(fixed):.1f} s mesh-independent + "
                      f"{np.mean(per_pt) * 1e3:.1f} ms per point x {M}: {mean:.0f} s "
                      f"(samples {min(est):.0f}..{max(est):.0f} s)"}


def reference_program_leg(steps: int) -> dict:
    """The reference program itself (nrlmol_exe, built from the reference sources, serial) against the library on one
    input: the config-1 H2O (NRLMOL default basis, the reference's own 65238-point mesh and converged LDA orbitals,
    5+5 FODs, tests/golden/reference_runs/h2o_lsda).  The reference's LDA SCF runs first (CPU) to make the state its
    non-SCF FLO-SIC evaluation starts from; then the same evaluation (NONSCF with forces) by the reference (its own
    'MAIN TIME FOR SIC' timer: CPU seconds of the SIC block on one core) and by the reference linked with the
    library's drop-ins (GPU); and the library alone on the run directory read by formats.load_run (FP64 and
    split-BF16x3, wall time per evaluation).  Skipped where the programs were not built."""
    import re
    import shutil
    import subprocess
    import tempfile

    import numpy as np

    nrl = os.path.join(ROOT, "oracle", "_ref", "nrlmol")
    fix = os.path.join(ROOT, "tests", "golden", "reference_runs", "h2o_lsda")
    if not (os.path.exists(os.path.join(nrl, "nrlmol_exe")) and os.path.exists(os.path.join(nrl, "nrlmol_flosic_hip"))):
        return {"skipped": "the reference program was not built (needs the reference sources at build time)"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from reference_runs import load

    from publicrelease_2020_amd import flosic as F
    tmp = tempfile.mkdtemp(prefix="flosic_ref_")
    out = {"input": "config 1: H2O, NRLMOL default basis (73 functions), the reference's 65238-point mesh and "
                    "converged LDA orbitals, 5+5 FODs", "cpu_model": cpu_model(), "reference_cores": 1}
    try:
        scf = os.path.join(tmp, "scf")
        shutil.copytree(os.path.join(fix, "inputs"), scf, copy_function=shutil.copyfile)

        def run(exe, d):
            t0 = time.perf_counter()
            with open(os.path.join(d, "out.txt"), "w") as f:
                subprocess.run([os.path.join(nrl, exe)], cwd=d, stdout=f, stderr=subprocess.STDOUT, timeout=600)
            return open(os.path.join(d, "out.txt")).read(), time.perf_counter() - t0

        _, dt = run("nrlmol_exe", scf)
        log(f"reference program: LDA SCF {dt:.1f} s")
        for name, exe in (("reference", "nrlmol_exe"), ("dropin", "nrlmol_flosic_hip")):
            d = os.path.join(tmp, name)
            shutil.copytree(scf, d, copy_function=shutil.copyfile)
            shutil.copyfile(os.path.join(fix, "FRMORB"), os.path.join(d, "FRMORB"))
            p = os.path.join(d, "NRLMOL_INPUT.DAT")
            txt = open(p).read()
            for key in ("NONSCFV", "NONSCFFORCESV"):
                txt = re.sub(rf"^({key}\s*=\s*)'N'", r"\1'Y'", txt, flags=re.M)
            open(p, "w").write(txt)
            text, dt = run(exe, d)
            z = re.findall(r"MAIN : ZSIC =\s*(\S+)", text)
            t_sic = re.findall(r"MAIN TIME FOR SIC :\s*(\S+)", text)
            out[name] = {"zsic": float(z[-1]) if z else None, "wall_s": dt,
                         "sic_block_cpu_s": float(t_sic[-1]) if t_sic else None}
            log(f"reference program ({name}): ZSIC {out[name]['zsic']}, SIC block {out[name]['sic_block_cpu_s']} s, "
                f"wall {dt:.1f} s")
        sysm, ref = load("h2o_lsda")
        for prec, pname in ((F.PREC_FP64, "fp64"), (F.PREC_BF16X3, "bf16x3")):
            ctx = F.Flosic(0)
            try:
                ctx.load_system(sysm, prec)
                ctx.evaluate()
                t0 = time.perf_counter()
                for _ in range(max(1, steps)):
                    e = ctx.evaluate()
                dt = (time.perf_counter() - t0) / max(1, steps)
            finally:
                ctx.close()
            out[f"library_{pname}"] = {"eval_ms": 1e3 * dt, "zsic": e, "dE_vs_reference": abs(e - ref["zsic"])}
            log(f"library {pname} on the same input: {1e3 * dt:.2f} ms per evaluation, |dE| vs reference "
                f"{abs(e - ref['zsic']):.2e} Ha")
        t_ref = out["reference"]["sic_block_cpu_s"]
        if t_ref:
            out["speedup_sic_block_vs_reference"] = {
                k: t_ref / (out[f"library_{k}"]["eval_ms"] * 1e-3) for k in ("fp64", "bf16x3")}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


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


def bench_batch(args) -> int:
    """The batch workload: n_sys independent systems from the batch generator (K uniform in [4,96], LSDA or PBE
    at random, 4..128 FODs per spin), every context in mesh-block mode.  Timed: the batch in the reduced form
    `--batch-prec` with its device part captured once and replayed (the reported value), the same systems
    evaluated one at a time, and each system in the plain FP64 form (the accuracy reference; timed on its second
    evaluation, so no buffer allocation is inside the timed region)."""
    import numpy as np

    from publicrelease_2020_amd import flosic as F
    from publicrelease_2020_amd import system as S

    t0 = time.perf_counter()
    systems = []
    for i in range(args.batch_systems):
        systems.append(S.make_config(5, index=i))
        log(f"system {i}: K={systems[-1].meta['K']} nbas={systems[-1].nbas} npts={systems[-1].mesh.npts} "
            f"nocc={systems[-1].nocc} xc={'PBE' if systems[-1].xc else 'LSDA'} "
            f"({time.perf_counter() - t0:.0f} s)")

    def engines(prec):
        out = []
        for i, sysm in enumerate(systems):
            e = F.Flosic(0)
            e.load_system(sysm, prec)
            e.set_mesh_block(args.batch_block)
            out.append(e)
        log(f"{len(out)} contexts loaded ({time.perf_counter() - t0:.0f} s)")
        return out

    def forces(e):
        return np.concatenate([e.fod_forces(1), e.fod_forces(2)])

    form = args.batch_prec
    mixed = engines({"mixed": F.PREC_MIXED, "bf16x3": F.PREC_BF16X3}[form])
    batch = F.FlosicBatch(mixed, use_graph=True)
    z = batch.evaluate()  # eager: allocations, plan
    t_eager = batch.timings()
    log(f"{form} batch eager: {t_eager[0]:.0f} ms wall, Coulomb launches {t_eager[2]:.0f} ms")
    z = batch.evaluate()  # captured
    log(f"{form} batch captured: {batch.timings()[0]:.0f} ms wall")
    walls = []
    for _ in range(args.steps):
        z = batch.evaluate()
        walls.append(batch.timings()[0])
        log(f"{form} batch replay: {walls[-1]:.0f} ms wall, device part {batch.timings()[1]:.0f} ms")
    t_batch = float(np.mean(walls)) * 1e-3
    f_mixed = [forces(e) for e in mixed]
    peak = sum(e.peak_device_bytes() for e in mixed)
    # the same systems one at a time
    t1 = time.perf_counter()
    seq_coul = 0.0
    for e in mixed:
        e.evaluate()
        seq_coul += e.timings()[9]
    t_seq = time.perf_counter() - t1
    log(f"{form} one at a time: {t_seq * 1e3:.0f} ms, Coulomb kernels {seq_coul:.0f} ms")
    counts = [e.counts() for e in mixed]
    batch.close()
    for e in mixed:
        e.close()
    # plain FP64 form, one system at a time: accuracy reference and speed comparison, timed on a warm context (the
    # second evaluation: buffers allocated by the first, as the batch's contexts are when it is timed)
    t_fp64, z64, dF = 0.0, [], 0.0
    for i, sysm in enumerate(systems):
        e = F.Flosic(0)
        e.load_system(sysm, F.PREC_FP64)
        e.set_mesh_block(args.batch_block)
        e.evaluate()
        t1 = time.perf_counter()
        z64.append(e.evaluate())
        t_fp64 += time.perf_counter() - t1
        dF = max(dF, float(np.max(np.abs(f_mixed[i] - forces(e)))))
        log(f"fp64 system {i}: {e.timings()[8]:.0f} ms")
        e.close()
    log(f"fp64 one at a time: {t_fp64 * 1e3:.0f} ms")
    dE = float(np.max(np.abs(z - np.array(z64))))
    flops = sum(coulomb_flops(c) for c in counts)
    achieved = flops / (t_eager[2] * 1e-3) / 1e12
    n = len(systems)
    out = {
        "metric": "FLO-SIC E+FOD-force evals/s, mixed vs FP64 on one GPU at |dE|<1e-6 Ha",
        "value": n / t_batch, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": 2,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": form, "data": "synthetic",
        "config": {"workload": f"batch of {n} synthetic systems, K uniform in [4,96], LSDA/PBE, 60x194 Becke "
                               f"meshes, mesh blocks of {args.batch_block} points, one step = one evaluation of "
                               f"every system", "systems": n, "precision": form,
                   "atoms": [s.meta["K"] for s in systems], "npts_total": int(sum(s.mesh.npts for s in systems)),
                   "nflo_total": int(sum(sum(s.nocc) for s in systems))},
        "batch_graph_s": t_batch, "one_at_a_time_s": t_seq, "fp64_one_at_a_time_s": t_fp64,
        "speedup_vs_one_at_a_time": t_seq / t_batch, "speedup_vs_fp64": t_fp64 / t_batch,
        "max_dE_Ha": dE, "max_dF_Ha_per_bohr": dF, "accepted": bool(dE < E_TOL and dF < F_TOL),
        "peak_device_bytes": int(peak),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK[form][1], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[form][1], "traffic": None,
                     "kernel": ("k_coulomb_ws_batch_all<float>" if form == "mixed" else "k_coulomb_ws_bf16_batch_all<*>"),
                     "kernel_ms": t_eager[2],
                     "flops_per_launch": flops, "one_at_a_time_kernel_ms": seq_coul},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_adjoint(args) -> int:
    """The differentiation workload: K=32, PBE, 48 FODs per spin.  FP64 evaluations with resident basis planes
    and with mesh-block recomputation (time and peak device memory of each), the FOD gradient against central
    differences of E_SIC (h = 1e-4 bohr) on a few coordinates, and the mixed-precision gradient against FP64."""
    import numpy as np

    from publicrelease_2020_amd import flosic as F
    from publicrelease_2020_amd import system as S

    sysm = S.make_config(4)
    log(f"workload K={sysm.meta['K']} nbas={sysm.nbas} npts={sysm.mesh.npts} nocc={sysm.nocc} PBE")
    modes = {}
    for name, block in (("resident", 0), ("mesh_block", args.adjoint_block)):
        e = F.Flosic(0)
        e.load_system(sysm, F.PREC_FP64)
        e.set_mesh_block(block)
        e.evaluate()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            z = e.evaluate()
        dt = (time.perf_counter() - t0) / args.steps
        g = np.concatenate([e.fod_gradient(1), e.fod_gradient(2)])
        modes[name] = dict(ms_per_eval=dt * 1e3, peak_device_bytes=e.peak_device_bytes(), esic=z, grad=g,
                           debdax=np.concatenate([e.fod_forces(1), e.fod_forces(2)]), engine=e)
        log(f"{name}: {dt * 1e3:.1f} ms/eval, peak device memory {e.peak_device_bytes() / 2**30:.2f} GiB")
        modes[name]["next_rows"] = next_rows(e, sysm, e.counts())
    e = modes["mesh_block"]["engine"]
    ref = modes["resident"]
    h = 1e-4
    rng = np.random.default_rng(4)
    coords = [(int(rng.integers(2)), int(rng.integers(sysm.nocc[0])), int(rng.integers(3)))
              for _ in range(args.fd_coords)]
    fd_err, deb_err = [], []
    for (s, i, x) in coords:
        fods = [np.array(f, dtype=np.float64) for f in sysm.fod]
        fods[s][i, x] += h
        e.set_fods(1, fods[0])
        e.set_fods(2, fods[1])
        ep = e.evaluate()
        fods[s][i, x] -= 2 * h
        e.set_fods(1, fods[0])
        e.set_fods(2, fods[1])
        em = e.evaluate()
        fd = (ep - em) / (2 * h)
        k = s * sysm.nocc[0] + i
        fd_err.append(abs(fd - ref["grad"][k, x]))
        deb_err.append(abs(fd - ref["debdax"][k, x]))
        log(f"coordinate {(s, i, x)}: FD {fd:.9f} adjoint {ref['grad'][k, x]:.9f} FRMORB2 {ref['debdax'][k, x]:.9f}")
    mixed = F.Flosic(0)
    mixed.load_system(sysm, F.PREC_MIXED)
    mixed.set_mesh_block(args.adjoint_block)
    zm = mixed.evaluate()
    gm = np.concatenate([mixed.fod_gradient(1), mixed.fod_gradient(2)])
    out = {
        "metric": "FOD gradient of E_SIC through Lowdin and the mesh contractions, K=32 PBE",
        "config": {"workload": "synthetic K=32 ladder basis, 60x194 Becke mesh, PBE, 48 FODs per spin",
                   "nbas": sysm.nbas, "npts": sysm.mesh.npts, "mesh_block_points": args.adjoint_block},
        "resident": {k: modes["resident"][k] for k in ("ms_per_eval", "peak_device_bytes", "esic")},
        "mesh_block": {k: modes["mesh_block"][k] for k in ("ms_per_eval", "peak_device_bytes", "esic")},
        "dE_block_vs_resident": abs(modes["mesh_block"]["esic"] - ref["esic"]),
        "max_dgrad_block_vs_resident": float(np.abs(modes["mesh_block"]["grad"] - ref["grad"]).max()),
        "fd_h_bohr": h, "fd_coordinates": coords, "max_fd_error_adjoint": float(max(fd_err)),
        "max_fd_error_frmorb2_formula": float(max(deb_err)),
        "mixed_dE": abs(zm - ref["esic"]), "mixed_max_dgrad": float(np.abs(gm - ref["grad"]).max()),
        "next_rows": {k: modes[k]["next_rows"] for k in ("resident", "mesh_block")},
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--atoms", type=int, default=64, choices=(16, 32, 64),
                    help="K of the synthetic workload: 64 (precision workload), 16 (plain FP64), 32 (PBE)")
    ap.add_argument("--forms", default="mixed,bf16x3", help="reduced-precision forms to time after FP64")
    ap.add_argument("--coulomb-cols", type=int, default=0,
                    help="FLO columns per Coulomb block: 0 widest, -1 autotuned, or 64/128/192/256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=128, help="mesh points in each CPU oracle sample")
    ap.add_argument("--cpu-samples", type=int, default=3, help="disjoint strided samples of the mesh for the CPU oracle")
    ap.add_argument("--no-reference-program", action="store_true",
                    help="skip the reference program (nrlmol_exe) against the library on the config-1 H2O inputs")
    ap.add_argument("--workload", default="precision", choices=("precision", "batch", "adjoint", "reference"),
                    help="precision: one system, FP64 and reduced forms; batch: many systems in one batch; "
                         "adjoint: K=32 PBE gradient, recomputation and memory; reference: only the reference program "
                         "against the library on the config-1 H2O inputs")
    ap.add_argument("--adjoint-block", type=int, default=65536, help="mesh-block points of the adjoint workload")
    ap.add_argument("--fd-coords", type=int, default=6, help="FOD coordinates checked by central differences")
    ap.add_argument("--batch-systems", type=int, default=48)
    ap.add_argument("--batch-block", type=int, default=8192, help="mesh-block points of the batch contexts")
    ap.add_argument("--batch-prec", default="mixed", choices=("mixed", "bf16x3"), help="form of the batch workload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run device copies of this size (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the evaluation runs on a single device)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 5)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "reference":
        print(json.dumps({"reference_program": reference_program_leg(args.steps)}), flush=True)
        return 0

    import numpy as np

    from publicrelease_2020_amd import system as S

    t0 = time.perf_counter()
    sysm = S.make_config({16: 2, 64: 3, 32: 4}[args.atoms])
    log(f"workload K={sysm.meta['K']} nbas={sysm.nbas} npts={sysm.mesh.npts} nocc={sysm.nocc} "
        f"xc={'PBE' if sysm.xc else 'LSDA'} (generated in {time.perf_counter() - t0:.1f} s)")

    ref = run_form(sysm, "fp64", args.steps, args.warmup, args.coulomb_cols, measure_next=True)
    forms = [ref]
    for name in [f for f in args.forms.split(",") if f]:
        r = run_form(sysm, name, args.steps, args.warmup, args.coulomb_cols)
        r["dE"] = abs(r["esic"] - ref["esic"])
        r["dF"] = float(np.max(np.abs(r["forces"] - ref["forces"])))
        r["dVab"] = float(max(np.max(np.abs(a - b)) for a, b in zip(r["vab"], ref["vab"])))
        r["accepted"] = bool(r["dE"] < E_TOL and r["dF"] < F_TOL)
        log(f"{name}: |dE| = {r['dE']:.2e} Ha, max|dF| = {r['dF']:.2e} Ha/bohr, max|dVab| = {r['dVab']:.2e} Ha, "
            f"accepted={r['accepted']}, "
            f"speed-up {r['evals_s'] / ref['evals_s']:.2f}x")
        forms.append(r)
    ok = [r for r in forms[1:] if r["accepted"]]
    best = max(ok, key=lambda r: r["evals_s"]) if ok else ref
    flops = coulomb_flops(best["counts"])
    arith, peak = PEAK[best["prec"]]
    kname = coulomb_kernel_name(best["prec"], int(best["counts"][4]), int(best["counts"][8]))
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
                              "dE_Ha": r["dE"], "max_dF_Ha_per_bohr": r["dF"], "max_dVab_Ha": r["dVab"],
                              "accepted": r["accepted"],
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
        "phase_roofline": phase_roofline(sysm, best["counts"], best["phases"], best["prec"]),
        "esic": best["esic"],
        "next_rows": ref["next_rows"],
    }
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"CPU oracle samples on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(sysm, args.cpu_sample, threads, args.cpu_samples)
    if not args.no_reference_program:
        out["reference_program"] = reference_program_leg(args.steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

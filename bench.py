#!/usr/bin/env python3
"""Benchmark: EBSD master-pattern directions/s on one MI355X.

One step = the whole master pattern of the workload: every incident direction of both
Lambert hemispheres through beam selection (ctem_plan_select, run again in every step),
DynMat assembly, the depth-integrated Lgh and the Sgh contraction, with the inputs
(per-energy setup and wave vectors) resident in HBM before the timed region. `value` is
directions per second of the mixed-precision form; the FP64 form is timed in the same run
and reported beside it with the ratio and the error of the mixed form against it.

roofline: the dominant kernel is the matrix-exponential GEMM. `achieved` prices its work
with the path's algorithmic count, 8 n^3 (q + j) per direction with the Pade degree q and
squaring count j that the reference's MatrixExponential would pick for that direction's
matrix (math.f90:208-219; ctem_plan_alg_flops), divided by the GEMM kernels' time measured
with HIP events on the plan's stream; `kernel_flops_tflops` is the GEMM's own product count
(8 n^3, or n^2 (n+1) 4 in the symmetric mode, per product launched) over the same time.

    python bench.py [--steps K] [--warmup W] [--workload fe_c1w] [--precision both|fp64|mixed]

The per-energy setup (Ucg look-up table, reflection list, sites) is the reference's own, as
its host computes it for bcc Fe at 20 kV (tests/golden/<workload>.npz); the wave vectors,
depth profile and Bethe parameters are generated here.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK = {"fp64": 78.6e12, "mixed": 2.5e15}  # MI355X dense FP64 MFMA / BF16 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_setup(name):
    from ctemsoft_amd import ebsd, workloads
    wl = workloads.get(name)
    path = os.path.join(REPO, "tests", "golden", f"{name}.npz")
    fx = np.load(path)
    # a reference-sized LUT does not hold every g-h of a hexagonal list (the reference reads
    # outside its array there); use the fixture's full-box LUT when there is one
    base = os.path.join(REPO, "tests", "golden", f"{name[:-1] if name.endswith('w') else name}.npz")
    full = np.load(base) if os.path.exists(base) else None
    st = ebsd.Setup.from_fixture(fx, full if full is not None and "lut_full" in full.files else None)
    return wl, st, fx


def load_jobs(name):
    """One job per energy bin: (label, setup, lam, dz, thick, fixture-or-None). Single-energy
    workloads read their fixture; energy series read the per-bin setups archive."""
    from ctemsoft_amd import ebsd, workloads
    wl = workloads.get(name)
    if wl.energies:
        arc = np.load(os.path.join(REPO, "tests", "golden", f"{name}_setups.npz"))
        jobs = []
        for i in range(int(arc["nenergy"])):
            view = {k[len(f"e{i}_"):]: arc[k] for k in arc.files if k.startswith(f"e{i}_")}
            st = ebsd.Setup.from_fixture(view)
            e = float(view["ekev"])
            dz = wl.dz_for(e)
            jobs.append((f"{e:g} kV", st, wl.depth_weights(e), dz, float(int(wl.nz * dz)), None))
        return wl, jobs
    wl, st, fx = load_setup(name)
    return wl, [(f"{wl.ekev:g} kV", st, wl.depth_weights(), wl.dz, float(int(wl.nz * wl.dz)), fx)]


def directions(wl, st):
    from ctemsoft_amd import ebsd
    npx, npy = ebsd.lambert_dims(wl.npx, st.usehex)
    ij, k, kn = ebsd.kvectors(npx, npy, st.isym, st.usehex, st.dsm, st.mlambda, 1)
    if wl.hemispheres == 2:
        ij2, k2, kn2 = ebsd.kvectors(npx, npy, st.isym, st.usehex, st.dsm, st.mlambda, -1)
        ij = np.concatenate([ij, ij2]); k = np.concatenate([k, k2]); kn = np.concatenate([kn, kn2])
    return ij, k, kn


def measured_traffic(workload, head):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy kernel), when one exists for this
    workload and precision."""
    path = os.path.join(REPO, "profiles", f"{workload}_{head}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_dominant:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(wl, lam, k_count, seconds=20.0, ekev=None, all_cores=True):
    """Reference beam loop (compiled from the reference sources, MKL sequential) on one
    host core over a seeded sample of the workload's northern directions; then the same loop
    as one process per host core (the reference's loop is serial, so its all-cores form is
    independent processes over disjoint directions), over a sample sized for ~`seconds`."""
    from oracle import refrun
    if not refrun.have_reference():
        return None, None
    rng = np.random.default_rng(wl.seed)
    with tempfile.TemporaryDirectory(prefix="ctem_cpu_") as work:
        # calibrate on 4 directions, then size the sample to ~`seconds`
        ks = np.sort(rng.choice(k_count, 4, replace=False))
        t4 = max(refrun.time_reference(work, wl, lam, ks, ekev), 1e-3)
        n = int(max(8, min(512, 4 * seconds / t4)))
        ks = np.sort(rng.choice(k_count, n, replace=False))
        t = refrun.time_reference(work, wl, lam, ks, ekev)
        one = dict(value=n / t, unit="directions/s", cores=1, kind="reference",
                   sample=f"{n} seeded directions of {wl.name}"
                          f"{'' if ekev is None else f' at {ekev:g} kV'} (northern hemisphere), Compute_DynMat + "
                          f"CalcLgh3 per direction, amdflang -O2 + MKL 2021.4 sequential, {t:.1f} s")
        allc = None
        if all_cores:
            procs = max(1, min(16, len(os.sched_getaffinity(0))))
            if procs > 1:
                m = int(min(k_count, max(procs, procs * n)))
                ks2 = np.sort(rng.choice(k_count, m, replace=False))
                t2 = refrun.time_reference_parallel(work, wl, lam, ks2, procs, ekev)
                allc = dict(value=m / t2, unit="directions/s", cores=procs, kind="reference",
                            sample=f"{m} seeded directions in {procs} processes (one per core, disjoint "
                                   f"directions), same build, {t2:.1f} s of beam loop")
    return one, allc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fe_c1w")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--terms", default="16,16",
                    help="split schemes of the mixed form, taylor,square: 3 = BF16x3, 6 = BF16x6, 16 = FP16x3")
    ap.add_argument("--symmetry", default="auto", choices=["auto", "off"],
                    help="symmetric products of the split mixed form for an inversion-symmetric LUT")
    ap.add_argument("--tune", default="on", choices=["on", "off", "fresh"],
                    help="launch autotuner (bucket multiple, chunk size) before timing; fresh ignores its cache")
    ap.add_argument("--graph", default="auto", choices=["auto", "on", "off"],
                    help="timed steps as captured HIP graphs (auto: on for energy series, all bins in one graph)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {a.gpus} requested); the directions of a "
              "master pattern would shard across GPUs without exchange, which this build does "
              "not run", file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401  (HIP runtime present)
    from ctemsoft_amd import ebsd

    if a.workload == "ebsd_det640":
        return bench_detector(a)
    if a.workload == "ecp_ni30":
        return bench_ecp(a)
    if a.workload == "mc_ni30":
        return bench_mc(a)
    if a.workload == "ni_c4w_adjoint":
        return bench_adjoint(a)
    wl, jobs = load_jobs(a.workload)
    tt, ts = (int(v) for v in a.terms.split(","))
    # one plan per energy bin, each holding its bin's crystal, all resident before timing; the
    # plans share the context's work space
    ctx = ebsd.Context(0)
    plans, ndir, npix = [], 0, 0
    ns_all, nw_all = [], []
    for label, st, lam, dz, thick, _ in jobs:
        ij, k, kn = directions(wl, st)
        lx, ly = ebsd.lambert_dims(wl.npx, st.usehex)
        npix += wl.hemispheres * (2 * lx + 1) * (2 * ly + 1)
        ctx.upload_crystal(st)
        plan = ctx.plan(k, kn, lam, dz, thick)
        ns, nw = plan.select()
        plan.set_mixed_terms(tt, ts)
        sym = plan.set_symmetry(1 if a.symmetry == "auto" else 0)
        ns_all.append(ns); nw_all.append(nw)
        plans.append(plan)
        ndir += len(k)
    ns, nw = np.concatenate(ns_all), np.concatenate(nw_all)

    use_graph = a.graph == "on" or (a.graph == "auto" and len(plans) > 1)
    tuned = None
    if a.tune != "off":
        from ctemsoft_amd import autotune
        tprec = "mixed" if a.precision != "fp64" else "fp64"
        tuned, _ = autotune.tune(ctx, plans, tprec, use_cache=(a.tune == "on"), log=log)

    def step(prec, graph=use_graph):
        # one pass of the path: beam selection and the whole master pattern of every bin (with
        # graphs: every bin's assembly .. contraction replayed as one captured launch)
        for plan in plans:
            plan.select()
        if graph:
            ebsd.run_plans(plans, prec)
        else:
            for plan in plans:
                plan.run(prec)
    name, cus = ctx.device_info()
    log(f"device {name}, {cus} CUs; workload {wl.name}: {len(jobs)} energy bin(s), {ndir} directions, nz={wl.nz}")
    log(f"strong beams {ns.min()}..{ns.max()} (mean {ns.mean():.1f}); weak {nw.min()}..{nw.max()} (mean {nw.mean():.1f})")
    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    res = {}
    for prec in precs:
        try:
            for w in range(a.warmup):
                step(prec)
                log(f"{prec}: warmup {w + 1}/{a.warmup} done")
        except ebsd.CtemError as e:
            log(f"{prec}: not available ({e})")
            continue
        times = []
        for s_ in range(a.steps):
            t0 = time.perf_counter()
            step(prec)
            times.append(time.perf_counter() - t0)
            log(f"{prec}: step {s_ + 1}/{a.steps} {times[-1] * 1e3:.1f} ms")
        tot_gpu_ms = sum(plan.timing()["total_ms"] for plan in plans)
        if use_graph:
            # a graph replay records no per-kernel events: one more (untimed) eager pass gives the
            # GEMM kernels' event times for the roofline
            step(prec, graph=False)
        tims = [plan.timing() for plan in plans]
        g = dict(gemm_ms=sum(t["gemm_ms"] for t in tims), gemm_flops=sum(t["gemm_flops"] for t in tims),
                 gemm_launches=sum(t["gemm_launches"] for t in tims))
        dots = np.concatenate([plan.dots().copy() for plan in plans])
        if prec == "fp64":
            afl = [plan.alg_flops() for plan in plans]
            alg = {k: sum(f[k] for f in afl) for k in ("total", "products", "solve")}
        ms = float(np.mean(times) * 1e3)
        res[prec] = dict(ms=ms, dps=ndir / (ms * 1e-3), gemm=g, dots=dots, gpu_ms=tot_gpu_ms,
                         gemm_tflops=g["gemm_flops"] / (g["gemm_ms"] * 1e-3) / 1e12 if g["gemm_ms"] > 0 else 0.0)
        log(f"{prec}: {ms:.1f} ms/step, {res[prec]['dps']:.0f} directions/s, GEMM {g['gemm_ms']:.1f} ms "
            f"({res[prec]['gemm_tflops']:.2f} TFLOP/s over {g['gemm_launches']} launches)")
    head = "mixed" if "mixed" in res else "fp64"
    r = res[head]
    err, nonfinite = None, None
    if "mixed" in res and "fp64" in res:
        d64, dmx = res["fp64"]["dots"], res["mixed"]["dots"]
        # a direction whose Bethe-corrected matrix grows without bound (a weak beam with an
        # excitation error near zero) overflows in FP64 as in the reference; it is compared
        # only for being non-finite in both forms
        fin = np.isfinite(d64).all(1)
        nonfinite = int((~fin).sum())
        if nonfinite:
            log(f"{nonfinite} directions non-finite in FP64: {np.argwhere(~fin)[:8, 0].tolist()}")
        err = float(np.max(np.abs(dmx[fin] - d64[fin]) / np.abs(d64[fin]))) if fin.any() else None
        if np.isfinite(dmx[~fin]).all(1).any():
            err = float("inf")
    cpu, cpu_all = (None, None) if a.no_cpu else cpu_baseline(
        wl, jobs[-1][2], len(plans[-1].k) // wl.hemispheres,
        ekev=None if not wl.energies else float(jobs[-1][0].split()[0]))
    gflops = r["gemm"]["gemm_flops"]
    gsec = r["gemm"]["gemm_ms"] * 1e-3
    # the path's algorithmic product flops (reference Pade rule) over the GEMM kernels' time
    ach = alg["products"] / gsec / 1e12 if ("fp64" in res and gsec > 0) else None
    line = {
        "metric": "master-pattern directions/s, 1 GPU: mixed vs FP64 at \u22641e-4 rel err",
        "value": round(r["dps"], 2),
        "unit": "directions/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(r["ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": (("fp16" if a.terms == "16,16" else "bf16/fp16") + "-split MFMA, f32 accumulate; f32 Lgh MFMA; f64 depth and contraction")
                 if head == "mixed" else "f64",
        "data": "synthetic wave vectors and depth profile; reference per-energy setup",
        "config": {"workload": f"{wl.name}: {wl.describe()}",
                   "directions": int(ndir), "strong_beams_mean": round(float(ns.mean()), 1),
                   "weak_beams_mean": round(float(nw.mean()), 1), "precision": head,
                   "mixed_terms": a.terms, "symmetric_products": bool(sym),
                   "launch": {"graph": use_graph, "bins_per_graph": len(plans) if use_graph else 0,
                              "bucket_multiple": tuned[0] if tuned else 32,
                              "chunk_gib": tuned[1] if tuned else 24.0,
                              "autotuned": tuned is not None}},
        "fp64_directions_per_s": round(res["fp64"]["dps"], 2) if "fp64" in res else None,
        "mixed_directions_per_s": round(res["mixed"]["dps"], 2) if "mixed" in res else None,
        # Lambert pixels of the master pattern(s) the step's directions determine through the Laue
        # equivalents (every hemisphere and energy bin), per second of the step (scatter not timed)
        "master_pixels_per_step": int(npix),
        "pixels_per_s": round(npix / (r["ms"] * 1e-3), 1),
        "mixed_over_fp64": round(res["mixed"]["dps"] / res["fp64"]["dps"], 3) if err is not None else None,
        "mixed_max_rel_err_vs_fp64": err,
        "nonfinite_directions": nonfinite if err is not None else None,
        "roofline": {"bound": "mfma", "achieved": None if ach is None else round(ach, 3),
                     "peak": PEAK[head] / 1e12, "unit": "TFLOP/s",
                     "frac": None if ach is None else round(ach * 1e12 / PEAK[head], 4),
                     "traffic": measured_traffic(wl.name, head),
                     "kernel": "k_zgemm" if head == "fp64" else ("k_sgemm" if a.terms == "16,16" else "k_cgemm"),
                     "alg_flops_per_step": alg["products"] if "fp64" in res else None,
                     "kernel_ms_per_step": round(r["gemm"]["gemm_ms"], 3),
                     "kernel_ms_source": ("HIP events of an eager pass after the timed graph steps"
                                          if use_graph else "HIP events over the timed steps"),
                     "kernel_launches_per_step": r["gemm"]["gemm_launches"],
                     "kernel_flops_tflops": round(r["gemm_tflops"], 3),
                     "kernel_flops_frac": round(r["gemm_tflops"] * 1e12 / PEAK[head], 4),
                     "kernel_flops_per_step": gflops},
        "path_alg_flops_per_step": alg["total"] if "fp64" in res else None,
        "path_alg_tflops": round(alg["total"] / (r["ms"] * 1e-3) / 1e12, 3) if "fp64" in res else None,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(line), flush=True)
    for plan in plans:
        plan.close()
    ctx.close()


# ---- EBSD detector patterns (include/ctem_detector.h) ------------------------------------------

def detector_inputs(seed=1):
    """The detector-pattern workload: the program's default 640 x 480 detector (&EBSDdata defaults,
    CTEMEBSD.f90:180-198, with thetac 10 and pattern centre 0), a master of the fe_c3w series'
    shape (npx 150, 21 energy bins 10..20 keV, one site) and Monte Carlo exit-energy histograms on
    a 101 x 101 Lambert grid, all synthetic (seeded smooth fields); energy window of every bin."""
    from ctemsoft_amd import detector as D
    rng = np.random.default_rng(seed)
    npx, nE, nsx = 150, 21, 50
    yy, xx = np.mgrid[-npx:npx + 1, -npx:npx + 1]
    base = (np.cos(xx / 9.0) * np.sin(yy / 13.0) + 1.5).astype(np.float32)
    sr = np.stack([base * np.float32(1 + 0.05 * e) + rng.random(base.shape, dtype=np.float32) * np.float32(0.01)
                   for e in range(nE)])[None].astype(np.float32)
    ys, xs = np.mgrid[-nsx:nsx + 1, -nsx:nsx + 1]
    acc = np.stack([np.round(300 * np.exp(-(xs ** 2 + ys ** 2) / 2500.0) * (e + 1) + rng.integers(0, 5, xs.shape))
                    for e in range(nE)], axis=-1).astype(np.int32)
    geo = D.Geometry(L=20000.0, thetac=10.0, delta=25.0, numsx=640, numsy=480, xpc=0.0, ypc=0.0,
                     energymin=9.5, energymax=20.5, beamcurrent=14.513, dwelltime=100.0)
    return geo, D.MCEnergy(acc, 10.0, 0.5, 70.0), sr


def detector_cpu_baseline(geo, mc, sr, quats, seconds=15.0):
    """The reference program itself (oracle/_ref/CTEMEBSD_clean: CTEMEBSD.f90 built from its
    sources, POIDEV replaced by the identity so that both sides compute the noise-free patterns)
    on a bounded sample of the orientations, one core."""
    import subprocess
    import tempfile
    from ctemsoft_amd.masterfile import write_master
    from oracle.refrun import REF_BIN, run_ebsd_program, write_mc_energy_file
    if not os.path.exists(os.path.join(REF_BIN, "CTEMEBSD_clean")):
        return None
    n = 4
    with tempfile.TemporaryDirectory(prefix="ctemdet_") as work:
        write_mc_energy_file(os.path.join(work, "mc.data"), mc.accum_e, mc.Ehistmin, mc.Ebinsize, mc.MCsig)
        write_master(os.path.join(work, "master.data"), sr, np.arange(sr.shape[1], dtype=np.float32) * 0.5 + 10,
                     [26])
        nml = dict(L=geo.L, thetac=geo.thetac, delta=geo.delta, numsx=geo.numsx, numsy=geo.numsy, xpc=geo.xpc,
                   ypc=geo.ypc, anglefile="angles.txt", eulerconvention="tsl", FZfile="master.data",
                   energyfile="mc.data", datafile="out.data", beamcurrent=geo.beamcurrent, dwelltime=geo.dwelltime,
                   energymin=geo.energymin, energymax=geo.energymax, binning=1, gammavalue=1.0, scalingmode="not",
                   axisangle=[0.0, 0.0, 0.0, 0.0])
        while True:
            with open(os.path.join(work, "angles.txt"), "w") as f:
                f.write(f"qu\n{n}\n" + "".join(" ".join(repr(float(v)) for v in q) + "\n" for q in quats[:n]))
            t0 = time.perf_counter()
            try:
                run_ebsd_program(work, nml, clean=True, timeout=120)
            except (subprocess.CalledProcessError, OSError):
                return None
            dt = time.perf_counter() - t0
            if dt > seconds / 4 or n >= len(quats):
                break
            n = min(len(quats), n * max(2, int(seconds / 4 / max(dt, 1e-3))))
    return {"value": round(n / dt, 3), "unit": "patterns/s", "cores": 1, "kind": "reference",
            "sample": f"{n} orientations through the reference program CTEMEBSD (noise-free build, file "
                      f"input/output included), {dt:.1f} s"}


def bench_detector(a):
    import torch
    from ctemsoft_amd import detector as D
    geo, mc, sr = detector_inputs()
    det = D.Detector(geo, mc)
    det.set_master(sr)
    info = det.info()
    n = 4096
    rng = np.random.default_rng(7)
    q = rng.normal(size=(n, 4))
    quats = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    npix = geo.numsx * geo.numsy
    dq = torch.from_numpy(quats).cuda()
    dp = torch.empty((n, geo.numsy, geo.numsx), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        det.patterns_device(n, dq.data_ptr(), dp.data_ptr())
        det.sync()
    kms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        det.patterns_device(n, dq.data_ptr(), dp.data_ptr())
        det.sync()
        kms.append(det.last_ms())
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / a.steps
    kms_mean = float(np.mean(kms))
    nwin = info["Emax"] - info["Emin"] + 1
    kw, kp = (nwin + 3) // 4 * 4, (sr.shape[1] - info["Emin"] + 1 + 3) // 4 * 4
    # algorithmic bytes of one launch: every pattern written once, the pixel directions, the
    # detector histograms and the master read once, the quaternions
    alg_bytes = 4 * npix * n + 12 * npix + 4 * kw * npix + 4 * kp * (sr.shape[2] * sr.shape[3]) + 16 * n
    ach = alg_bytes / (kms_mean * 1e-3) / 1e9
    cpu = None if a.no_cpu else detector_cpu_baseline(geo, mc, sr, quats)
    line = {
        "metric": "EBSD detector patterns/s, 1 GPU (noise-free, bit-equal to the reference program)",
        "value": round(n / wall, 2), "unit": "patterns/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 prefactor)", "data": "synthetic master, MC histograms and orientations",
        "config": {"workload": "ebsd_det640: 640 x 480 detector (program defaults, thetac 10), master npx 150 with "
                               "21 energy bins in the window, MC grid 101 x 101, 4096 orientations per step",
                   "orientations_per_step": n, "energy_bins": nwin},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(ach / 8000.0, 4), "traffic": None, "kernel": "k_ebsd_patterns",
                     "alg_bytes_per_launch": int(alg_bytes), "kernel_ms_per_step": round(kms_mean, 3),
                     "kernel_ms_source": "HIP events around the launch on the detector's stream",
                     "note": "VALU-bound: float32 Lambert geometry with correctly rounded division and "
                             "square root per pixel and orientation, 21 ordered bilinear bins"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    det.close()


# ---- electron channeling patterns (include/ctem_ecp.h) -----------------------------------------

# the reference's own template (templatefolder/CTEMECP.template: 30 kV, dmin 0.025, ktmax 5,
# npix 256, ten thicknesses 2..20 nm, Blochwv) with the SEM Bethe cutoffs
# (BetheParametersSEM.template) on fcc Ni, zone axis and foil normal [001]
ECP_NML = dict(voltage=30000.0, k=[0, 0, 1], fn=[0, 0, 1], dmin=0.025, ktmax=5.0, startthick=2.0, thickinc=2.0,
               numthick=10, npix=256, compmode="Blochwv", zintstep=1.0)
ECP_CUTOFFS = (30.0, 60.0, 0.05)


def ecp_cpu_baseline(xtal, seconds=20.0):
    """The reference program itself (oracle/_ref/CTEMECP: CTEMECP.f90 built from its sources) on a
    bounded disk of the same workload (npix grown until the run takes about `seconds` / 4..
    `seconds`), one core (the program's beam loop is serial); directions per second of the whole
    run, set-up and file output included."""
    import subprocess
    import tempfile
    from oracle.refrun import REF_BIN, run_ecp_program
    if not os.path.exists(os.path.join(REF_BIN, "CTEMECP")):
        return None
    npix = 2
    with tempfile.TemporaryDirectory(prefix="ctemecp_") as work:
        while True:
            nml = dict(ECP_NML, npix=npix)
            t0 = time.perf_counter()
            try:
                r = run_ecp_program(work, xtal, nml, cutoffs=ECP_CUTOFFS, timeout=300)
            except (subprocess.CalledProcessError, OSError):
                return None
            dt = time.perf_counter() - t0
            if dt > seconds / 4 or npix >= 64:
                break
            npix = min(64, int(npix * max(1.5, min(4.0, (seconds / 2 / max(dt, 1e-3)) ** 0.5))))
    return {"value": round(r["numk"] / dt, 3), "unit": "directions/s", "cores": 1, "kind": "reference",
            "sample": f"the reference program CTEMECP on the same workload with npix {npix} ({r['numk']} "
                      f"directions, set-up and file output included), {dt:.1f} s"}


def bench_ecp(a):
    import torch  # noqa: F401
    from ctemsoft_amd import ebsd, ecp, workloads, xtal as xtal_mod
    x = workloads.get("ni_c0").xtal
    nml = ECP_NML
    # the program's set-up (host, outside the timed step, as the per-crystal set-up of the master bench)
    st0, v0 = xtal_mod.native_setup(x, nml["voltage"] / 1000.0, nml["dmin"], ECP_CUTOFFS, ecp=(nml["voltage"], nml["k"], 0.0))
    dmt, rmt, mlambda = v0["dmt"], v0["rmt"], v0["mLambda"]
    dynfn = ecp.foil_normal(nml["fn"], dmt, rmt)
    ga = ecp.shortest_g(nml["k"], dmt, rmt)
    thetac, ktmax, bragg, galen = ecp.convergence(ga, rmt, mlambda, 0.0, nml["ktmax"])
    st, _ = xtal_mod.native_setup(x, nml["voltage"] / 1000.0, nml["dmin"], ECP_CUTOFFS,
                                  ecp=(nml["voltage"], nml["k"], np.float32(thetac) / np.float32(1000.0)))
    kv = ecp.kvectors(nml["k"], ga, ktmax, nml["npix"], dmt, rmt, mlambda)
    ctx = ebsd.Context(0)
    ctx.upload_crystal(st)
    occ = [s["occ"] for s in x["sites"]]
    numk = len(kv["kn"])
    res = {}
    for mode in ("Blochwv", "ScatMat"):
        p = ecp.EcpPlan(ctx, kv["k"], kv["kn"], mode, nml["startthick"], nml["thickinc"], nml["numthick"],
                        nml["zintstep"], dynfn, occ)
        ns, nw = p.select()
        for _ in range(a.warmup):
            p.select()
            p.run()
        t0 = time.perf_counter()
        tm = []
        for _ in range(a.steps):
            p.select()  # beam selection is part of the step
            p.run()
            tm.append(p.timing())
        wall = (time.perf_counter() - t0) / a.steps
        sr = p.sr()
        assert np.all(np.isfinite(sr)) and np.all(sr > 0)
        res[mode] = dict(wall=wall, gemm_ms=float(np.mean([t["gemm_ms"] for t in tm])),
                         gemm_flops=tm[-1]["gemm_flops"], total_ms=float(np.mean([t["total_ms"] for t in tm])),
                         ns=ns, nw=nw)
        log(f"{mode}: {numk} directions, strong {ns.min()}..{ns.max()} (mean {ns.mean():.1f}), weak mean "
            f"{nw.mean():.1f}; {wall * 1e3:.1f} ms per step, run {res[mode]['total_ms']:.1f} ms, GEMM "
            f"{res[mode]['gemm_ms']:.1f} ms / {res[mode]['gemm_flops'] / 1e12:.3f} TFLOP")
        p.close()
    ctx.close()
    b = res["Blochwv"]
    ach = b["gemm_flops"] / (b["gemm_ms"] * 1e-3) / 1e12
    cpu = None if a.no_cpu else ecp_cpu_baseline(x)
    line = {
        "metric": "ECP directions/s, 1 GPU (Blochwv, intensities at 10 thicknesses; beam selection included)",
        "value": round(numk / b["wall"], 1), "unit": "directions/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(b["wall"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic: the reference's own ECP template on fcc Ni",
        "config": {"workload": f"ecp_ni30: fcc Ni, 30 kV, zone axis and foil normal [001], dmin 0.025 nm, ktmax 5 "
                               f"(thetac {thetac:.1f} mrad), npix 256 ({numk} directions), thicknesses 2..20 nm "
                               f"(10), Bethe cutoffs 30/60/0.05",
                   "directions": numk, "scatmat_directions_per_s": round(numk / res["ScatMat"]["wall"], 1),
                   "scatmat_ms_per_step": round(res["ScatMat"]["wall"] * 1e3, 3)},
        "roofline": {"bound": "mfma", "achieved": round(ach, 2), "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(ach / 78.6, 4), "traffic": None, "kernel": "k_zgemm",
                     "alg_flops_per_step": b["gemm_flops"], "kernel_ms_per_step": round(b["gemm_ms"], 3),
                     "kernel_ms_source": "HIP events around the GEMM launches on the context's stream",
                     "note": "FP64 complex products of the block-matrix exponentials (2n = 2 x strong beams), "
                             "8 (2n)^3 flops per product on the true sizes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# ---- reverse mode (ctem_plan_adjoint) on the adjoint workload ---------------------------------

def bench_adjoint(a):
    """The adjoint workload ni_c4w (fcc Ni as ni_c0w on its own npx 100 grid, 5,151 directions): the
    master-pattern residual against the perturbed target model and its gradient with respect to the
    LUT and the site B (ctemsoft_amd.fit.master_loss_grad). One step = forward, scatter and
    residual, then the reverse pass (which reruns its forward); `value` for the workload's mixed
    form (ctem_plan_adjoint_mixed), the FP64 form beside it, with the two gradients' difference."""
    import torch  # noqa: F401
    from ctemsoft_amd import ebsd, fit, workloads
    fx = np.load(os.path.join(REPO, "tests", "golden", "ni_c0w.npz"))
    st = ebsd.Setup.from_fixture(fx)
    wl = workloads.get("ni_c4w")
    ij, k, kn = ebsd.kvectors(wl.npx, wl.npx, st.isym, st.usehex, st.dsm, st.mlambda, 1)
    npx = int(ij[:, 0].max())
    isym = int(st.isym)
    lam = wl.depth_weights()
    thick = float(int(wl.nz * wl.dz))
    ctx = ebsd.Context(0)
    ctx.upload_crystal(fit.perturbed_target_setup(st, wl.seed))
    pt = ctx.plan(k, kn, lam, wl.dz, thick)
    pt.run("fp64")
    target = pt.scatter(ij, npx, npx, isym)
    pt.close()
    ctx.upload_crystal(st)
    plan = ctx.plan(k, kn, lam, wl.dz, thick)
    ns, nw = plan.select()
    owner = fit.pixel_owner(ij, npx, npx, isym)
    res = {}
    for prec in ("fp64", "mixed"):
        for _ in range(a.warmup):
            fit.master_loss_grad(plan, ij, npx, npx, isym, target, owner, prec)
        t_rev, t_all = [], []
        for _ in range(a.steps):
            t0 = time.perf_counter()
            fit.master_loss(plan, ij, npx, npx, isym, target, prec)
            t1 = time.perf_counter()
            loss, glut, gb = fit.master_loss_grad(plan, ij, npx, npx, isym, target, owner, prec)
            t2 = time.perf_counter()
            t_all.append(t2 - t1)
            t_rev.append((t2 - t1) - (t1 - t0))
        plan.run(prec)
        t0 = time.perf_counter()
        plan.run(prec)
        fwd = time.perf_counter() - t0
        res[prec] = dict(step=float(np.median(t_all)), rev=float(np.median(t_rev)), fwd=fwd, loss=loss, glut=glut,
                         gb=np.asarray(gb, float))
    nk = len(k)
    g64, gmx = res["fp64"], res["mixed"]
    eg = float(np.abs(gmx["glut"] - g64["glut"]).max() / np.abs(g64["glut"]).max())
    eb = float(np.abs(gmx["gb"] - g64["gb"]).max() / np.abs(g64["gb"]).max())
    for prec, r in res.items():
        log(f"ni_c4w adjoint {prec}: {nk} directions, strong {ns.min()}..{ns.max()}; gradient step "
            f"{r['step'] * 1e3:.1f} ms (reverse {r['rev'] * 1e3:.1f} ms), forward {r['fwd'] * 1e3:.1f} ms; loss "
            f"{r['loss']:.6e}, |grad LUT| {np.abs(r['glut']).max():.3e}, grad B {r['gb']}")
    log(f"mixed vs fp64 gradient: LUT {eg:.2e}, B {eb:.2e} (max abs difference / max abs)")
    line = {
        "metric": "master-pattern gradient directions/s, 1 GPU (mixed forward + reverse: dloss/dLUT and dloss/dB)",
        "value": round(nk / gmx["step"], 1), "unit": "directions/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(gmx["step"] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp16-split MFMA, f32 accumulate (forward and the exponential's derivative); f64 depth recursions",
        "data": "synthetic: the reference-built Ni setup, a perturbed target model",
        "config": {"workload": f"ni_c4w: fcc Ni 20 kV, npx {wl.npx} ({nk} directions), strong beams "
                               f"{ns.min()}-{ns.max()}, residual against B x 1.05 and LUT x (1 + 0.01 N(0,1))",
                   "reverse_ms": round(gmx["rev"] * 1e3, 3), "forward_ms": round(gmx["fwd"] * 1e3, 3),
                   "fp64_directions_per_s": round(nk / g64["step"], 1), "fp64_ms_per_step": round(g64["step"] * 1e3, 3),
                   "fp64_reverse_ms": round(g64["rev"] * 1e3, 3), "fp64_forward_ms": round(g64["fwd"] * 1e3, 3),
                   "mixed_over_fp64": round(g64["step"] / gmx["step"], 2),
                   "grad_lut_rel_diff": eg, "grad_b_rel_diff": eb},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    plan.close()
    ctx.close()


# ---- Monte Carlo electron trajectories (include/ctem_mc.h) --------------------------------------

# the reference's own template (templatefolder/CTEMMC.template) on fcc Ni: 30 kV, 70 degree tilt,
# a 1501 x 1501 Lambert grid, 51 energy bins 5..30 keV, 101 depth bins, 12.5 million electrons
MC_NML = dict(sig=70.0, omega=0.0, numsx=1501, primeseed=932117, num_el=12_500_000, EkeV=30.0, Ehistmin=5.0,
              Ebinsize=0.5, depthmax=100.0, depthstep=1.0)
# double-precision arithmetic of one trajectory step (CTEMMC.f90:434-521) besides its six library
# calls (two log, acos, two sin, cos): energy loss 6, deflection angle 6, azimuth 1, the new
# direction 24, its normalisation with the compensated SUM 20, the next step length 14, the new
# position 9, the surface test 1
MC_FLOPS_PER_STEP = 81


def mc_cpu_baseline(xtal, seconds=20.0):
    """The reference program itself (oracle/_ref/CTEMMC: CTEMMC.f90 built from its sources) on a
    bounded number of electrons of the same workload with a 301 x 301 grid (the histograms' size
    does not change the per-electron work; the template's 1501 x 1501 output file alone is 930 MB),
    one core (the program runs its electrons on one stream)."""
    import subprocess
    import tempfile
    from oracle.refrun import REF_BIN, run_mc_program
    if not os.path.exists(os.path.join(REF_BIN, "CTEMMC")):
        return None
    n = 2000
    with tempfile.TemporaryDirectory(prefix="ctemmc_") as work:
        while True:
            nml = dict(MC_NML, numsx=301, num_el=n, nthreads=1)
            t0 = time.perf_counter()
            try:
                run_mc_program(work, xtal, nml, timeout=300)
            except (subprocess.CalledProcessError, OSError):
                return None
            dt = time.perf_counter() - t0
            if dt > seconds / 4 or n >= 2_000_000:
                break
            n = min(2_000_000, int(n * max(1.5, min(8.0, seconds / 2 / max(dt, 1e-3)))))
    return {"value": round(n / dt, 1), "unit": "electrons/s", "cores": 1, "kind": "reference",
            "sample": f"the reference program CTEMMC on {n} electrons of the same workload with a 301 x 301 grid "
                      f"(set-up and file output included), {dt:.1f} s"}


def bench_mc(a):
    import torch  # noqa: F401
    from ctemsoft_amd import mc, workloads, xtal as xtal_mod
    x = workloads.get("ni_c0").xtal
    p = MC_NML
    dens, avz, ava = xtal_mod.density(x)
    m = mc.MonteCarlo(p["sig"], p["omega"], p["numsx"], p["primeseed"], p["num_el"], p["EkeV"], p["Ehistmin"],
                      p["Ebinsize"], p["depthmax"], p["depthstep"], float(avz), float(dens), float(ava))
    for _ in range(a.warmup):
        m.reset()
        m.run("blocked")
    kms, steps = [], 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        m.reset()  # a fresh simulation: the histograms are zeroed in the step
        m.run("blocked")
        kms.append(m.last_ms())
        steps = m.steps()
    wall = (time.perf_counter() - t0) / a.steps
    ae, az, dropped = m.histograms()
    hits = int(ae.sum())
    m.close()
    kms_mean = float(np.mean(kms))
    ach = MC_FLOPS_PER_STEP * steps / (kms_mean * 1e-3) / 1e12
    log(f"mc_ni30: {p['num_el']} electrons, {steps} steps ({steps / p['num_el']:.1f} per electron), {hits} hits, "
        f"{dropped} dropped; kernel {kms_mean:.1f} ms, step {wall * 1e3:.1f} ms")
    cpu = None if a.no_cpu else mc_cpu_baseline(x)
    line = {
        "metric": "Monte Carlo electrons/s, 1 GPU (backscattered-electron energy and depth histograms)",
        "value": round(p["num_el"] / wall, 1), "unit": "electrons/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic: the reference's own MC template on fcc Ni",
        "config": {"workload": "mc_ni30: fcc Ni, 30 kV, tilt 70 deg, Lambert grid 1501 x 1501, energy bins 5..30 keV "
                               "by 0.5, depth bins 0..100 nm by 1, 12.5 million electrons (blocked random streams)",
                   "electrons": p["num_el"], "trajectory_steps": steps, "hits": hits},
        "roofline": {"bound": "mfma", "achieved": round(ach, 3), "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(ach / 78.6, 4), "traffic": None, "kernel": "k_mc",
                     "alg_flops_per_step": MC_FLOPS_PER_STEP, "kernel_ms_per_step": round(kms_mean, 3),
                     "kernel_ms_source": "HIP events around the launch on the object's stream",
                     "note": "FP64 vector work (78.6 TFLOP/s on MI355X as for the matrix cores); the count is "
                             "the steps' arithmetic without their six library calls (log, acos, sin, cos), "
                             "which take most of the executed instructions"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: integral-direct RHF Fock build (ERI engine + J/K digestion) on
one MI355X.

Workload (default): (H2O)_64/6-31G, 832 basis functions, the largest of the
water-cluster configurations.  A step is one full (non-incremental) Fock
build G(D) with the converged density already resident in HBM.  The
headline value is ERI quartets/s: canonical basis-function quartets held by
the shell quartets that survive the screen, divided by the ERI-kernel wall
time of one Fock build (HIP events around the class launches on the
library's stream), averaged over the timed steps.

Prints one JSON line (rank 0).  Extra fields carry the per-class launch
breakdown, the roofline of the dominant kernel, the same record for
(H2O)_8/6-31G, the mixed-precision and gradient records and the CPU
baseline (the oracle's reference algorithm on the host cores, on a bounded
sample of canonical quartets of the same molecule).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "haskell-abinitio_amd"), os.path.join(ROOT, "oracle")]

METRIC = "ERI quartets/s + SCF wall time, mixed vs FP64 at |ΔE|≤1µEh; gradient time"
PEAK_FP64 = 78.6      # TFLOP/s, MI355X FP64 (vector and matrix)
PEAK_FP32 = 157.3     # TFLOP/s, MI355X FP32 vector

WORKLOADS = {
    "h2o8": (8, "(H2O)_8 RHF/6-31G, all FP64: 8 rigid waters, seeded orientations, jittered 3x3x3 lattice "
                "(5.5 bohr, +-0.5), 104 basis functions, 80 electrons"),
    "h2o16": (16, "(H2O)_16 RHF/6-31G, FP64, 208 basis functions"),
    "h2o32": (32, "(H2O)_32 RHF/6-31G, 4x4x4 lattice, 416 basis functions"),
    "h2o64": (64, "(H2O)_64 RHF/6-31G, 6x6x6 lattice, 832 basis functions"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _finite(x):
    """x, or None where JSON has no number (the energy change after a
    single SCF step is +inf)."""
    import math
    return x if math.isfinite(x) else None


def cpu_baseline(mol, target_s=12.0):
    """The oracle (McMurchie-Davidson with the reference screen and Boys
    table, OpenMP) on a uniform sample of canonical quartets of the same
    molecule.  Returns quartets/s on the host threads."""
    import numpy as np
    import hfk_oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.boys_table()
    probe = O.random_canonical_quartets(mol.nbf, 4000, seed=1)
    t = time.perf_counter()
    O.eri_list(mol, probe)
    rate0 = len(probe) / max(time.perf_counter() - t, 1e-6)
    count = int(min(5e7, max(2e4, rate0 * target_s)))
    sample = O.random_canonical_quartets(mol.nbf, count, seed=2)
    t = time.perf_counter()
    O.eri_list(mol, sample)
    dt = time.perf_counter() - t
    m = mol.nbf * (mol.nbf + 1) // 2
    return {"value": count / dt, "unit": "quartets/s", "cores": threads, "kind": "port",
            "sample": "%d uniformly sampled canonical quartets of %d (%.1f s; oracle C restatement of "
                      "contracted4Centers, reference screen + Boys table, OpenMP)" % (count, m * (m + 1) // 2, dt)}


def cpu_extra(sample_s=4.0):
    """The oracle beyond the quartet-rate sample: end-to-end SCF (ERIs, then
    the DIIS loop on the flat ERI array, the reference's calcIntegrals + fock
    scheme) for water/STO-3G and (H2O)_8/6-31G, and, for (H2O)_32 and _64,
    the time to evaluate every canonical quartet once (one integral-direct
    Fock build) extrapolated from a uniform sample."""
    import hfk_oracle as O
    from hfk import systems
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.boys_table()
    out = {"cores": threads, "kind": "port", "end_to_end": [], "fock_build_extrapolated": []}
    for name, mol, kw in (("water STO-3G, reference SCF settings", systems.water("sto-3g"), {}),
                          ("(H2O)_8 6-31G, DIIS 6 vectors from step 1", systems.water_cluster(8, "6-31g"),
                           dict(ndiis=6, diis_start=float("inf"), max_steps=60))):
        t0 = time.perf_counter()
        S, H = O.overlap(mol), O.hcore(mol)
        flat, _kept = O.eri_flat(mol)
        t1 = time.perf_counter()
        r = O.scf(mol, lambda D: O.gmatrix(mol.nbf, flat, D), H, S, **kw)
        t2 = time.perf_counter()
        out["end_to_end"].append({"workload": name, "energy": r["energy"], "steps": r["steps"],
                                  "eri_s": t1 - t0, "scf_loop_s": t2 - t1, "total_s": t2 - t0})
        log("bench: cpu e2e %s %.2fs" % (name, t2 - t0))
    for nw in (32, 64):
        mol = systems.water_cluster(nw, "6-31g")
        m = mol.nbf * (mol.nbf + 1) // 2
        total = m * (m + 1) // 2
        probe = O.random_canonical_quartets(mol.nbf, 2000, seed=1)
        t = time.perf_counter()
        O.eri_list(mol, probe)
        rate0 = len(probe) / max(time.perf_counter() - t, 1e-6)
        count = int(min(5e7, max(2e4, rate0 * sample_s)))
        sample = O.random_canonical_quartets(mol.nbf, count, seed=2)
        t = time.perf_counter()
        O.eri_list(mol, sample)
        dt = time.perf_counter() - t
        out["fock_build_extrapolated"].append({
            "workload": "(H2O)_%d 6-31G" % nw, "canonical_quartets": total, "sample": count, "sample_s": dt,
            "quartets_per_s": count / dt, "fock_build_s": total * dt / count,
            "label": "extrapolated from a uniform sample of canonical quartets"})
        log("bench: cpu extrapolated (H2O)_%d build %.0fs" % (nw, total * dt / count))
    return out


def mixed_vs_fp64(workload, thr, sweep=(1e-6, 1e-5, 1e-3, 1e-2)):
    """SCF wall time FP64 vs mixed (FP32 for shell quartets whose Schwarz
    bound is below thr) on the same molecule; |dE|, orbital-energy error and
    one Fock build each."""
    import numpy as np
    import hfk
    from hfk import systems
    nwat, desc = WORKLOADS[workload]
    mol = systems.water_cluster(nwat, "6-31g")
    res = {"workload": desc, "fp32_threshold": thr}
    runs = {}
    for name, opts in (("fp64", hfk.options()), ("mixed", hfk.options(fp32_threshold=thr))):
        t0 = time.perf_counter()
        ctx = hfk.Context(mol, opts)
        t_ctx = time.perf_counter() - t0
        r = ctx.scfHF(diis_start=float("inf"), diis_size=6, max_steps=60)
        ms = []
        for _ in range(3):
            ctx.calcGmatrix(r.getDensity)
            ms.append(ctx.stats()["last_eri_ms"])
        cs = ctx.class_stats()
        ctx.close()
        runs[name] = r
        res[name] = {"energy": r.getEnergy, "steps": r.steps, "scf_s": r.seconds, "wall_s": r.seconds + t_ctx,
                     "fock_ms": min(ms),
                     "fp32_shell_quartets": sum(c["shell_quartets"] for c in cs if c["prec"] == 1),
                     "fp64_shell_quartets": sum(c["shell_quartets"] for c in cs if c["prec"] == 0)}
        log("bench: %s %s E=%.10f steps=%d scf %.2fs fock %.2fms" % (workload, name, r.getEnergy, r.steps,
                                                                       r.seconds, min(ms)))
    res["dE"] = runs["mixed"].getEnergy - runs["fp64"].getEnergy
    res["max_dEps"] = float(np.max(np.abs(runs["mixed"].getOrbE - runs["fp64"].getOrbE)))
    res["fock_speedup"] = res["fp64"]["fock_ms"] / res["mixed"]["fock_ms"]
    res["scf_speedup"] = res["fp64"]["scf_s"] / res["mixed"]["scf_s"]
    # threshold sweep: error and speed of the mixed SCF against FP64
    res["sweep"] = []
    for t in sweep:
        with hfk.Context(mol, hfk.options(fp32_threshold=t)) as ctx:
            try:
                r = ctx.scfHF(diis_start=float("inf"), diis_size=6, max_steps=60)
            except hfk.HfkError as e:
                res["sweep"].append({"fp32_threshold": t, "error": str(e)})
                log("bench: %s sweep thr %g: %s" % (workload, t, e))
                continue
            ctx.calcGmatrix(r.getDensity)
            fock_ms = ctx.stats()["last_eri_ms"]
            cs = ctx.class_stats()
        q32 = sum(c["shell_quartets"] for c in cs if c["prec"] == 1)
        q64 = sum(c["shell_quartets"] for c in cs if c["prec"] == 0)
        res["sweep"].append({"fp32_threshold": t, "steps": r.steps, "dE": r.getEnergy - runs["fp64"].getEnergy,
                             "max_dEps": float(np.max(np.abs(r.getOrbE - runs["fp64"].getOrbE))),
                             "fock_ms": fock_ms, "fp32_quartet_fraction": q32 / max(1, q32 + q64)})
        log("bench: %s sweep thr %g dE %.2e fock %.2fms fp32 %.1f%%" % (workload, t, res["sweep"][-1]["dE"], fock_ms,
                                                                        100.0 * q32 / max(1, q32 + q64)))
    return res


def gradient_bench(workload, reps=3):
    """Nuclear gradient wall time (hfk_gradient: derivative ERIs, one-electron
    and Pulay terms, W = 1/2 DFD) at the converged FP64 density, and the same
    with mixed-precision SCF density vs FP64."""
    import numpy as np
    import hfk
    from hfk import systems
    if workload == "c12h26":
        mol = systems.alkane(12, "6-31g*")
        desc = "C12H26 RHF/6-31G* (Cartesian d), 38 atoms, 232 basis functions"
    else:
        nwat, desc = WORKLOADS[workload]
        mol = systems.water_cluster(nwat, "6-31g")
    res = {"workload": desc, "natom": mol.natom, "nbf": mol.nbf}
    grads = {}
    for name, opts in (("fp64", hfk.options()), ("mixed", hfk.options(fp32_threshold=1e-4))):
        t0 = time.perf_counter()
        with hfk.Context(mol, opts) as ctx:
            t_ctx = time.perf_counter() - t0
            r = ctx.scfHF(diis_start=float("inf"), diis_size=6, max_steps=60)
            ms = []
            for _ in range(reps):
                t = time.perf_counter()
                g = ctx.gradient(r.getDensity)
                ms.append((time.perf_counter() - t) * 1e3)
            ctx.calcGmatrix(r.getDensity)
            fock_ms = ctx.stats()["last_eri_ms"]
        grads[name] = (r, g)
        res[name] = {"energy": r.getEnergy, "steps": r.steps, "scf_wall_s": r.seconds + t_ctx,
                     "fock_ms": fock_ms, "gradient_ms": min(ms), "max_abs_grad": float(np.max(np.abs(g))),
                     "translation_residual": float(np.max(np.abs(g.sum(axis=0))))}
        log("bench: gradient %s %s E=%.10f scf %.2fs grad %.1fms" % (workload, name, r.getEnergy,
                                                                     r.seconds + t_ctx, min(ms)))
    res["dE"] = grads["mixed"][0].getEnergy - grads["fp64"][0].getEnergy
    res["max_dgrad"] = float(np.max(np.abs(grads["mixed"][1] - grads["fp64"][1])))
    return res


def target_bench(nwat=64, thr=1e-4, tau=1e-12):
    """(H2O)_64/6-31G: energy and gradient in FP64 and in mixed precision
    (FP32 below Schwarz thr), both from the fragment guess, with incremental
    Fock builds (incremental_tau=tau); plus one FP64 SCF with a full Fock build
    every step (the reference's scheme) for its wall time and energy."""
    import numpy as np
    import hfk
    from hfk import systems
    mol = systems.water_cluster(nwat, "6-31g")
    res = {"workload": "(H2O)_%d RHF/6-31G, %d basis functions, %d atoms" % (nwat, mol.nbf, mol.natom),
           "fp32_threshold": thr, "incremental_tau": tau}
    scf = dict(diis_start=float("inf"), diis_size=8, max_steps=100)
    t0 = time.perf_counter()
    with hfk.Context(mol) as ctx:
        Dg = ctx.fragment_guess()
    res["fragment_guess_s"] = time.perf_counter() - t0
    log("bench: target %s fragment guess %.2fs" % (res["workload"], res["fragment_guess_s"]))
    out = {}
    for name, opts, inc in (("fp64_full_builds", hfk.options(), 0.0), ("fp64", hfk.options(), tau),
                            ("mixed", hfk.options(fp32_threshold=thr), tau)):
        t0 = time.perf_counter()
        with hfk.Context(mol, opts) as ctx:
            t_ctx = time.perf_counter() - t0
            r = ctx.scfHF(guess_density=Dg, incremental_tau=inc, **scf)
            rec = {"energy": r.getEnergy, "steps": r.steps, "max_error": r.max_error, "ctx_s": t_ctx,
                   "scf_s": r.seconds, "fock_s": r.fock_seconds,
                   "wall_s": t_ctx + res["fragment_guess_s"] + r.seconds}
            if name != "fp64_full_builds":
                t = time.perf_counter()
                g = ctx.gradient(r.getDensity)
                rec["gradient_s"] = time.perf_counter() - t
                rec["translation_residual"] = float(np.max(np.abs(g.sum(axis=0))))
                out[name] = (r, g)
        res[name] = rec
        log("bench: target %s E=%.10f %d steps scf %.2fs%s" % (name, r.getEnergy, r.steps, r.seconds,
                                                                 " grad %.2fs" % rec["gradient_s"] if "gradient_s" in rec else ""))
    (r64, g64), (r32, g32) = out["fp64"], out["mixed"]
    res["dE"] = r32.getEnergy - r64.getEnergy
    res["max_dgrad"] = float(np.max(np.abs(g32 - g64)))
    res["max_deps"] = float(np.max(np.abs(np.sort(r32.getOrbE) - np.sort(r64.getOrbE))))
    res["dE_incremental_vs_full"] = r64.getEnergy - res["fp64_full_builds"]["energy"]
    return res


def fock_rate(ctx, D, steps, warmup, workload):
    """Time `steps` full Fock builds G(D) with D resident in HBM: canonical
    function quartets of the surviving shell quartets / ERI-kernel time
    (HIP events on the library's stream around the class launches), the
    per-class breakdown and the roofline of the dominant launch."""
    import torch
    from hfk.flops import launch_flops
    n = ctx.n
    nf = n * (n + 1) // 2
    dD = torch.tensor(D, dtype=torch.float64, device="cuda")
    dG = torch.empty(nf, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for _ in range(warmup):
        ctx.fock_dev(None, dD.data_ptr(), dG.data_ptr())
    # the first statistics request counts each class's surviving quartets once
    # (one k_count_tasks launch per class); keep that out of the timed steps
    ctx.stats()
    torch.cuda.synchronize()
    eri_ms, per_launch = [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.fock_dev(None, dD.data_ptr(), dG.data_ptr())
        st = ctx.stats()
        eri_ms.append(st["last_eri_ms"])
        cs = ctx.class_stats()
        if per_launch is None:
            per_launch = [dict(c, ms_sum=0.0) for c in cs]
        # a launch's time is its own only if it ran alone in every timed build
        # (a context's first two builds run the classes one after another,
        # later ones concurrently with only the longest class alone)
        for acc, c in zip(per_launch, cs):
            acc["ms_sum"] += c["ms"]
            acc["solo"] = bool(acc["solo"]) and bool(c["solo"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = ctx.stats()
    fq = st["func_quartets"]
    eri_avg = sum(eri_ms) / len(eri_ms)
    m = n * (n + 1) // 2
    classes = []
    total_flops = 0
    for c in per_launch:
        ms = c["ms_sum"] / steps
        fl = launch_flops(c)
        total_flops += fl
        classes.append({"class": "(%s%s|%s%s)" % tuple("spdf"[c[k]] for k in ("la", "lb", "lc", "ld")),
                        "prec": "fp32" if c["prec"] else "fp64", "ms": round(ms, 4), "solo": bool(c["solo"]),
                        "tasks": c["tasks"], "shell_quartets": c["shell_quartets"], "prim_quartets": c["prim_quartets"],
                        "func_quartets": c["func_quartets"], "alg_flops": fl,
                        "alg_tflops": round(fl / (ms * 1e-3) / 1e12, 3) if ms > 0 else None})
    # the dominant kernel: the longest launch that had the GPU to itself (the
    # library runs it first and alone; the other launches overlap each other)
    dom = max([x for x in classes if x["solo"]] or classes, key=lambda x: x["ms"])
    achieved = dom["alg_flops"] / (dom["ms"] * 1e-3) / 1e12
    peak = PEAK_FP32 if dom["prec"] == "fp32" else PEAK_FP64
    # HBM bytes per launch of the dominant kernel, from the committed
    # FETCH_SIZE/WRITE_SIZE summary of the same workload (calibrated).
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "jk_tiles_%s_traffic.json" % workload)
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))["traffic"].get("%s %s" % (dom["class"], dom["prec"]))
    return {
        "value": fq / (eri_avg * 1e-3), "eri_ms": eri_avg, "ms_per_step": (t1 - t0) * 1e3 / steps,
        "canonical_quartets": m * (m + 1) // 2, "surviving_func_quartets": fq, "shell_pairs": st["nshell_pair"],
        "unscreened_rate": m * (m + 1) // 2 / (eri_avg * 1e-3),
        "alg_tflops_all_classes": total_flops / (eri_avg * 1e-3) / 1e12,
        "classes": classes,
        "roofline": {"bound": "fp64_valu" if dom["prec"] == "fp64" else "fp32_valu", "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": "k_jk_tiles %s %s (ERI + fused J/K digestion)" % (dom["class"], dom["prec"]),
                     "launch_ms": dom["ms"], "alg_flops_per_launch": dom["alg_flops"],
                     "note": "vector-ALU bound: MD ERI arithmetic per primitive quartet; peak is the MI355X "
                             + ("FP64 vector figure" if dom["prec"] == "fp64" else
                                "FP32 vector figure (packed FP32; the scalar FP32 this library issues runs at half of it)")
                             + "; achieved = Head-Gordon-Pople flops of the launch's surviving quartets / its "
                             "HIP-event time"},
    }


def serial_context(mol, opts=None):
    """A context whose class launches run one after another (HFK_CONCURRENT=0
    at creation), so every per-class time is the launch's own."""
    import hfk
    old = os.environ.get("HFK_CONCURRENT")
    os.environ["HFK_CONCURRENT"] = "0"
    try:
        return hfk.Context(mol, opts)
    finally:
        if old is None:
            del os.environ["HFK_CONCURRENT"]
        else:
            os.environ["HFK_CONCURRENT"] = old


def gemm_record(mol, reps=5):
    """The SCF's FP64 matrix-core GEMMs at the workload's size: X^T F X
    (hfk_unitary, two n^3 GEMMs) and D = 2 C_occ C_occ^T (hfk_density,
    n x n x occ), timed by the library's HIP events around the GEMM
    launches (the host copies of the ABI call excluded); roofline against
    the FP64 MFMA peak."""
    import numpy as np
    import hfk
    n = mol.nbf
    occ = int(round(mol.Z.sum() - mol.charge)) // 2
    rng = np.random.default_rng(5)
    X = rng.standard_normal((n, n))
    A = rng.standard_normal((n, n))
    Ff = (A + A.T)[np.triu_indices(n)]
    Cm = rng.standard_normal((n, n))
    out = {"n": n, "occ": occ, "peak": PEAK_FP64, "unit": "TFLOP/s", "bound": "fp64_mfma"}
    with hfk.Context(mol) as ctx:
        ctx.stats()   # the one-time quartet count, outside the timed calls
        for name, call in (("unitary_XtFX", lambda: ctx.unitaryTransf(X, Ff)),
                           ("density_2CCt", lambda: ctx.calcDensity(Cm, occ))):
            call()
            ms = []
            for _ in range(reps):
                call()
                st = ctx.stats()
                ms.append(st["last_gemm_ms"])
            t = sum(ms) / len(ms)
            achieved = st["last_gemm_flops"] / (t * 1e-3) / 1e12
            out[name] = {"ms": t, "flops": st["last_gemm_flops"], "achieved": achieved, "frac": achieved / PEAK_FP64}
            log("bench: gemm %s n=%d %.3f ms %.2f TFLOP/s (%.1f%% of FP64 MFMA peak)"
                % (name, n, t, achieved, 100 * achieved / PEAK_FP64))
    return out


def grad_roofline(mol, D, desc):
    """Per-class gradient launches (derivative integrals contracted with the
    two-particle density) on a context whose classes run one after another:
    each launch's time is its own.  Algorithmic flops from hfk.flops
    (f_grad_prim); roofline of the longest launch against the FP64 vector
    peak."""
    from hfk.flops import grad_launch_flops
    ctx = serial_context(mol)
    try:
        # an explicit energy-weighted density (here D itself: the kernels' work
        # does not depend on the values) keeps the W = 1/2 DFD Fock build out
        ctx.gradient(D, D)
        t0 = time.perf_counter()
        ctx.gradient(D, D)
        wall = time.perf_counter() - t0
        cs = [c for c in ctx.class_stats() if c["prec"] == 0]
    finally:
        ctx.close()
    classes = []
    for c in cs:
        fl = grad_launch_flops(c)
        ms = c["grad_ms"]
        classes.append({"class": "(%s%s|%s%s)" % tuple("spdf"[c[k]] for k in ("la", "lb", "lc", "ld")), "ms": ms,
                        "prim_quartets": c["prim_quartets"], "alg_flops": fl,
                        "alg_tflops": fl / (ms * 1e-3) / 1e12 if ms > 0 else None})
    dom = max(classes, key=lambda x: x["ms"])
    ach = dom["alg_flops"] / (dom["ms"] * 1e-3) / 1e12
    tot_ms = sum(x["ms"] for x in classes)
    tot_fl = sum(x["alg_flops"] for x in classes)
    log("bench: gradient kernels %s: %.1f ms serial, dominant %s %.1f ms %.1f%% of FP64 peak"
        % (desc, tot_ms, dom["class"], dom["ms"], 100 * ach / PEAK_FP64))
    return {"workload": desc, "gradient_wall_s": wall, "kernel_ms_sum": tot_ms,
            "alg_tflops_all_classes": tot_fl / (tot_ms * 1e-3) / 1e12, "classes": classes,
            "roofline": {"bound": "fp64_valu", "achieved": ach, "peak": PEAK_FP64, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP64, "kernel": "gradient %s (derivative ERIs x 2-PDM)" % dom["class"],
                         "launch_ms": dom["ms"], "alg_flops_per_launch": dom["alg_flops"]}}


def dclass_record(steps=3):
    """C12H26/6-31G* (Cartesian d, 232 functions): per-class J/K launches
    run one after another, with the cooperative d-class kernels; each
    class's algorithmic rate against the FP64 vector peak."""
    import numpy as np
    from hfk import systems
    from hfk.flops import launch_flops
    mol = systems.alkane(12, "6-31g*")
    n = mol.nbf
    rng = np.random.default_rng(0)
    A = rng.standard_normal((n, n)) * 0.01
    D = (A + A.T)[np.triu_indices(n)]
    ctx = serial_context(mol)
    try:
        ctx.stats()
        ms = {}
        for _ in range(steps):
            ctx.calcGmatrix(D)
            for c in ctx.class_stats():
                k = "(%s%s|%s%s)" % tuple("spdf"[c[x]] for x in ("la", "lb", "lc", "ld"))
                ms.setdefault(k, []).append((c["ms"], c))
        eri_ms = ctx.stats()["last_eri_ms"]
    finally:
        ctx.close()
    classes = []
    for k, v in ms.items():
        t = sum(x[0] for x in v) / len(v)
        c = v[-1][1]
        fl = launch_flops(c)
        coop = max(c["la"], c["lb"], c["lc"], c["ld"]) == 2 and \
            (c["la"] + 1) * (c["la"] + 2) // 2 * ((c["lb"] + 1) * (c["lb"] + 2) // 2) * \
            ((c["lc"] + 1) * (c["lc"] + 2) // 2) * ((c["ld"] + 1) * (c["ld"] + 2) // 2) > 54
        classes.append({"class": k, "kernel": "k_coop" if coop else "k_jk_tiles", "ms": t,
                        "prim_quartets": c["prim_quartets"], "alg_flops": fl,
                        "alg_tflops": fl / (t * 1e-3) / 1e12 if t > 0 else None})
    co = [x for x in classes if x["kernel"] == "k_coop"]
    dom = max(co, key=lambda x: x["ms"])
    ach = dom["alg_flops"] / (dom["ms"] * 1e-3) / 1e12
    log("bench: c12h26 Fock build %.2f ms (classes serial); d-class %s %.2f ms %.1f%% of FP64 peak"
        % (eri_ms, dom["class"], dom["ms"], 100 * ach / PEAK_FP64))
    return {"workload": "C12H26 RHF/6-31G* (Cartesian d), 232 basis functions, all FP64; random density",
            "eri_ms_serial": eri_ms, "classes": classes,
            "roofline": {"bound": "fp64_valu", "achieved": ach, "peak": PEAK_FP64, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP64, "kernel": "k_coop %s" % dom["class"], "launch_ms": dom["ms"],
                         "alg_flops_per_launch": dom["alg_flops"]}}


def headline_density(mol, big):
    """Converged density of the workload: from the fragment guess with
    incremental builds for the large clusters, else from the core guess."""
    import hfk
    t0 = time.perf_counter()
    with hfk.Context(mol) as ctx:
        ctx_s = time.perf_counter() - t0
        if big:
            t1 = time.perf_counter()
            Dg = ctx.fragment_guess()
            guess_s = time.perf_counter() - t1
            r = ctx.scfHF(guess_density=Dg, diis_start=float("inf"), diis_size=8, max_steps=100,
                          incremental_tau=1e-12)
            scf = {"guess": "fragment", "guess_s": guess_s, "incremental_tau": 1e-12}
        else:
            r = ctx.scfHF(diis_start=float("inf"), diis_size=6, max_steps=60)
            scf = {"guess": "core"}
    scf.update({"energy": r.getEnergy, "steps": r.steps, "scf_s": r.seconds, "ctx_s": ctx_s,
                "max_error": r.max_error, "last_step_dE": _finite(r.delta_energy), "fock_s": r.fock_seconds,
                "precision": "fp64"})
    log("bench: scf E=%.10f steps=%d %.3fs" % (r.getEnergy, r.steps, r.seconds))
    return r.getDensity, scf


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h2o64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-extra", action="store_true",
                    help="also time the oracle end to end (water, (H2O)_8) and extrapolate (H2O)_32/_64 builds")
    ap.add_argument("--no-scf", action="store_true", help="time the builds on a random density (no SCF first)")
    ap.add_argument("--fock-thr", type=float, default=0.0,
                    help="timed Fock builds in mixed precision, FP32 below this Schwarz bound (0: all FP64, the headline)")
    ap.add_argument("--no-small", action="store_true", help="skip the (H2O)_8 Fock-build record")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-workload", default="h2o32", choices=sorted(WORKLOADS))
    ap.add_argument("--mixed-thr", type=float, default=1e-4)
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-target", action="store_true", help="skip the (H2O)_64 FP64-vs-mixed energy+gradient record")
    ap.add_argument("--grad-workloads", default="h2o16,c12h26",
                    help="comma list of gradient/SCF workloads (h2oN or c12h26)")
    ap.add_argument("--no-gemm", action="store_true", help="skip the GEMM roofline record")
    ap.add_argument("--no-gradroof", action="store_true", help="skip the per-class gradient roofline record")
    ap.add_argument("--no-dclass", action="store_true", help="skip the C12H26 d-class record")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before timing, run one device copy of this size (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import hfk
    from hfk import systems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    nwat, desc = WORKLOADS[args.workload]
    mol = systems.water_cluster(nwat, "6-31g")
    n = mol.nbf
    log("bench: workload %s nbf=%d" % (args.workload, n))
    scf, D = None, None
    if not args.no_scf:
        try:
            D, scf = headline_density(mol, nwat >= 32)
        except hfk.HfkError as e:
            scf = {"error": str(e)}
            log("bench: scf failed: %s" % e)
    if D is None:
        rng = np.random.default_rng(0)
        A = rng.standard_normal((n, n)) * 0.01
        D = (A + A.T)[np.triu_indices(n)]
    if args.calibrate_copy_mib > 0:
        src = torch.ones(args.calibrate_copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t_ctx = time.perf_counter()
    ctx = hfk.Context(mol, hfk.options(fp32_threshold=args.fock_thr)) if args.fock_thr > 0 else hfk.Context(mol)
    ctx_s = time.perf_counter() - t_ctx
    rec = fock_rate(ctx, D, args.steps, args.warmup, args.workload)
    ctx.close()
    log("bench: %s Fock build %.2f ms, %.3e quartets/s, dominant %s %.1f%% of its %s peak"
        % (args.workload, rec["eri_ms"], rec["value"], rec["roofline"]["kernel"], 100 * rec["roofline"]["frac"],
           rec["roofline"]["bound"]))
    out = {
        "metric": METRIC,
        "value": rec["value"],
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.fock_thr <= 0 else "f64+f32",
        "data": "synthetic (seeded water cluster geometry; converged SCF density of it)",
        "config": {"workload": desc + ("; one full integral-direct Fock build G(D) per step, all FP64" if args.fock_thr <= 0
                                       else "; one full integral-direct Fock build G(D) per step, FP32 below Schwarz "
                                            "bound %g" % args.fock_thr),
                   "nbf": n, "screen": "reference per-primitive 1e-8", "boys": "reference table",
                   "canonical_quartets": rec["canonical_quartets"],
                   "surviving_func_quartets": rec["surviving_func_quartets"], "shell_pairs": rec["shell_pairs"]},
        "eri_ms": rec["eri_ms"],
        "ctx_s": ctx_s,
        "unscreened_rate": rec["unscreened_rate"],
        "alg_tflops_all_classes": rec["alg_tflops_all_classes"],
        "classes": rec["classes"],
        "roofline": rec["roofline"],
        "scf": scf,
    }
    if args.workload != "h2o8" and not args.no_small:
        m8 = systems.water_cluster(8, "6-31g")
        D8, scf8 = headline_density(m8, False)
        with hfk.Context(m8) as c8:
            r8 = fock_rate(c8, D8, 20, 3, "h2o8")
        out["h2o8"] = {"workload": WORKLOADS["h2o8"][1], "value": r8["value"], "eri_ms": r8["eri_ms"],
                       "ms_per_step": r8["ms_per_step"], "roofline": r8["roofline"], "classes": r8["classes"],
                       "scf": scf8}
        log("bench: h2o8 Fock build %.3f ms, %.3e quartets/s" % (r8["eri_ms"], r8["value"]))
    if not args.no_gemm:
        out["gemm"] = gemm_record(mol)
    if not args.no_gradroof:
        out["gradient_roofline"] = grad_roofline(mol, D, desc)
    if not args.no_dclass:
        out["dclass_c12h26"] = dclass_record()
    if not args.no_mixed:
        out["mixed_vs_fp64"] = mixed_vs_fp64(args.mixed_workload, args.mixed_thr)
    if not args.no_grad:
        out["gradient"] = [gradient_bench(w) for w in args.grad_workloads.split(",") if w]
    if not args.no_target:
        out["target_h2o64"] = target_bench()
    if not args.no_cpu:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(mol)
        if args.cpu_extra:
            out["cpu_baseline"]["extra"] = cpu_extra()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

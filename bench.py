#!/usr/bin/env python3
"""Benchmark of the lattice update on one MI355X.

Workload (default): D3Q15 pseudopotential MRT static vapour bubble in liquid, 256^3
periodic box, van der Waals EOS at T = 0.8 Tc, bubble radius 40 with a tanh interface
of width 5 between the Maxwell-construction densities, seeded +-1e-4 noise, FP32.
The f field uses the D3Q15 MRT scheme (Collision liMRT mapped onto the reference's
D3Q15 xuMRT scheme); --thermal adds the myMRT energy field with heat transfer.

One "step" is one multiPhasePP time step over the whole lattice; the metric is MLUPS
= stored nodes x steps / time / 1e6, timed with HIP events on the library's stream
after W warm-up steps, inputs resident in HBM.  Prints one JSON line, which also
carries sub-records timed in the same run: `thermal` (the same bubble with the myMRT
energy field and heat transfer), `ensemble` (the batched D2Q9 ensemble step on 256
cases, FP32, 2000 steps), `copy_ceiling` (flat copies of the same bytes) and
`cpu_baseline` (the FP64 oracle on the host cores).

Other workloads (--workload):
  ensemble   1024 D2Q9 liMRT droplet/bubble cases run to steady state in one batch per
             precision (FP64, FP32, FP16 storage): cases/s and Laplace-law errors vs FP64;
  warmstart  learned warm start: solver-generated training profiles, GPU training of the
             network, held-out D3Q15 128^3 bubbles run cold and warm to the same steady
             criterion: time-to-steady ratio and final-state deviations;
  accuracy   FP32 and FP16 storage against FP64 on the 256^3 bubble and the
             128 x 128 x 384 thermal column: coexistence densities, Laplace jump,
             spurious currents, temperature profile.
--pattern-probe times the single pass's memory pattern without arithmetic (flat copy,
pulls, pulls + ring re-pulls) with multiphaselb_amd/stream_probe.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
LAM15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)
GLAM15 = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)


def bytes_per_lu(precision, thermal, q=15):
    """Algorithmic bytes per lattice update: every population read and written once
    and every macroscopic field carried across steps written and read once."""
    s_pdf = {"fp64": 8, "fp32": 4, "fp16s": 2, "fp16n": 2}[precision]
    s_mac = 8 if precision == "fp64" else 4
    # FP16N also carries each node's FP16 reference density from step to step
    s_ref = 2 * 2 if precision == "fp16n" else 0
    if thermal:
        return 2 * q * s_pdf * 2 + 2 * s_mac * 5 + s_ref
    return 2 * q * s_pdf + 2 * s_mac * 1 + s_ref


def make_case(n, thermal, steps_total, cv=4.0):
    import multiphaselb_amd as M
    from multiphaselb_amd.casefiles import CaseSpec
    from multiphaselb_amd.initfields import maxwell, noise, sphere, vdw_critical

    rc, tc = vdw_critical(0.5, 4.0)
    T = 0.8 * tc
    rv, rl = maxwell(T, a=0.5, b=4.0)
    spec = CaseSpec(lbmodel=1, nx=n, ny=n, nz=n, btype=5, f_model=0, f_Lambda=LAM15, f_sigma=0.125,
                    g_model=1, g_Lambda=GLAM15, g_alpha1=0.25, g_alpha2=0.25, Cv=cv,
                    end=10 ** 8, write_interval=10 ** 8)
    d = tempfile.mkdtemp(prefix="mplb_bench_")
    M.write_case(d, spec)
    shape = (n, n, n)
    c0 = (n - 1) / 2.0
    radius = 40.0 * n / 256.0
    rho = sphere(shape, (c0, c0, c0), radius, rv, rl, 5.0) + noise(shape, 1e-4, 20240601 + 1)
    return d, spec, rho.ravel(), T, (rv, rl)


def run_gpu(args, thermal=None):
    import multiphaselb_amd as M

    thermal = args.thermal if thermal is None else thermal
    prec = {"fp64": M.FP64, "fp32": M.FP32, "fp16s": M.FP16S, "fp16n": M.FP16N}[args.precision]
    d, spec, rho, T, dens = make_case(args.size, thermal, args.steps + args.warmup, args.cv)
    case = M.Case(d, precision=prec, heat_transfer=thermal, fused=not args.generic,
                  extensions=M.EXT_LIMRT_D3Q15)
    N = case.N
    case.set("rho", rho)
    case.set("U", np.zeros((N, 3)))
    case.set("T", np.full(N, T))
    case.init()
    if args.warmup:
        case.step(args.warmup)
    done, total_ms, kern = case.time_steps(args.steps, per_kernel=True)
    assert done == args.steps, done
    rho_end = case.get("rho")
    ok = bool(np.isfinite(rho_end).all())
    fused = bool(case.get_info().fused_active)
    case.close()
    return dict(N=N, total_ms=total_ms, kern=kern, finite=ok, fused=fused, dens=dens,
                rho_range=(float(rho_end.min()), float(rho_end.max())))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_run(args, threads, budget_s):
    """The oracle on `threads` host threads: D3Q15 bubble at 64^3, as many steps as fit
    the time budget; returns (MLUPS, steps, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from _cases import oracle_params

    n = 64
    d, spec, rho, T, _ = make_case(n, args.thermal, 0, args.cv)
    spec.f_model = 2        # the D3Q15 MRT scheme liMRT is mapped onto
    o = O.Oracle(oracle_params(spec, ht=args.thermal, nthreads=threads))
    N = n ** 3
    o.set("rho", rho)
    o.set("U", np.zeros((N, 3)))
    o.set("T", np.full(N, T))
    o.init()
    o.run(1)                 # first touch of the arrays outside the timed loop
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 2000:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    return N * steps / dt / 1e6, steps, dt


def cpu_baseline(args, budget_s=10.0):
    """The reference-faithful FP64 C restatement of the step (the oracle: AoS storage,
    dense mat-vecs, the reference's operation order, OpenMP over nodes) on the host
    cores, beside the FP32 GPU number: one thread and all the threads the host offers
    (OMP_NUM_THREADS caps them).  Sample: the same D3Q15 bubble at 64^3."""
    threads = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", threads))
    threads = max(1, min(threads, cap))
    v1, s1, t1 = cpu_run(args, 1, budget_s)
    vn, sn, tn = cpu_run(args, threads, budget_s)
    what = "liMRT+myMRT thermal" if args.thermal else "liMRT"
    return dict(value=round(vn, 3), unit="MLUPS", cores=threads, kind="port",
                sample="D3Q15 %s bubble 64^3, FP64 oracle (reference-faithful C, not the FP32 GPU "
                       "arithmetic): %d steps in %.1f s on %d threads" % (what, sn, tn, threads),
                cpu_model=cpu_model(),
                single_thread={"value": round(v1, 3), "unit": "MLUPS", "cores": 1,
                               "sample": "same case, %d steps in %.1f s" % (s1, t1)})


def laplace(rho, p, L, rho_v, rho_l):
    """Interior/exterior density and pressure (averages over the disc r < R/2 around the
    box centre and over the corner region r > r_max - 3), the equivalent radius from
    the mass on the L-1 periodic sites, and the Laplace product dp * R of one steady
    case; None when the droplet or bubble has dissolved."""
    ii, jj = np.meshgrid(np.arange(L), np.arange(L), indexing="xy")
    c = (L - 1) / 2.0
    r = np.sqrt((ii - c) ** 2 + (jj - c) ** 2).ravel()
    uniq = ((ii < L - 1) & (jj < L - 1)).ravel()
    outer = r > r.max() - 3.0
    rout, pout = float(rho[outer].mean()), float(p[outer].mean())
    core = r < 2.0
    if abs(float(rho[core].mean()) - rout) < 0.25 * (rho_l - rho_v):
        return None
    area = (float(rho[uniq].sum()) - rout * (L - 1) ** 2) / (float(rho[core].mean()) - rout)
    R = float(np.sqrt(max(area, 0.0) / np.pi))
    inner = r < max(1.5, 0.5 * R)
    rin, pin = float(rho[inner].mean()), float(p[inner].mean())
    return dict(rho_in=rin, rho_out=rout, dp=pin - pout, R=R, gamma=(pin - pout) * R)


def ensemble_record(args, ncases=256, steps=2000):
    """The batched-ensemble step (the C3 kernel) timed inside the default run: the first
    ncases cases of the seed-7 ensemble in one FP32 batch, launch shape from the
    autotuner, `steps` steps in graph-captured 100-step runs with no steady criterion (no
    case leaves the table).  MLUPS of all case steps."""
    import multiphaselb_amd as M
    from multiphaselb_amd.casefiles import CaseSpec
    from multiphaselb_amd.initfields import ensemble, ensemble_density

    ens = ensemble(ncases, seed=7)
    hosts = {}
    for c in ens:
        if c["L"] not in hosts:
            d = tempfile.mkdtemp(prefix="mplb_ens_")
            M.write_case(d, CaseSpec(lbmodel=0, nx=c["L"], ny=c["L"], btype=3, f_model=0, a=0.5, b=4.0,
                                     end=10 ** 8, write_interval=10 ** 8))
            hosts[c["L"]] = M.Case(d, heat_transfer=False, device=-1)
    b = M.Batch([hosts[c["L"]] for c in ens], [ensemble_density(c) for c in ens], [c["T"] for c in ens],
                precision=M.FP32)
    b.init()
    block, xcd = b.tune(args.tune_steps)
    rows = b.launch_info()["march_rows"]
    s0 = np.array([b.case_steps(k) for k in range(len(ens))])
    t0 = time.perf_counter()
    nconv, s1, delta, state = b.run_to_steady(0.0, check_every=100, max_steps=steps)
    t1 = time.perf_counter()
    lu = float(sum(c["L"] ** 2 * int(s1[k] - s0[k]) for k, c in enumerate(ens)))
    b.close()
    mlups = lu / (t1 - t0) / 1e6
    # populations read once and written once per node-step: 2 x 9 x 4 B
    gbs = mlups * 1e6 * 72 / 1e9
    return {"value": round(mlups, 1), "unit": "MLUPS", "cases": len(ens),
            "nodes": int(sum(c["L"] ** 2 for c in ens)), "steps": steps, "precision": "fp32",
            "wall_s": round(t1 - t0, 3), "finite": bool((state != 2).all()),
            "hbm_gbs_effective": round(gbs, 1), "step_roofline_frac": round(gbs / PEAK_HBM_GBS, 4),
            "bytes_per_lu": 72,
            "launch": {"kernel": "k_batch_march" if rows else "k_batch_psi + k_batch_collide",
                       "rows_per_item": rows, "block": block, "xcd_order": xcd}}


def run_ensemble(args):
    """Batched ensemble: n independent D2Q9 liMRT droplet/bubble cases (sizes, radii,
    temperatures and shapes drawn from a seeded generator) advanced together until each
    is steady, in FP64 (the reference for the Laplace comparison) and the requested
    reduced precisions."""
    import multiphaselb_amd as M
    from multiphaselb_amd.casefiles import CaseSpec
    from multiphaselb_amd.initfields import ensemble, ensemble_density

    ens = ensemble(args.cases, seed=7)
    hosts = {}
    for c in ens:
        if c["L"] not in hosts:
            d = tempfile.mkdtemp(prefix="mplb_ens_")
            M.write_case(d, CaseSpec(lbmodel=0, nx=c["L"], ny=c["L"], btype=3, f_model=0, a=0.5, b=4.0,
                                     end=10 ** 8, write_interval=10 ** 8))
            hosts[c["L"]] = M.Case(d, heat_transfer=False, device=-1)
    rhos = [ensemble_density(c) for c in ens]
    nodes = sum(c["L"] ** 2 for c in ens)
    thr = {"fp64": args.steady_fp64, "fp32": args.steady_fp32, "fp16s": args.steady_fp16s,
           "fp16n": args.steady_fp16n}
    code = {"fp64": M.FP64, "fp32": M.FP32, "fp16s": M.FP16S, "fp16n": M.FP16N}
    precs = ["fp64"] + [x for x in args.ens_precisions.split(",") if x != "fp64"]
    res = {}
    for prec in precs:
        b = M.Batch([hosts[c["L"]] for c in ens], rhos, [c["T"] for c in ens], precision=code[prec])
        t0 = time.perf_counter()
        b.init()
        block, xcd = b.tune(args.tune_steps)
        rows = b.launch_info()["march_rows"]
        if args.persistent and rows:
            b.set_persistent(True)
        t1 = time.perf_counter()
        nconv, steps, delta, state = b.run_to_steady(thr[prec], check_every=100, max_steps=args.max_steps)
        t2 = time.perf_counter()
        lap = [laplace(b.get(k, "rho"), b.get(k, "p"), c["L"], c["rho_v"], c["rho_l"]) if state[k] == 1 else None
               for k, c in enumerate(ens)]
        lu = float(sum(c["L"] ** 2 * int(steps[k]) for k, c in enumerate(ens)))
        res[prec] = dict(steady=int(nconv), diverged=int((state == 2).sum()), threshold=thr[prec],
                         dissolved=int(sum(1 for k in range(len(ens)) if state[k] == 1 and lap[k] is None)),
                         wall_s=t2 - t0, run_s=t2 - t1, tune_s=t1 - t0, block=block, xcd_order=xcd, rows=rows,
                         persistent=bool(args.persistent and rows),
                         steps_median=float(np.median(steps)), steps_max=int(steps.max()),
                         mlups=lu / (t2 - t0) / 1e6, lap=lap, state=state)
        b.close()
        print("# %s: %d/%d steady, %d diverged, %.1f s (tune %.1f s), %s, xcd %d, steps median %d max %d" % (
            prec, nconv, len(ens), res[prec]["diverged"], t2 - t0, t1 - t0,
            "single pass, %d rows per item" % rows if rows else "two passes, block %d" % block, xcd,
            res[prec]["steps_median"], res[prec]["steps_max"]), file=sys.stderr, flush=True)
    ref = res["fp64"]["lap"]
    tc_bins = np.arange(0.6, 0.951, 0.05)
    out_prec = {}
    for prec in precs:
        r = res[prec]
        both = [k for k in range(len(ens)) if ref[k] is not None and r["lap"][k] is not None]
        e = {}
        for q in ("dp", "gamma", "rho_in", "rho_out", "R"):
            v = np.array([abs(r["lap"][k][q] - ref[k][q]) / abs(ref[k][q]) for k in both]) if both else np.zeros(1)
            e[q] = {"median": float(np.median(v)), "max": float(np.max(v))}
        # Laplace law per T/Tc bin: least-squares dp = s / R + c0
        fits = []
        for lo, hi in zip(tc_bins[:-1], tc_bins[1:]):
            ks = [k for k in both if lo <= ens[k]["tr"] < hi]
            if len(ks) < 3:
                continue
            A = np.array([[1.0 / r["lap"][k]["R"], 1.0] for k in ks])
            y = np.array([r["lap"][k]["dp"] for k in ks])
            Ar = np.array([[1.0 / ref[k]["R"], 1.0] for k in ks])
            yr = np.array([ref[k]["dp"] for k in ks])
            s = np.linalg.lstsq(A, y, rcond=None)[0][0]
            sr = np.linalg.lstsq(Ar, yr, rcond=None)[0][0]
            fits.append({"tr": [round(lo, 2), round(hi, 2)], "cases": len(ks), "slope": float(s),
                         "rel_err_vs_fp64": float(abs(s - sr) / abs(sr))})
        out_prec[prec] = {
            "cases_per_s": round(r["steady"] / r["wall_s"], 3), "steady": r["steady"], "diverged": r["diverged"],
            "dissolved": r["dissolved"], "laplace_cases": len(both),
            "threshold": r["threshold"], "wall_s": round(r["wall_s"], 2), "tune_s": round(r["tune_s"], 2),
            "mlups": round(r["mlups"], 1), "steps_median": r["steps_median"], "steps_max": r["steps_max"],
            "launch": {"kernel": ("k_batch_persist" if r["persistent"] else "k_batch_march") if r["rows"]
                       else "k_batch_psi + k_batch_collide",
                       "rows_per_item": r["rows"], "block": r["block"], "xcd_order": r["xcd_order"],
                       "persistent": r["persistent"]},
            "laplace_err_vs_fp64": e, "laplace_fit": fits,
            "laplace_fit_max_rel_err_vs_fp64": max([f["rel_err_vs_fp64"] for f in fits], default=None),
        }
    # the headline is the FP32 rate (its Laplace results match FP64 to the stated errors);
    # FP16 storage is reported beside it
    main_prec = "fp32" if "fp32" in precs else precs[-1]
    out = {
        "metric": "batched ensemble cases/s to steady state",
        "value": out_prec[main_prec]["cases_per_s"],
        "unit": "cases/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16s": "f16-storage/f32", "fp16n": "f16-node-ref-storage/f32"}[main_prec],
        "data": "synthetic (seeded ensemble, Maxwell densities)",
        "config": {"workload": "%d D2Q9 liMRT droplet/bubble cases, L in {64,96,128,192,256}, R in [6, L/4], "
                               "T/Tc in [0.6, 0.95], VdW a=0.5 b=4, seed 7, one batched launch per pass, "
                               "graph-captured 100-step runs" % len(ens),
                   "nodes": nodes, "steady": "max|drho|/rho_l over 100 steps < threshold"},
        "precisions": out_prec,
    }
    print(json.dumps(out))
    return 0


def accuracy_variants(args, log, run, rel):
    """The C1 bubble (256^3 periodic, R = 40, 2000 steps) with the model options the
    configurations leave out: xuSurfTen at kappa = 0.5 (van der Waals, T = 0.8 Tc), and
    the Carnahan-Starling (a = 1, b = 4) and Peng-Robinson (a = 2/49, b = 2/21,
    w = 0.344) equations of state at T = 0.85 Tc with their own Maxwell densities.  FP64
    (per-routine kernels) against FP32 and FP16 storage (fused kernels): bulk densities,
    Laplace jump, spurious currents; FP32 MLUPS over the last 500 steps."""
    import multiphaselb_amd as M
    from multiphaselb_amd.casefiles import CaseSpec
    from multiphaselb_amd.initfields import critical_temperature, maxwell, noise, sphere

    n, steps = args.acc_size, 2000
    N = n ** 3
    c0 = (n - 1) / 2.0
    R = 40.0 * n / 256.0
    idx = np.indices((n, n, n)).reshape(3, -1)[::-1]
    r = np.sqrt(((idx - c0) ** 2).sum(0))
    inner, outer = r < 0.5 * R, r > R + 20.0 * n / 256.0
    variants = [("xusurften_kappa0.5", 3, "vdw", dict(a=0.5, b=4.0), 0.8, dict(f_st=1, f_kappa=0.5)),
                ("carnahan_starling", 1, "cs", dict(a=1.0, b=4.0), 0.85, {}),
                ("peng_robinson", 2, "pr", dict(a=2.0 / 49.0, b=2.0 / 21.0, w=0.344), 0.85, {})]
    out = {}
    for name, code, eos, kw, tr, extra in variants:
        T = tr * critical_temperature(eos, **kw)
        rv, rl = maxwell(T, eos, **kw)
        spec = CaseSpec(lbmodel=1, nx=n, ny=n, nz=n, btype=5, f_model=0, f_Lambda=LAM15, f_sigma=0.125, g_model=None,
                        eos=code, a=kw["a"], b=kw["b"], w_pr=kw.get("w", 0.344), eos_keys_first=True,
                        end=10 ** 8, write_interval=steps, **extra)
        d = tempfile.mkdtemp(prefix="mplb_f4_")
        M.write_case(d, spec)
        rho = (sphere((n, n, n), (c0, c0, c0), R, rv, rl, 5.0) + noise((n, n, n), 1e-4 * rl, 20240603)).ravel()
        obs = {}
        for pname, prec, fused in (("fp64", M.FP64, False), ("fp32", M.FP32, True), ("fp16s", M.FP16S, True),
                                   ("fp16n", M.FP16N, True)):
            t0 = time.perf_counter()
            c = M.Case(d, precision=prec, heat_transfer=False, fused=fused, extensions=M.EXT_LIMRT_D3Q15)
            c.set("rho", rho)
            c.set("U", np.zeros((N, 3)))
            c.set("T", np.full(N, T))
            c.init()
            mlups = None
            if pname == "fp32":
                c.step(steps - 500)
                done, ms, _ = c.time_steps(500)
                mlups = N * done / (ms / 1e3) / 1e6
            else:
                c.step(steps)
            f = {k: np.asarray(c.get(k)) for k in ("rho", "U", "p")}
            c.close()
            um = np.sqrt((f["U"] ** 2).sum(1))
            obs[pname] = dict(rho_in=float(f["rho"][inner].mean()), rho_out=float(f["rho"][outer].mean()),
                              dp=float(f["p"][inner].mean() - f["p"][outer].mean()), umax=float(um.max()),
                              finite=bool(np.isfinite(f["rho"]).all()), wall_s=round(time.perf_counter() - t0, 1))
            if mlups:
                obs[pname]["mlups"] = round(mlups, 1)
            log("F4 %s %s: %s" % (name, pname, obs[pname]))
        ref = obs["fp64"]
        out[name] = {"T_over_Tc": tr, "maxwell": [rv, rl], "fp64": ref,
                     **{k: {**obs[k], "rel_err_vs_fp64": {q: rel(obs[k][q], ref[q]) for q in
                                                          ("rho_in", "rho_out", "dp", "umax")}}
                        for k in ("fp32", "fp16s", "fp16n")}}
    return out


def run_to_steady_case(case, rho_l, thr, check_every, max_steps):
    """Steps one case until max|drho|/rho_l over check_every steps < thr; returns
    (steps, wall seconds, steady flag)."""
    t0 = time.perf_counter()
    case.density_change()
    steps = 0
    while steps < max_steps:
        case.step(check_every)
        steps += check_every
        d = case.density_change()
        if not np.isfinite(d):
            return steps, time.perf_counter() - t0, False
        if d / rho_l < thr:
            return steps, time.perf_counter() - t0, True
    return steps, time.perf_counter() - t0, False


def bubble3d_cases(n, seed, L, droplet_share, rmin=6.0):
    from multiphaselb_amd.initfields import splitmix64, vdw_critical, maxwell

    u = splitmix64(seed, 3 * n).reshape(n, 3)
    tc = vdw_critical(0.5, 4.0)[1]
    out = []
    for k in range(n):
        R = rmin + u[k, 0] * (L / 4.0 - rmin)
        tr = 0.6 + 0.35 * u[k, 1]
        rv, rl = maxwell(tr * tc, a=0.5, b=4.0)
        out.append(dict(L=L, radius=R, tr=tr, T=tr * tc, droplet=bool(u[k, 2] < droplet_share), rho_v=rv, rho_l=rl,
                        noise_seed=20240601 + seed * 100000 + k))
    return out


def run_warmstart(args):
    """Learned warm start: solver-generated training profiles (2D ensemble cases and 3D
    cases), GPU training of the network, then held-out 3D bubbles run cold and warm to
    the same steady criterion."""
    import multiphaselb_amd as M
    from multiphaselb_amd import warmstart as W
    from multiphaselb_amd.casefiles import CaseSpec
    from multiphaselb_amd.initfields import ensemble, ensemble_density, noise, sphere

    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    L3 = args.ws_size
    thr = args.steady_fp32
    samples_X, samples_Y = [], []
    t_data = time.perf_counter()
    # 2D cases from the ensemble generator, advanced together in one batch
    if args.ws_2d:
        ens = ensemble(args.ws_2d, seed=11)
        hosts = {}
        for c in ens:
            if c["L"] not in hosts:
                d = tempfile.mkdtemp(prefix="mplb_ws2_")
                M.write_case(d, CaseSpec(lbmodel=0, nx=c["L"], ny=c["L"], btype=3, f_model=0, a=0.5, b=4.0,
                                         end=10 ** 8, write_interval=10 ** 8))
                hosts[c["L"]] = M.Case(d, heat_transfer=False, device=-1)
        b = M.Batch([hosts[c["L"]] for c in ens], [ensemble_density(c) for c in ens], [c["T"] for c in ens],
                    precision=M.FP32)
        b.init()
        b.tune(10)
        nconv, steps, delta, state = b.run_to_steady(thr, check_every=100, max_steps=args.max_steps)
        geo = {}
        kept2 = flat2 = 0
        for k, c in enumerate(ens):
            if state[k] != 1:
                continue
            L = c["L"]
            if L not in geo:
                geo[L] = (W.node_radius((L, L), ((L - 1) / 2.0, (L - 1) / 2.0)), W.distinct_sites((L, L)))
            try:
                y, kept = W.steady_profile(b.get(k, "rho"), *geo[L], c["rho_v"], c["rho_l"], c["droplet"])
            except ValueError:
                continue
            kept2 += kept
            flat2 += not kept
            samples_X.append(W.features(c["tr"], c["radius"], L, c["droplet"], 2))
            samples_Y.append(np.concatenate([W.target_from_profile(y, c["droplet"]), np.zeros(W.NPROF)]))
        b.close()
        log("2D: %d/%d steady in %.1f s, steps median %d; targets: %d interfaces, %d dissolved" % (
            nconv, len(ens), time.perf_counter() - t_data, int(np.median(steps)), kept2, flat2))
    d3 = tempfile.mkdtemp(prefix="mplb_ws3_")
    M.write_case(d3, CaseSpec(lbmodel=1, nx=L3, ny=L3, nz=L3, btype=5, f_model=0, f_Lambda=LAM15, f_sigma=0.125,
                              g_model=1, g_Lambda=GLAM15, end=10 ** 8, write_interval=10 ** 8))
    case = M.Case(d3, precision=M.FP32, heat_transfer=False, fused=True, extensions=M.EXT_LIMRT_D3Q15)
    N = case.N
    cen = ((L3 - 1) / 2.0,) * 3
    r3 = W.node_radius((L3, L3, L3), cen)
    sites3 = W.distinct_sites((L3, L3, L3))

    def cold_field(c):
        inside, outside = (c["rho_l"], c["rho_v"]) if c["droplet"] else (c["rho_v"], c["rho_l"])
        rho = sphere((L3, L3, L3), cen, c["radius"], inside, outside, 5.0) + noise((L3, L3, L3), 1e-4, c["noise_seed"])
        return rho.ravel()

    def start(rho=None):
        if rho is not None:
            case.set("rho", rho)
        case.set("U", np.zeros((N, 3)))
        case.set("T", np.full(N, c["T"]))
        case.init()

    t3 = time.perf_counter()
    # training cases: seed 13 over the evaluation's ranges; --ws-3d-rmin > 6 draws another
    # set (seed 19, bubbles only) from the larger radii, where interfaces survive
    if args.ws_3d_rmin > 6.0:
        train3 = bubble3d_cases(args.ws_3d_offset + args.ws_3d, 19, L3, 0.0, args.ws_3d_rmin)[args.ws_3d_offset:]
    else:
        train3 = bubble3d_cases(args.ws_3d_offset + args.ws_3d, 13, L3, 0.5)[args.ws_3d_offset:]
    kept3 = flat3 = moving3 = 0
    plateau_rows = []          # (T/Tc, R, droplet, outside and inside plateaus) of steady 3D interfaces
    for i, c in enumerate(train3):
        start(cold_field(c))
        st, wall, ok = run_to_steady_case(case, c["rho_l"], thr, 100, args.max_steps)
        y, kept = None, False
        rho = case.get("rho")
        if np.isfinite(rho).all():
            try:
                y, kept = W.steady_profile(rho, r3, sites3, c["rho_v"], c["rho_l"], c["droplet"])
            except ValueError:
                y = None
            # a run still moving at the step limit is used only for an interface profile
            # (its near-steady state); a uniform target needs a steady run
            if y is not None and not ok and not kept:
                y = None
        if y is not None:
            kept3 += kept and ok
            moving3 += kept and not ok
            flat3 += not kept
            if kept and ok:
                _, Rs = W.extract_profile(rho, r3, c["rho_v"], c["rho_l"], c["droplet"])
                plateau_rows.append((c["tr"], Rs, float(c["droplet"]), float(y[-1]), float(y[0])))
            samples_X.append(W.features(c["tr"], c["radius"], L3, c["droplet"], 3))
            samples_Y.append(np.concatenate([W.target_from_profile(y, c["droplet"]), np.zeros(W.NPROF)]))
        if i % 16 == 0:
            log("3D train case %d: %d steps %.2f s steady %s" % (i, st, wall, ok))
    log("3D training runs: %.1f s; targets: %d interfaces (+%d still moving at the step limit), %d dissolved" % (
        time.perf_counter() - t3, kept3, moving3, flat3))
    X = np.array(samples_X, np.float32).reshape(-1, W.N_IN)
    Y = np.array(samples_Y, np.float32).reshape(-1, W.N_OUT)
    P = np.array(plateau_rows, np.float64).reshape(-1, 5)
    loaded = {}
    if args.ws_load_data:
        # profiles from earlier solver runs of this workload (same generators, saved with
        # --ws-save-data by runs over other case indices)
        for path in args.ws_load_data.split(","):
            z = np.load(path)
            X = np.concatenate([X, z["X"]])
            Y = np.concatenate([Y, z["Y"]])
            if "P" in z.files and z["P"].shape[1] == 5:
                P = np.concatenate([P, z["P"]])
            loaded[os.path.basename(path)] = int(z["X"].shape[0])
    if args.ws_save_data:
        np.savez(args.ws_save_data, X=X, Y=Y, P=P)
    if args.ws_eval == 0:
        print(json.dumps({"saved_samples": int(X.shape[0]), "path": args.ws_save_data}))
        return 0
    # the 3D profiles are a few percent of the samples: repeat them so that the
    # minibatches see the evaluation's dimension often enough
    is3 = X[:, 7] > 0
    if args.ws_3d_repeat > 1 and is3.any():
        X = np.concatenate([X] + [X[is3]] * (args.ws_3d_repeat - 1))
        Y = np.concatenate([Y] + [Y[is3]] * (args.ws_3d_repeat - 1))
    t_train = time.perf_counter()
    net = M.Mlp(W.WIDTHS, args.ws_batch)
    net.set_params(W.init_params(W.WIDTHS, 0))
    hist = W.train(net, X, Y, iters=args.ws_iters, seed=1)
    # second network: the interface profiles alone (the first one decides flat or not)
    keep = ~W.flat_targets(X, Y)
    # (--ws-kept-nets of them from different seeds; the proposal is their mean)
    net_k = []
    for k in range(args.ws_kept_nets):
        nk = M.Mlp(W.WIDTHS, args.ws_batch)
        nk.set_params(W.init_params(W.WIDTHS, 1 + 10 * k))
        hist_k = W.train(nk, X[keep], Y[keep], iters=args.ws_iters_kept, seed=2 + 10 * k)
        net_k.append(nk)
    t_train = time.perf_counter() - t_train
    log("trained %d samples (%d interface profiles), 1 + %d x %d params, %d + %d iters in %.1f s, loss %.3e -> %.3e "
        "(interfaces, last network: %.3e -> %.3e)" % (
            X.shape[0], int(keep.sum()), args.ws_kept_nets, net.P, args.ws_iters, args.ws_iters_kept, t_train,
            hist[:20].mean(), hist[-50:].mean(), hist_k[:20].mean(), hist_k[-50:].mean()))
    # the outside plateau from the steady 3D interfaces (a second model; see warmstart.py)
    pfit, prms = W.fit_plateau(P[:, 0], P[:, 1], P[:, 2] > 0, P[:, 3]) if len(P) else ({}, {})
    pfit_in, prms_in = W.fit_plateau(P[:, 0], P[:, 1], P[:, 2] > 0, P[:, 4]) if len(P) else ({}, {})
    log("plateau fits: %d steady 3D interfaces (%d droplets), rms residual outside %s / %s, inside %s / %s "
        "(droplets / bubbles, of the jump)" % (len(P), int((P[:, 2] > 0).sum()) if len(P) else 0, prms.get(True),
                                              prms.get(False), prms_in.get(True), prms_in.get(False)))
    evals = bubble3d_cases(args.ws_eval, 17, L3, args.ws_eval_droplets)
    inner, outer = r3 < 2.0, r3 > r3.max() - 2.0
    rows = []
    for i, c in enumerate(evals):
        rho0 = cold_field(c)
        mass = float(rho0[sites3].sum())
        start(rho0)
        st_c, wall_c, ok_c = run_to_steady_case(case, c["rho_l"], thr, 100, args.max_steps)
        rc = case.get("rho")
        finite_c = bool(np.isfinite(rc).all())
        cin, cout = rc[inner].mean(), rc[outer].mean()
        kept = finite_c and abs(cin - cout) >= 0.25 * (c["rho_l"] - c["rho_v"])
        # warm: proposal -> mass-matched raster -> the solver to its own criterion; a
        # proposal outside the physical range, or a warm run that diverges, falls back to
        # the cold start, whose time is then added to the warm total.  Two rasters of the
        # same proposal: with the fitted outside plateau (the method) and the network's own
        # (for comparison); a uniform proposal is the same for both.
        tw = time.perf_counter()
        prof, flat = W.propose(net, net_k, W.features(c["tr"], c["radius"], L3, c["droplet"], 3), c["droplet"],
                               c["rho_v"], c["rho_l"])
        used = W.proposal_ok(prof, c["rho_v"], c["rho_l"])
        flat = used and flat
        t_prop = time.perf_counter() - tw
        res = {}
        for variant in ("plateau", "net"):
            if variant == "net" and (flat or not used):
                res["net"] = res["plateau"]
                continue
            # timed: the proposal and its raster (the set-up of U, T and the first collision
            # that follows is the cold run's set-up too, untimed in both)
            tr0 = time.perf_counter()
            Rw = None
            ok_raster = used
            if used:
                if variant == "plateau" and not flat:
                    y = (np.asarray(prof, np.float64) - c["rho_v"]) / (c["rho_l"] - c["rho_v"])
                    Rw = W.raster_case_plateau(case, y, pfit, c["tr"], c["droplet"], c["rho_v"], c["rho_l"], cen, mass,
                                               c["radius"], int(sites3.sum()), fit_in=pfit_in)
                else:
                    Rw = W.raster_case(case, prof, cen, mass, c["radius"], int(sites3.sum()), uniform=flat)
            t_r = time.perf_counter() - tr0
            if used:
                start(None)
                ok_raster = bool(np.isfinite(case.get("rho")).all())
            fallback = not ok_raster
            st_w = wall_w = 0.0
            ok_w = False
            if ok_raster:
                st_w, wall_w, ok_w = run_to_steady_case(case, c["rho_l"], thr, 100, args.max_steps)
                fallback = not ok_w and not np.isfinite(case.get("rho")).all()
            if fallback:
                start(rho0)
                s2, w2, ok_w = run_to_steady_case(case, c["rho_l"], thr, 100, args.max_steps)
                st_w, wall_w = st_w + s2, wall_w + w2
            wall_w += t_prop + t_r
            rw = case.get("rho")
            win, wout = rw[inner].mean(), rw[outer].mean()
            dev = float(max(abs(win - cin) / cin, abs(wout - cout) / cout)) if finite_c else float("nan")
            res[variant] = dict(steps=st_w, wall=wall_w, ok=ok_w, fallback=fallback, dev=dev, R=Rw)
        w, wn = res["plateau"], res["net"]
        perr = None
        if kept and used and not flat and ok_c:
            # the proposal against the cold run's steady radial profile (normalised): inside
            # plateau, interface (|d| <= 6) and outside plateau, and the interface radius
            yt, Rc = W.extract_profile(rc, r3, c["rho_v"], c["rho_l"], c["droplet"])
            yp = (np.asarray(prof, np.float64) - c["rho_v"]) / (c["rho_l"] - c["rho_v"])
            ypl = W.plateau_profile(yp, pfit, pfit_in, c["tr"], w["R"] if w["R"] else c["radius"], c["droplet"])
            band = np.abs(W.DK) <= 6.0
            perr = {"in": float(ypl[0] - yt[0]), "out": float(ypl[-1] - yt[-1]),
                    "out_net": float(yp[-1] - yt[-1]), "in_net": float(yp[0] - yt[0]),
                    "shape_rms": float(np.sqrt(np.mean((ypl[band] - yt[band]) ** 2))),
                    "R": float((w["R"] or 0.0) - Rc)}
        rows.append(dict(steps_cold=st_c, steps_warm=w["steps"], wall_cold=wall_c, wall_warm=w["wall"], ok_cold=ok_c,
                         ok_warm=w["ok"], finite_cold=finite_c, kept=kept, droplet=c["droplet"], dev=w["dev"],
                         fallback=w["fallback"], flat=flat, prop_s=t_prop, steps_net=wn["steps"], wall_net=wn["wall"],
                         ok_net=wn["ok"], dev_net=wn["dev"], perr=perr))
        log("eval %d (%s R %.1f T/Tc %.3f): cold %d steps %.2f s%s, warm %d steps %.2f s%s%s (%s), network-only "
            "plateau %d steps, cold in/out %.5f %.5f%s, dev %.1e / %.1e" % (
                i, "drop" if c["droplet"] else "bubl", c["radius"], c["tr"], st_c, wall_c, "" if ok_c else " (not steady)",
                w["steps"], w["wall"], "" if w["ok"] else " (not steady)", " (fell back to cold)" if w["fallback"] else "",
                "uniform" if flat else ("R %.2f" % w["R"] if w["R"] is not None else "not used"), wn["steps"],
                cin, cout, "" if kept else " (interface gone)", w["dev"], wn["dev"]))
        if perr:
            log("   proposal vs steady profile: inside %+.1e (network %+.1e), outside %+.1e (network %+.1e), "
                "interface rms %.1e, radius %+.3f" % (perr["in"], perr["in_net"], perr["out"], perr["out_net"],
                                                      perr["shape_rms"], perr["R"]))
    # population: held-out cases whose cold run stays finite (the plain solver's result
    # exists); a cold run still moving at the step limit counts with the limit (a lower
    # bound of its time).  Agreement is judged where both runs met the steady criterion.
    pop = [r for r in rows if r["finite_cold"]]
    both = [r for r in pop if r["ok_cold"] and r["ok_warm"]]

    def ratio(rs, key):
        return round(float(np.sum([r[key + "_cold"] for r in rs]) / np.sum([r[key + "_warm"] for r in rs])), 3) if rs else None

    def per_case(rs, key):
        v = np.array([r[key + "_cold"] / r[key + "_warm"] for r in rs]) if rs else None
        return {"median": round(float(np.median(v)), 3), "min": round(float(v.min()), 3),
                "max": round(float(v.max()), 3)} if rs else None

    devs = np.array([r["dev"] for r in both]) if both else None
    groups = {"interface_kept": [r for r in pop if r["kept"]], "dissolved": [r for r in pop if not r["kept"]]}
    out = {
        "metric": "time-to-steady-state NN-warm vs plain",
        "value": ratio(pop, "wall"),
        "unit": "x (total cold wall time / total warm wall time, proposal and fallbacks included)",
        "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (solver), bf16 MFMA / f32 accumulate (network)",
        "data": "synthetic: training profiles from solver runs (2D ensemble + 3D cases), held-out 3D cases",
        "config": {"workload": "D3Q15 liMRT static %s %d^3 periodic FP32, R in [6, %d], T/Tc in [0.6, 0.95]" % (
            "bubbles" if args.ws_eval_droplets == 0 else "droplets/bubbles", L3, L3 // 4),
                   "steady": "max|drho|/rho_l over 100 steps < %g" % thr, "max_steps": args.max_steps,
                   "train_2d": args.ws_2d, "train_3d": args.ws_3d, "loaded_samples": loaded, "eval": len(evals),
                   "network": {"widths": W.WIDTHS, "params": net.P, "networks": 1 + args.ws_kept_nets,
                               "roles": "all steady states (dissolves or not); interface profiles (mean of %d)" %
                                        args.ws_kept_nets,
                               "iters": [args.ws_iters, args.ws_iters_kept], "batch": args.ws_batch,
                               "3d_repeat": args.ws_3d_repeat}},
        "cases": {"eval": len(rows), "cold_unstable": len(rows) - len(pop), "compared": len(pop),
                  "cold_steady": int(sum(r["ok_cold"] for r in pop)), "warm_steady": int(sum(r["ok_warm"] for r in pop)),
                  "both_steady": len(both), "fallbacks": int(sum(r["fallback"] for r in pop)),
                  "interface_kept_cold": len(groups["interface_kept"]),
                  "proposed_uniform": int(sum(r["flat"] for r in pop)),
                  "proposed_uniform_but_kept_cold": int(sum(r["flat"] and r["kept"] for r in pop)),
                  "proposed_interface_but_dissolved_cold": int(sum(not r["flat"] and not r["kept"] for r in pop))},
        "speedup_wall": {k: ratio(g, "wall") for k, g in groups.items()},
        "speedup_steps": {"all": ratio(pop, "steps"), **{k: ratio(g, "steps") for k, g in groups.items()}},
        "speedup_steps_per_case": per_case(pop, "steps"),
        "network_plateau_only": {
            "speedup_wall": {k: round(float(np.sum([r["wall_cold"] for r in g]) / np.sum([r["wall_net"] for r in g])), 3)
                             if g else None for k, g in groups.items()},
            "speedup_steps": {k: round(float(np.sum([r["steps_cold"] for r in g]) / np.sum([r["steps_net"] for r in g])), 3)
                              if g else None for k, g in groups.items()},
            "dev_max": float(max([r["dev_net"] for r in pop if r["ok_cold"] and r["ok_net"]], default=0.0))},
        "proposal_error_kept": {k: round(float(np.median(np.abs([r["perr"][k] for r in pop if r["perr"]]))), 8)
                                for k in ("in", "in_net", "out", "out_net", "shape_rms", "R")}
        if any(r["perr"] for r in pop) else None,
        "plateau_fit": {"samples": int(len(P)), "rms_outside_droplets": prms.get(True),
                        "rms_outside_bubbles": prms.get(False), "rms_inside_droplets": prms_in.get(True),
                        "rms_inside_bubbles": prms_in.get(False)},
        "final_state_rel_dev": {"median": float(np.median(devs)), "max": float(devs.max()),
                                "failures_over_1e-4": int((devs > 1e-4).sum()),
                                **{k + "_max": float(max([r["dev"] for r in g if r in both], default=0.0))
                                   for k, g in groups.items()}} if both else None,
        "train_samples": int(X.shape[0]), "train_s": round(t_train, 2),
        "data_s": round(time.perf_counter() - t_data, 1),
    }
    print(json.dumps(out))
    return 0


def run_accuracy(args):
    """Reduced precision against FP64 on the configurations' own sizes.  The FP64
    per-routine path is bit-identical to the oracle (1000-step parity tests), so it
    stands in for the CPU oracle at sizes the oracle cannot finish in minutes.
    C1: 256^3 static bubble, 2000 steps: coexistence densities (inside r < R/2, outside
    r > R + 20), Laplace pressure jump between the same regions, spurious currents (max
    |U|).  C2: 128 x 128 x 384 thermal column (gravity -1e-5, pool in the lower third,
    bounceBack floor and ceiling, fixedRandomT floor at 0.95 Tc +-0.5%, fixedT ceiling at
    0.8 Tc): liquid and vapour densities (z in [20, 100] and [160, 360]), the
    plane-averaged temperature profile, max |U|.  With the examples' Cv = 1 the thermally
    coupled column goes unstable in FP64 (the reference arithmetic) after about a thousand
    steps: the heat source's pressure-work term, applied one step late, amplifies long
    sound waves in the liquid when (dp/dT)/(rho Cv) is large (2.8 at Cv = 1).  The column
    therefore runs with --acc-c2-cv (4 by default, stable) for --acc-c2-steps steps."""
    import multiphaselb_amd as M
    from multiphaselb_amd.casefiles import BC, CaseSpec
    from multiphaselb_amd.initfields import maxwell, noise, vdw_critical

    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    tc = vdw_critical(0.5, 4.0)[1]
    modes = [("fp64", M.FP64, False), ("fp32", M.FP32, True), ("fp16s", M.FP16S, True), ("fp16n", M.FP16N, True)]

    def run(d, N, rho, T, prec, fused, ht, steps):
        t0 = time.perf_counter()
        c = M.Case(d, precision=prec, heat_transfer=ht, fused=fused, extensions=M.EXT_LIMRT_D3Q15)
        c.set("rho", rho)
        c.set("U", np.zeros((N, 3)))
        c.set("T", T)
        c.init()
        done = c.step(steps)
        out = {k: np.asarray(c.get(k)) for k in ("rho", "U", "p", "T")}
        c.close()
        return done, out, time.perf_counter() - t0

    def rel(a, b):
        return float(abs(a - b) / abs(b)) if b != 0 else float("nan")

    res = {}
    # C1
    n, steps1 = args.acc_size, 2000
    d, spec, rho, T, dens = make_case(n, False, steps1)
    spec.write_interval = steps1
    M.write_case(d, spec)
    N = n ** 3
    c0 = (n - 1) / 2.0
    R = 40.0 * n / 256.0
    idx = np.indices((n, n, n)).reshape(3, -1)[::-1]
    r = np.sqrt(((idx - c0) ** 2).sum(0))
    inner, outer = r < 0.5 * R, r > R + 20.0 * n / 256.0
    obs = {}
    for name, prec, fused in modes:
        done, f, wall = run(d, N, rho, np.full(N, T), prec, fused, False, steps1)
        um = np.sqrt((f["U"] ** 2).sum(1))
        obs[name] = dict(rho_in=float(f["rho"][inner].mean()), rho_out=float(f["rho"][outer].mean()),
                         dp=float(f["p"][inner].mean() - f["p"][outer].mean()), umax=float(um.max()),
                         finite=bool(np.isfinite(f["rho"]).all()), steps=int(done), wall_s=round(wall, 1))
        log("C1 %s: %s" % (name, obs[name]))
    ref = obs["fp64"]
    res["c1_bubble"] = {"size": [n, n, n], "steps": steps1, "fp64": ref,
                        **{k: {**obs[k], "rel_err_vs_fp64": {q: rel(obs[k][q], ref[q]) for q in
                                                             ("rho_in", "rho_out", "dp", "umax")}}
                           for k in ("fp32", "fp16s", "fp16n")}}
    # C2
    nx, nz, steps2 = args.acc_size // 2, (3 * args.acc_size) // 2, args.acc_c2_steps
    rv, rl = maxwell(0.8 * tc, a=0.5, b=4.0)
    spec2 = CaseSpec(lbmodel=1, nx=nx, ny=nx, nz=nz, btype=3, f_model=0, f_Lambda=LAM15, f_sigma=0.125,
                     g_model=1, g_Lambda=GLAM15, g_alpha1=0.25, g_alpha2=0.25, g=(0.0, 0.0, -1e-5),
                     Cv=args.acc_c2_cv, end=10 ** 8, write_interval=steps2,
                     f_bc={"Z0": BC("bounceBack"), "Z1": BC("bounceBack")},
                     g_bc={"Z0": BC("fixedRandomT", (0.95 * tc,), pert=0.5, duration=10 ** 8),
                           "Z1": BC("fixedT", (0.8 * tc,))})
    d2 = tempfile.mkdtemp(prefix="mplb_acc_")
    M.write_case(d2, spec2)
    N2 = nx * nx * nz
    k = np.arange(nz)
    prof = 0.5 * (rl + rv) + 0.5 * (rv - rl) * np.tanh(2.0 * (k - nz / 3.0) / 5.0)
    rho2 = (np.broadcast_to(prof[:, None, None], (nz, nx, nx)) + noise((nz, nx, nx), 1e-4, 20240602,
                                                                       (True, True, False))).ravel()
    zi = np.repeat(k, nx * nx)
    liq, vap = (zi >= 20 * nz // 384) & (zi < 100 * nz // 384), (zi >= 160 * nz // 384) & (zi < 360 * nz // 384)
    obs2, tprof = {}, {}
    for name, prec, fused in modes:
        done, f, wall = run(d2, N2, rho2, np.full(N2, 0.8 * tc), prec, fused, True, steps2)
        um = np.sqrt((f["U"] ** 2).sum(1))
        tprof[name] = f["T"].reshape(nz, -1).mean(1)
        obs2[name] = dict(rho_liquid=float(f["rho"][liq].mean()), rho_vapour=float(f["rho"][vap].mean()),
                          umax=float(um.max()), finite=bool(np.isfinite(f["rho"]).all()), steps=int(done),
                          wall_s=round(wall, 1))
        log("C2 %s: %s" % (name, obs2[name]))
    ref2 = obs2["fp64"]
    res["c2_column"] = {"size": [nx, nx, nz], "steps": steps2, "Cv": args.acc_c2_cv, "fp64": ref2,
                        **{k: {**obs2[k], "rel_err_vs_fp64": {
                            **{q: rel(obs2[k][q], ref2[q]) for q in ("rho_liquid", "rho_vapour", "umax")},
                            "T_profile_max": float(np.max(np.abs(tprof[k] - tprof["fp64"]) / tprof["fp64"]))}}
                           for k in ("fp32", "fp16s", "fp16n")}}
    # F4: the surface-tension tensor and the other equations of state on the C1 bubble
    res["f4_variants"] = accuracy_variants(args, log, run, rel)
    out = {"metric": "reduced precision vs FP64 on the configuration sizes", "value": None, "unit": "relative error",
           "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32, f16-storage/f32 vs f64", "data": "synthetic",
           "config": {"workload": "C1 bubble %d^3 2000 steps; C2 thermal column %dx%dx%d %d steps, Cv=%g; "
                                  "C1 with xuSurfTen, Carnahan-Starling, Peng-Robinson" % (
               n, nx, nx, nz, steps2, args.acc_c2_cv)}, **res}
    print(json.dumps(out))
    return 0


def copy_probe(mib, reps=5):
    """A plain device-to-device copy of known size; its FETCH_SIZE / WRITE_SIZE counter
    readings give the scale factor applied to the lattice kernels' counters."""
    import torch

    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "reps": reps}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "fp16s", "fp16n"])
    ap.add_argument("--thermal", action="store_true")
    ap.add_argument("--cv", type=float, default=4.0,
                    help="heat capacity Cv of the thermal workload (the examples' Cv = 1 makes the coupled "
                         "model's liquid unstable to long waves within ~2000 steps at 256^3, also in FP64)")
    ap.add_argument("--generic", action="store_true", help="per-routine kernels instead of the fused kernel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-thermal-record", action="store_true",
                    help="default workload: skip the thermal sub-record")
    ap.add_argument("--no-ensemble-record", action="store_true",
                    help="default workload: skip the batched-ensemble sub-record")
    ap.add_argument("--workload", default="bubble3d", choices=["bubble3d", "ensemble", "warmstart", "accuracy"])
    ap.add_argument("--acc-size", type=int, default=256, help="accuracy: bubble edge (the column is n/2 x n/2 x 3n/2)")
    ap.add_argument("--acc-c2-steps", type=int, default=5000, help="accuracy: column steps compared")
    ap.add_argument("--acc-c2-cv", type=float, default=4.0, help="accuracy: heat capacity of the column")
    ap.add_argument("--ws-2d", type=int, default=4096, help="warmstart: 2D training cases")
    ap.add_argument("--ws-3d", type=int, default=256, help="warmstart: 3D training cases")
    ap.add_argument("--ws-eval", type=int, default=256, help="warmstart: held-out 3D bubbles")
    ap.add_argument("--ws-size", type=int, default=128)
    ap.add_argument("--ws-eval-droplets", type=float, default=0.0, help="warmstart: share of droplets among the eval cases")
    ap.add_argument("--ws-iters", type=int, default=4000, help="warmstart: Adam iterations of the all-cases network")
    ap.add_argument("--ws-iters-kept", type=int, default=16000, help="warmstart: iterations of the interface network")
    ap.add_argument("--ws-kept-nets", type=int, default=3, help="warmstart: interface networks (the proposal is their mean)")
    ap.add_argument("--ws-batch", type=int, default=256)
    ap.add_argument("--ws-3d-repeat", type=int, default=4, help="warmstart: times each 3D profile enters the training set")
    ap.add_argument("--ws-3d-offset", type=int, default=0, help="warmstart: first 3D training case index")
    ap.add_argument("--ws-3d-rmin", type=float, default=6.0,
                    help="warmstart: smallest training radius (> 6: a separate set of larger bubbles)")
    ap.add_argument("--ws-save-data", default="", help="warmstart: save the training profiles (npz)")
    ap.add_argument("--ws-load-data", default="", help="warmstart: add training profiles from npz files")
    ap.add_argument("--cases", type=int, default=1024, help="ensemble: number of cases")
    ap.add_argument("--ens-precisions", default="fp32,fp16s,fp16n", help="ensemble: precisions besides fp64")
    ap.add_argument("--steady-fp64", type=float, default=1e-9)
    ap.add_argument("--steady-fp32", type=float, default=5e-6)
    ap.add_argument("--steady-fp16s", type=float, default=1e-4)
    ap.add_argument("--steady-fp16n", type=float, default=1e-4)
    ap.add_argument("--max-steps", type=int, default=200000)
    ap.add_argument("--tune-steps", type=int, default=20)
    ap.add_argument("--persistent", action="store_true",
                    help="ensemble: each 100-step run as one persistent cooperative launch of the single pass")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM traffic counters)")
    ap.add_argument("--pattern-probe", action="store_true",
                    help="only run the memory-pattern probe of the single pass (streaming ceilings)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.pattern_probe:
        return pattern_probe(args)
    if args.gpus != 1:
        print("bench.py runs on one GPU: the whole lattice lives on a single MI355X "
              "(multi-GPU sharding is a design note, not built); got --gpus %d" % args.gpus, file=sys.stderr)
        return 2
    if args.workload == "ensemble":
        return run_ensemble(args)
    if args.workload == "warmstart":
        return run_warmstart(args)
    if args.workload == "accuracy":
        return run_accuracy(args)

    out = measure(args, args.thermal)
    if not args.thermal and not args.no_thermal_record and args.precision != "fp64":
        # the thermal step (xuMRT + myMRT with heat transfer) on the same bubble, timed the
        # same way: the configuration the >= 70% roofline target is stated for
        t = measure(args, True)
        out["thermal"] = {k: t[k] for k in ("value", "unit", "ms_per_step", "config", "hbm_gbs_effective",
                                            "step_roofline_frac", "roofline", "kernels_ms", "finite")}
    if not args.thermal and not args.no_ensemble_record and args.precision != "fp64":
        # the batched-ensemble step of C3, timed the same way (a short fixed-step run)
        out["ensemble"] = ensemble_record(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    if out["roofline"]["achieved"]:
        out["copy_ceiling"] = copy_ceiling(out["roofline"]["achieved"], args.size)
    print(json.dumps(out))
    return 0


def copy_ceiling(achieved_gbs, size):
    """A flat copy of the 15 population planes of an size^3 lattice (read + write counted;
    multiphaselb_amd/stream_probe, run as a child process so that its HIP context is its
    own): the streaming rate the same box reaches without any stencil, next to the
    dominant kernel's algorithmic rate.  The probe times two copies (grid-stride over 8
    workgroups per CU; one float4 per work item over the whole grid); the faster one is
    the ceiling."""
    exe = os.path.join(ROOT, "multiphaselb_amd", "stream_probe")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe, str(size), "copy"], capture_output=True, text=True, timeout=120)
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    rows = [w for w in rows if w["pattern"].startswith("copy")]
    if r.returncode != 0 or not rows:
        return None
    best = max(rows, key=lambda w: w["GBs"])
    gbs = best["GBs"]
    return {"GBs": gbs, "pattern": best["pattern"], "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4),
            "bytes": 120 * size ** 3, "kernel_over_copy": round(achieved_gbs / gbs, 4),
            "copies": {w["pattern"]: w["GBs"] for w in rows}}


def pattern_probe(args):
    """Runs the memory-pattern probe (multiphaselb_amd/stream_probe, built with the library)
    on an --size^3 lattice: two flat copies of the 15 population planes, the single pass's
    pull pattern without and with the ring re-pulls, at 1-3 z chunks; no collision
    arithmetic."""
    exe = os.path.join(ROOT, "multiphaselb_amd", "stream_probe")
    if not os.path.exists(exe):
        print("stream_probe is not built; run `python -m multiphaselb_amd.build`", file=sys.stderr)
        return 2
    r = subprocess.run([exe, str(args.size)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        return r.returncode
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    print(json.dumps({"metric": "memory-pattern ceilings of the D3Q15 single pass", "unit": "GB/s",
                      "bytes_per_node": 120, "size": args.size, "patterns": rows}))
    return 0


def measure(args, thermal):
    r = run_gpu(args, thermal)
    N, steps = r["N"], args.steps
    sec = r["total_ms"] / 1e3
    mlups = N * steps / sec / 1e6
    blu = bytes_per_lu(args.precision, thermal)
    # Dominant kernel: the fused collide pass (or the collision kernel of the per-routine
    # path).  Its algorithmic bytes per node: populations pulled once and written once,
    # plus the macroscopic values it reads or writes once (psi; thermal: T, S, U).
    s_pdf = {"fp64": 8, "fp32": 4, "fp16s": 2, "fp16n": 2}[args.precision]
    s_mac = 8 if args.precision == "fp64" else 4
    kname = "fused" if "fused" in r["kern"] else "collide"
    k_ms, k_launches = r["kern"].get(kname, (0.0, 0))
    # isothermal periodic D3Q15 (FP32 or FP16 storage) runs the single-pass kernel (psi stays
    # on chip): one timed kernel, where the two-pass kernels also report their first pass
    single = kname == "fused" and not thermal and "fused_aux" not in r["kern"]
    # thermal periodic D3Q15 with FP32 storage runs the thermal single pass (one timed
    # kernel: f and g pulled once, T and U of the previous step read once)
    tsingle = kname == "fused" and thermal and "fused_aux" not in r["kern"]
    s_ref = 2 * 2 if args.precision == "fp16n" else 0             # FP16N: node reference in/out
    if kname == "collide":
        per_unit = 2 * 15 * s_mac + (1 + 3 + 3 + 1) * s_mac      # f in/out; rho, U, Fi, psi in
    elif tsingle:
        per_unit = 4 * 15 * s_pdf + 2 * (1 + 3) * s_mac + s_ref   # f, g in/out; T, U in and out
    elif thermal:
        per_unit = 4 * 15 * s_pdf + (1 + 1 + 1 + 3) * s_mac + s_ref   # f, g in/out; psi, T, S in; U out
    elif single:
        per_unit = 2 * 15 * s_pdf + s_ref                         # f in/out
    else:
        per_unit = 2 * 15 * s_pdf + s_mac + s_ref                 # f in/out; psi in
    achieved = per_unit * N / (k_ms / 1e3) / 1e9 if k_ms > 0 else None
    step_gbs = mlups * 1e6 * blu / 1e9
    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3 counter
    # runs of this workload (profiles/fused_traffic.json), when one exists for it
    traffic = None
    staged = os.environ.get("MPLB_MARCH_LDS") == "1" and args.precision == "fp32"
    ahead = (not staged and os.environ.get("MPLB_MARCH_PF") == "1" and args.precision in ("fp32", "fp16s"))
    march_name = "k_fused_march_lds" if staged else ("k_fused_march_pf" if ahead else "k_fused_march")
    variant = "_staged" if staged else ("_ahead" if ahead else "")
    key = "d3q15_%d_%s_%s" % (args.size, args.precision,
                              ("thermal_single" if tsingle else "thermal") if thermal else
                              ("iso" + variant if single else "iso_twopass"))
    tpath = os.path.join(ROOT, "profiles", "fused_traffic.json")
    if kname == "fused" and os.path.exists(tpath):
        rec = json.load(open(tpath)).get("kernels", {}).get(key)
        if rec:
            traffic = rec["hbm_bytes_per_launch"]
    return {
        "metric": "D3Q15 pseudopot-MRT MLUPS",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["total_ms"] / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16s": "f16-storage/f32",
                  "fp16n": "f16-node-ref-storage/f32"}[args.precision],
        "data": "synthetic (seeded bubble, Maxwell densities %.5f/%.5f)" % r["dens"],
        "config": {"workload": "D3Q15 %s static bubble %d^3 periodic, VdW T=0.8Tc, R=%g%s" % (
            "liMRT+myMRT thermal" if thermal else "liMRT", args.size, 40.0 * args.size / 256.0,
            ", Cv=%g" % args.cv if thermal else ""),
            "precision": args.precision, "path": "fused" if r["fused"] else "per-routine",
            "bytes_per_lu": blu},
        "hbm_gbs_effective": round(step_gbs, 1),
        "step_roofline_frac": round(step_gbs / PEAK_HBM_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None,
                     "traffic": traffic, "kernel": (march_name if single else ("k_fused_tmarch" if tsingle else
                                                                    ("k_fused_collide" if kname == "fused" else "k_collide_*"))),
                     "kernel_ms": round(k_ms, 5),
                     "bytes_per_unit": per_unit, "units_per_launch": N},
        "kernels_ms": {k: round(v[0], 5) for k, v in r["kern"].items()},
        "finite": r["finite"],
    }


if __name__ == "__main__":
    sys.exit(main())

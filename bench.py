#!/usr/bin/env python3
"""Benchmark of the lattice update on one MI355X.

Workload (default): D3Q15 pseudopotential MRT static vapour bubble in liquid, 256^3
periodic box, van der Waals EOS at T = 0.8 Tc, bubble radius 40 with a tanh interface
of width 5 between the Maxwell-construction densities, seeded +-1e-4 noise, FP32.
The f field uses the D3Q15 MRT scheme (Collision liMRT mapped onto the reference's
D3Q15 xuMRT scheme); --thermal adds the myMRT energy field with heat transfer.

One "step" is one multiPhasePP time step over the whole lattice; the metric is MLUPS
= stored nodes x steps / time / 1e6, timed with HIP events on the library's stream
after W warm-up steps, inputs resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
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
    s_pdf = {"fp64": 8, "fp32": 4, "fp16s": 2}[precision]
    s_mac = 8 if precision == "fp64" else 4
    if thermal:
        return 2 * q * s_pdf * 2 + 2 * s_mac * 5
    return 2 * q * s_pdf + 2 * s_mac * 1


def make_case(n, thermal, steps_total):
    import multiphaselb_amd as M
    from multiphaselb_amd.casefiles import CaseSpec
    from multiphaselb_amd.initfields import maxwell, noise, sphere, vdw_critical

    rc, tc = vdw_critical(0.5, 4.0)
    T = 0.8 * tc
    rv, rl = maxwell(T, a=0.5, b=4.0)
    spec = CaseSpec(lbmodel=1, nx=n, ny=n, nz=n, btype=5, f_model=0, f_Lambda=LAM15, f_sigma=0.125,
                    g_model=1, g_Lambda=GLAM15, g_alpha1=0.25, g_alpha2=0.25,
                    end=10 ** 8, write_interval=10 ** 8)
    d = tempfile.mkdtemp(prefix="mplb_bench_")
    M.write_case(d, spec)
    shape = (n, n, n)
    c0 = (n - 1) / 2.0
    radius = 40.0 * n / 256.0
    rho = sphere(shape, (c0, c0, c0), radius, rv, rl, 5.0) + noise(shape, 1e-4, 20240601 + 1)
    return d, spec, rho.ravel(), T, (rv, rl)


def run_gpu(args):
    import multiphaselb_amd as M

    prec = {"fp64": M.FP64, "fp32": M.FP32, "fp16s": M.FP16S}[args.precision]
    d, spec, rho, T, dens = make_case(args.size, args.thermal, args.steps + args.warmup)
    case = M.Case(d, precision=prec, heat_transfer=args.thermal, fused=not args.generic,
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


def cpu_baseline(args, budget_s=12.0):
    """The oracle (reference-faithful FP64 C restatement, OpenMP over nodes) on the
    host cores: D3Q15 xuMRT bubble at 64^3, as many steps as fit the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from _cases import oracle_params

    n = 64
    d, spec, rho, T, _ = make_case(n, args.thermal, 0)
    spec.f_model = 2        # the D3Q15 MRT scheme liMRT is mapped onto
    threads = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", threads))
    threads = max(1, min(threads, cap))
    o = O.Oracle(oracle_params(spec, ht=args.thermal, nthreads=threads))
    N = n ** 3
    o.set("rho", rho)
    o.set("U", np.zeros((N, 3)))
    o.set("T", np.full(N, T))
    o.init()
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 2000:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    return dict(value=round(N * steps / dt / 1e6, 3), unit="MLUPS", cores=threads, kind="port",
                sample="D3Q15 %s bubble 64^3 FP64 oracle, %d steps in %.1f s" %
                       ("xuMRT+myMRT thermal" if args.thermal else "xuMRT", steps, dt))


def laplace(rho, p, L):
    """Interior/exterior density and pressure (2x2 nodes around the box centre, the box
    corner), the equivalent radius from the mass on the L-1 periodic sites, and the
    Laplace product dp * R of one steady case."""
    r = rho.reshape(L, L)[: L - 1, : L - 1]
    pp = p.reshape(L, L)
    c = (L - 1) // 2
    rin = float(rho.reshape(L, L)[c:c + 2, c:c + 2].mean())
    pin = float(pp[c:c + 2, c:c + 2].mean())
    rout, pout = float(r[0, 0]), float(pp[0, 0])
    area = (float(r.sum()) - rout * (L - 1) ** 2) / (rin - rout)
    R = float(np.sqrt(max(area, 0.0) / np.pi))
    return dict(rho_in=rin, rho_out=rout, dp=pin - pout, R=R, gamma=(pin - pout) * R)


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
    thr = {"fp64": args.steady_fp64, "fp32": args.steady_fp32, "fp16s": args.steady_fp16s}
    code = {"fp64": M.FP64, "fp32": M.FP32, "fp16s": M.FP16S}
    precs = ["fp64"] + [x for x in args.ens_precisions.split(",") if x != "fp64"]
    res = {}
    for prec in precs:
        b = M.Batch([hosts[c["L"]] for c in ens], rhos, [c["T"] for c in ens], precision=code[prec])
        t0 = time.perf_counter()
        b.init()
        block, xcd = b.tune(args.tune_steps)
        t1 = time.perf_counter()
        nconv, steps, delta, state = b.run_to_steady(thr[prec], check_every=100, max_steps=args.max_steps)
        t2 = time.perf_counter()
        lap = [laplace(b.get(k, "rho"), b.get(k, "p"), c["L"]) if state[k] == 1 else None
               for k, c in enumerate(ens)]
        lu = float(sum(c["L"] ** 2 * int(steps[k]) for k, c in enumerate(ens)))
        res[prec] = dict(steady=int(nconv), diverged=int((state == 2).sum()), threshold=thr[prec],
                         wall_s=t2 - t0, run_s=t2 - t1, tune_s=t1 - t0, block=block, xcd_order=xcd,
                         steps_median=float(np.median(steps)), steps_max=int(steps.max()),
                         mlups=lu / (t2 - t0) / 1e6, lap=lap, state=state)
        b.close()
        print("# %s: %d/%d steady, %d diverged, %.1f s (tune %.1f s), block %d xcd %d, steps median %d max %d" % (
            prec, nconv, len(ens), res[prec]["diverged"], t2 - t0, t1 - t0, block, xcd,
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
            "threshold": r["threshold"], "wall_s": round(r["wall_s"], 2), "tune_s": round(r["tune_s"], 2),
            "mlups": round(r["mlups"], 1), "steps_median": r["steps_median"], "steps_max": r["steps_max"],
            "launch": {"block": r["block"], "xcd_order": r["xcd_order"]},
            "laplace_err_vs_fp64": e, "laplace_fit": fits,
            "laplace_fit_max_rel_err_vs_fp64": max([f["rel_err_vs_fp64"] for f in fits], default=None),
        }
    main_prec = precs[-1]
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
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16s": "f16-storage/f32"}[main_prec],
        "data": "synthetic (seeded ensemble, Maxwell densities)",
        "config": {"workload": "%d D2Q9 liMRT droplet/bubble cases, L in {64,96,128,192,256}, R in [6, L/4], "
                               "T/Tc in [0.6, 0.95], VdW a=0.5 b=4, seed 7, one batched launch per pass, "
                               "graph-captured 100-step runs" % len(ens),
                   "nodes": nodes, "steady": "max|drho|/rho_l over 100 steps < threshold"},
        "precisions": out_prec,
    }
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
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "fp16s"])
    ap.add_argument("--thermal", action="store_true")
    ap.add_argument("--generic", action="store_true", help="per-routine kernels instead of the fused kernel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="bubble3d", choices=["bubble3d", "ensemble"])
    ap.add_argument("--cases", type=int, default=1024, help="ensemble: number of cases")
    ap.add_argument("--ens-precisions", default="fp32,fp16s", help="ensemble: precisions besides fp64")
    ap.add_argument("--steady-fp64", type=float, default=1e-9)
    ap.add_argument("--steady-fp32", type=float, default=1e-6)
    ap.add_argument("--steady-fp16s", type=float, default=1e-4)
    ap.add_argument("--max-steps", type=int, default=200000)
    ap.add_argument("--tune-steps", type=int, default=20)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM traffic counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print("bench.py runs on one GPU: the whole lattice lives on a single MI355X "
              "(multi-GPU sharding is a design note, not built); got --gpus %d" % args.gpus, file=sys.stderr)
        return 2
    if args.workload == "ensemble":
        return run_ensemble(args)

    r = run_gpu(args)
    N, steps = r["N"], args.steps
    sec = r["total_ms"] / 1e3
    mlups = N * steps / sec / 1e6
    blu = bytes_per_lu(args.precision, args.thermal)
    # Dominant kernel: the fused collide pass (or the collision kernel of the per-routine
    # path).  Its algorithmic bytes per node: populations pulled once and written once,
    # plus the macroscopic values it reads or writes once (psi; thermal: T, S, U).
    s_pdf = {"fp64": 8, "fp32": 4, "fp16s": 2}[args.precision]
    s_mac = 8 if args.precision == "fp64" else 4
    kname = "fused" if "fused" in r["kern"] else "collide"
    k_ms, k_launches = r["kern"].get(kname, (0.0, 0))
    if kname == "collide":
        per_unit = 2 * 15 * s_mac + (1 + 3 + 3 + 1) * s_mac      # f in/out; rho, U, Fi, psi in
    elif args.thermal:
        per_unit = 4 * 15 * s_pdf + (1 + 1 + 1 + 3) * s_mac     # f, g in/out; psi, T, S in; U out
    else:
        per_unit = 2 * 15 * s_pdf + s_mac                         # f in/out; psi in
    achieved = per_unit * N / (k_ms / 1e3) / 1e9 if k_ms > 0 else None
    step_gbs = mlups * 1e6 * blu / 1e9
    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3 counter
    # runs of this workload (profiles/fused_traffic.json), when one exists for it
    traffic = None
    key = "d3q15_%d_%s_%s" % (args.size, args.precision, "thermal" if args.thermal else "iso")
    tpath = os.path.join(ROOT, "profiles", "fused_traffic.json")
    if kname == "fused" and os.path.exists(tpath):
        rec = json.load(open(tpath)).get("kernels", {}).get(key)
        if rec:
            traffic = rec["hbm_bytes_per_launch"]
    out = {
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
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16s": "f16-storage/f32"}[args.precision],
        "data": "synthetic (seeded bubble, Maxwell densities %.5f/%.5f)" % r["dens"],
        "config": {"workload": "D3Q15 %s static bubble %d^3 periodic, VdW T=0.8Tc, R=%g" % (
            "liMRT+myMRT thermal" if args.thermal else "liMRT", args.size, 40.0 * args.size / 256.0),
            "precision": args.precision, "path": "fused" if r["fused"] else "per-routine",
            "bytes_per_lu": blu},
        "hbm_gbs_effective": round(step_gbs, 1),
        "step_roofline_frac": round(step_gbs / PEAK_HBM_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None,
                     "traffic": traffic, "kernel": kname, "kernel_ms": round(k_ms, 5),
                     "bytes_per_unit": per_unit, "units_per_launch": N},
        "kernels_ms": {k: round(v[0], 5) for k, v in r["kern"].items()},
        "finite": r["finite"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

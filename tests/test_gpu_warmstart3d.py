"""The learned warm start on D3Q15 cases, reduced: 40^3 periodic xuMRT droplets and
bubbles of radius 10-15 (smaller ones dissolve in this box; FP32 fused path), training
profiles from the solver's own steady states, the two networks (every steady state;
interface profiles alone) trained on the GPU by the library's kernels, then held-out
cases run from the cold start and from the network's
proposal to the same steady criterion (max |d rho| / rho_l over 100 steps < 5e-6).  The
solver judges both runs: the warm runs must reach the cold runs' final state and take at
least 1.2x fewer steps in total."""
import numpy as np
import pytest

import multiphaselb_amd as M
from multiphaselb_amd import warmstart as W
from multiphaselb_amd.casefiles import CaseSpec
from multiphaselb_amd.initfields import maxwell, noise, sphere, splitmix64, vdw_critical
from _cases import make_case_dir

pytestmark = pytest.mark.gpu

L = 40
THR_TRAIN, THR = 2e-5, 5e-6
LAM15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)


def cases(n, seed):
    u = splitmix64(seed, 3 * n).reshape(n, 3)
    tc = vdw_critical(0.5, 4.0)[1]
    out = []
    for k in range(n):
        tr = 0.65 + 0.25 * u[k, 1]
        rv, rl = maxwell(tr * tc, a=0.5, b=4.0)
        out.append(dict(radius=10.0 + 5.0 * u[k, 0], tr=tr, T=tr * tc, droplet=bool(u[k, 2] < 0.5), rho_v=rv, rho_l=rl,
                        seed=1000 * seed + k))
    return out


def run_to_steady(case, rho_l, thr, max_steps=40000):
    case.density_change()
    steps = 0
    while steps < max_steps:
        case.step(100)
        steps += 100
        d = case.density_change()
        if not np.isfinite(d):
            return steps, False
        if d / rho_l < thr:
            return steps, True
    return steps, False


def test_warm_start_3d_reaches_the_cold_state_in_fewer_steps():
    spec = CaseSpec(lbmodel=1, nx=L, ny=L, nz=L, btype=5, f_model=2, f_Lambda=LAM15, f_sigma=0.125, g_model=None,
                    end=10 ** 8, write_interval=10 ** 8)
    case = M.Case(make_case_dir(spec), precision=M.FP32, heat_transfer=False, fused=True)
    N = case.N
    cen = ((L - 1) / 2.0,) * 3
    r = W.node_radius((L, L, L), cen)
    sites = W.distinct_sites((L, L, L))
    inner, outer = r < 2.0, r > r.max() - 2.0

    def cold_field(c):
        inside, outside = (c["rho_l"], c["rho_v"]) if c["droplet"] else (c["rho_v"], c["rho_l"])
        return (sphere((L, L, L), cen, c["radius"], inside, outside, 5.0) + noise((L, L, L), 1e-4, c["seed"])).ravel()

    def start(c, rho=None):
        if rho is not None:
            case.set("rho", rho)
        case.set("U", np.zeros((N, 3)))
        case.set("T", np.full(N, c["T"]))
        case.init()

    X, Y = [], []
    for c in cases(40, 31):
        start(c, cold_field(c))
        st, ok = run_to_steady(case, c["rho_l"], THR_TRAIN)
        if not ok:
            continue
        try:
            y, kept = W.steady_profile(case.get("rho"), r, sites, c["rho_v"], c["rho_l"], c["droplet"])
        except ValueError:
            continue
        X.append(W.features(c["tr"], c["radius"], L, c["droplet"], 3))
        Y.append(np.concatenate([W.target_from_profile(y, c["droplet"]), np.zeros(W.NPROF)]))
    X, Y = np.array(X, np.float32), np.array(Y, np.float32)
    keep = ~W.flat_targets(X, Y)
    print("3D training profiles: %d (%d interfaces)" % (len(X), int(keep.sum())), flush=True)
    assert len(X) >= 30 and keep.sum() >= 8
    net = M.Mlp(W.WIDTHS, 64)
    net.set_params(W.init_params(W.WIDTHS, 0))
    W.train(net, X, Y, iters=2500, seed=1)
    net_k = M.Mlp(W.WIDTHS, 64)
    net_k.set_params(W.init_params(W.WIDTHS, 1))
    hist = W.train(net_k, X[keep], Y[keep], iters=4000, seed=2)
    print("interface network loss %.2e -> %.2e" % (hist[:10].mean(), hist[-20:].mean()), flush=True)

    cold_total = warm_total = 0
    for k, c in enumerate(cases(10, 32)):
        rho0 = cold_field(c)
        mass = float(rho0[sites].sum())
        start(c, rho0)
        st_c, ok_c = run_to_steady(case, c["rho_l"], THR)
        rc = case.get("rho")
        prof, flat = W.propose(net, net_k, W.features(c["tr"], c["radius"], L, c["droplet"], 3), c["droplet"],
                               c["rho_v"], c["rho_l"])
        assert W.proposal_ok(prof, c["rho_v"], c["rho_l"])
        W.raster_case(case, prof, cen, mass, c["radius"], int(sites.sum()), uniform=flat)
        start(c)
        st_w, ok_w = run_to_steady(case, c["rho_l"], THR)
        rw = case.get("rho")
        kept = abs(rc[inner].mean() - rc[outer].mean()) >= 0.25 * (c["rho_l"] - c["rho_v"])
        dev = max(abs(rw[m].mean() - rc[m].mean()) / rc[m].mean() for m in (inner, outer))
        print("  case %d %s R %.1f T/Tc %.3f: cold %d (%s) warm %d (%s) %s proposal, %s, final-state dev %.1e" % (
            k, "droplet" if c["droplet"] else "bubble", c["radius"], c["tr"], st_c, ok_c, st_w, ok_w,
            "flat" if flat else "interface", "kept" if kept else "dissolved", dev), flush=True)
        assert ok_c and ok_w
        # the same steady state: coexistence densities (kept interfaces) or the uniform
        # density (dissolved) to 1e-4 of the cold run
        assert dev < 1e-4
        cold_total += st_c
        warm_total += st_w
    print("3D warm start: %d cold steps, %d warm steps, gain %.2f" % (cold_total, warm_total, cold_total / warm_total),
          flush=True)
    assert warm_total * 1.2 < cold_total

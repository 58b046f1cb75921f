"""The learned warm start end to end on small D2Q9 cases: steady solver fields become
training profiles, the network is trained on the GPU by the library's own kernels, and
held-out cases started from its proposal reach the same steady state as their cold
starts in fewer steps.  The solver (FP64 batch, steady criterion 1e-9) judges both runs."""
import numpy as np
import pytest

import multiphaselb_amd as M
from multiphaselb_amd import warmstart as W
from multiphaselb_amd.casefiles import CaseSpec
from multiphaselb_amd.initfields import ensemble, ensemble_density
from _cases import make_case_dir

pytestmark = pytest.mark.gpu

L = 64
THR = 1e-9


def run_batch(cases, rhos, host):
    b = M.Batch([host] * len(cases), rhos, [c["T"] for c in cases], precision=M.FP64)
    b.init()
    nconv, steps, delta, state = b.run_to_steady(THR, check_every=100, max_steps=80000)
    fields = [b.get(k, "rho") for k in range(len(cases))]
    b.close()
    return steps, state, fields


def test_warm_start_gains_over_cold_start():
    host = M.Case(make_case_dir(CaseSpec(lbmodel=0, nx=L, ny=L, btype=3, f_model=0, a=0.5, b=4.0,
                                         end=10 ** 8, write_interval=10 ** 8)), heat_transfer=False, device=-1)
    r = W.node_radius((L, L), ((L - 1) / 2.0, (L - 1) / 2.0))
    sites = W.distinct_sites((L, L))
    train = ensemble(96, seed=21, sizes=(L,))
    steps, state, fields = run_batch(train, [ensemble_density(c) for c in train], host)
    X, Y = [], []
    for k, c in enumerate(train):
        if state[k] != 1:
            continue
        try:
            y, _ = W.extract_profile(fields[k], r, c["rho_v"], c["rho_l"], c["droplet"])
        except ValueError:          # the droplet or bubble dissolved: no profile to learn
            continue
        X.append(W.features(c["tr"], c["radius"], L, c["droplet"], 2))
        Y.append(np.concatenate([W.target_from_profile(y, c["droplet"]), np.zeros(W.NPROF)]))
    print("training profiles: %d of %d cases (others dissolved or still moving)" % (len(X), len(train)), flush=True)
    assert len(X) >= 40
    X, Y = np.array(X, np.float32), np.array(Y, np.float32)
    net = M.Mlp(W.WIDTHS, 64)
    net.set_params(W.init_params(W.WIDTHS, 0))
    hist = W.train(net, X, Y, iters=3000, seed=1)
    print("training loss %.3e -> %.3e" % (hist[:10].mean(), hist[-20:].mean()), flush=True)
    assert hist[-20:].mean() < 0.05 * hist[:10].mean()

    held = ensemble(16, seed=22, sizes=(L,))
    cold0 = [ensemble_density(c) for c in held]
    s_cold, st_cold, f_cold = run_batch(held, cold0, host)
    warm0 = []
    for c, rho0 in zip(held, cold0):
        o = net.forward(W.features(c["tr"], c["radius"], L, c["droplet"], 2)[None, :])[0, :W.NPROF]
        prof = W.profile_density(W.profile_from_output(o, c["droplet"]), c["rho_v"], c["rho_l"])
        rho_w, _ = W.raster_host(prof, r, float(rho0[sites].sum()), c["radius"], sites)
        warm0.append(rho_w)
    s_warm, st_warm, f_warm = run_batch(held, warm0, host)
    # a cold run still moving at the step limit counts with the limit (a lower bound)
    both = (st_cold == 1) & (st_warm == 1)
    assert (st_warm == 1).sum() >= (st_cold == 1).sum() and (st_warm == 1).sum() >= 14
    # gain over the cases that keep an interface (a dissolving one ends uniform either way)
    inner, outer = r < 2.0, r > r.max() - 2.0
    keep = np.array([abs(f_cold[k][inner].mean() - f_cold[k][outer].mean()) >= 0.25 * (held[k]["rho_l"] - held[k]["rho_v"])
                     for k in range(len(held))])
    sel = (st_warm == 1) & keep
    gain = np.array(s_cold, float)[sel] / np.array(s_warm, float)[sel]
    dev = []
    for k in range(len(held)):
        c = held[k]
        dk = [abs(f_warm[k][sel].mean() - f_cold[k][sel].mean()) / f_cold[k][sel].mean() for sel in (inner, outer)]
        dissolved = abs(f_cold[k][inner].mean() - f_cold[k][outer].mean()) < 0.25 * (c["rho_l"] - c["rho_v"])
        print("  case %2d %s R %.1f T/Tc %.3f: cold %6d (%d) warm %6d (%d) dev %.1e %.1e%s" % (
            k, "drop" if c["droplet"] else "bubl", c["radius"], c["tr"], s_cold[k], st_cold[k], s_warm[k], st_warm[k],
            dk[0], dk[1], " dissolved" if dissolved else ""), flush=True)
        if both[k]:
            dev += dk
    print("cold steady %d/16 steps median %d; warm steady %d/16 steps median %d; gain median %.2f min %.2f; "
          "coexistence dev max %.1e over %d cases" % (
              (st_cold == 1).sum(), np.median(s_cold), (st_warm == 1).sum(), np.median(s_warm), np.median(gain),
              gain.min(), max(dev) if dev else 0.0, both.sum()), flush=True)
    assert max(dev) < 1e-4
    assert np.median(gain) > 1.2

"""Batched ensembles (mplb_batch_*) through the C ABI: every case of a batch must follow
the reference's own time loop for that case alone, so each is compared with a CPU oracle
run of the same case (FP64 mode: bit-identical; FP32 and FP16-deviation storage: within
the stated tolerances), and the graph-captured, compacting, autotuned and persistent
paths must give exactly the plain stepping's result."""
import numpy as np
import pytest

import multiphaselb_amd as M
import _oracle as O
from _cases import make_case_dir, oracle_params
from multiphaselb_amd.casefiles import CaseSpec
from multiphaselb_amd.initfields import maxwell, noise, sphere, vdw_critical

pytestmark = pytest.mark.gpu

A, B = 0.5, 4.0
TC = vdw_critical(A, B)[1]

# (nx, ny, T/Tc, radius, droplet)
SMALL = [(16, 16, 0.7, 4.0, True), (24, 18, 0.85, 5.0, False), (12, 30, 0.6, 3.5, True), (20, 20, 0.9, 5.5, False)]


def spec_of(nx, ny):
    return CaseSpec(lbmodel=0, nx=nx, ny=ny, btype=3, f_model=0, a=A, b=B, end=10 ** 8, write_interval=10 ** 8)


def density(nx, ny, tr, radius, droplet, seed):
    rv, rl = maxwell(tr * TC, a=A, b=B)
    inside, outside = (rl, rv) if droplet else (rv, rl)
    rho = sphere((1, ny, nx), ((nx - 1) / 2.0, (ny - 1) / 2.0, 0.0), radius, inside, outside, 5.0)
    return (rho + noise((1, ny, nx), 1e-4, seed, (True, True, False))).ravel()


def host_case(nx, ny):
    return M.Case(make_case_dir(spec_of(nx, ny)), precision=M.FP64, heat_transfer=False, device=-1)


def make_batch(cases, precision, extra=()):
    hosts = {}
    handles, rhos, temps = [], [], []
    for k, (nx, ny, tr, radius, droplet) in enumerate(list(cases) + list(extra)):
        if (nx, ny) not in hosts:
            hosts[(nx, ny)] = host_case(nx, ny)
        handles.append(hosts[(nx, ny)])
        rhos.append(density(nx, ny, tr, radius, droplet, 100 + k))
        temps.append(tr * TC)
    b = M.Batch(handles, rhos, temps, precision=precision)
    b._hosts = hosts
    return b, rhos, temps


def oracle_run(nx, ny, rho, T, steps):
    o = O.Oracle(oracle_params(spec_of(nx, ny), ht=False))
    N = nx * ny
    o.set("rho", rho)
    o.set("U", np.zeros((N, 3)))
    o.set("T", np.full(N, T))
    o.init()
    o.run(steps)
    o.call("macro_pressure")
    return o.get("rho"), o.get("p")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_batch_fp64_matches_oracle_bit_for_bit():
    steps = 160
    b, rhos, temps = make_batch(SMALL, M.FP64)
    b.init()
    b.step(steps - 1)            # init runs the first step's collision
    for k, (nx, ny, *_rest) in enumerate(SMALL):
        print("case", k, nx, ny, flush=True)
        assert b.case_steps(k) == steps
        r_ref, p_ref = oracle_run(nx, ny, rhos[k], temps[k], steps)
        np.testing.assert_array_equal(b.get(k, "rho"), r_ref)
        np.testing.assert_array_equal(b.get(k, "p"), p_ref)


# FP32 arithmetic: psi = sqrt(2 (p - rho cs2) / G) loses digits to cancellation where p
# is close to rho cs2 (the vapour side), so after 300 steps the density differs from the
# FP64 oracle by a few 1e-5; FP16 storage keeps about 3 decimal digits of f - w rho_ref
# (FP16S, one reference density per case) or of each population's relative deviation from
# w R, R its node's own density (FP16N: both phases keep those digits; the deepest
# quench, T = 0.6 Tc, is the worst case).  Pressure adds the psi Laplacian, one more
# difference.
@pytest.mark.parametrize("precision,tol", [(M.FP32, 1e-4), (M.FP16S, 5e-3), (M.FP16N, 1e-3)])
def test_batch_reduced_precision_tracks_oracle(precision, tol):
    steps = 300
    b, rhos, temps = make_batch(SMALL, precision)
    b.init()
    b.step(steps - 1)
    for k, (nx, ny, *_rest) in enumerate(SMALL):
        r_ref, p_ref = oracle_run(nx, ny, rhos[k], temps[k], steps)
        e = rel(b.get(k, "rho"), r_ref)
        ep = rel(b.get(k, "p"), p_ref)
        print("case %d precision %d rho err %.2e p err %.2e" % (k, precision, e, ep), flush=True)
        assert e < tol and ep < 10 * tol


def test_graph_replay_equals_plain_steps():
    """run_to_steady (graph-captured runs, threshold 0 so no case stops) gives exactly the
    plain launches' state, for each launch shape the tuner can pick."""
    plain, _, _ = make_batch(SMALL, M.FP32)
    plain.set_launch(256, 1)
    plain.init()
    plain.step(120)
    for block, xcd, rows in ((64, 0, 0), (256, 1, 0), (1024, 1, 0), (256, 0, 7), (256, 1, 32)):
        g, _, _ = make_batch(SMALL, M.FP32)
        g.set_launch(block, xcd)
        if rows:
            g.set_march(rows)
        g.init()
        assert g.launch_info()["march_rows"] == rows
        nconv, steps, delta, state = g.run_to_steady(0.0, check_every=40, max_steps=120)
        assert nconv == 0 and (state == 0).all() and (steps == 121).all()
        for k in range(len(SMALL)):
            np.testing.assert_array_equal(g.get(k, "rho"), plain.get(k, "rho"))


def test_steady_cases_leave_the_table_and_keep_their_state():
    """A uniform case is steady at the first check and stops there; the others keep going
    and still match their plain runs; a case whose potential is undefined (T far above
    the range where p < rho cs2, so psi = sqrt(negative)) turns NaN and is reported as
    diverged."""
    b, rhos, temps = make_batch(SMALL, M.FP64, extra=[(16, 16, 0.7, 4.0, True), (16, 16, 0.7, 4.0, True)])
    nu, nn = len(SMALL), len(SMALL) + 1
    rhos[nu] = np.full(16 * 16, maxwell(0.7 * TC, a=A, b=B)[1])
    b.set_case(nu, rhos[nu], temps[nu])
    b.set_case(nn, rhos[nn], 30.0 * TC)
    b.init()
    nconv, steps, delta, state = b.run_to_steady(1e-12, check_every=20, max_steps=100)
    print("states", state, "steps", steps, "delta", delta, flush=True)
    assert nconv == 1
    assert state[nu] == 1 and steps[nu] == 21
    assert state[nn] == 2 and steps[nn] == 21
    assert (state[:nu] == 0).all() and (steps[:nu] == 101).all()
    for k, (nx, ny, *_rest) in enumerate(SMALL):
        r_ref, _ = oracle_run(nx, ny, rhos[k], temps[k], 101)
        np.testing.assert_array_equal(b.get(k, "rho"), r_ref)
    r_ref, _ = oracle_run(16, 16, rhos[nu], temps[nu], 21)
    np.testing.assert_array_equal(b.get(nu, "rho"), r_ref)


def test_autotuner_keeps_results_exact():
    """The tuner's trial steps advance every case; afterwards the state equals the oracle
    at the same time level."""
    b, rhos, temps = make_batch(SMALL[:2], M.FP64)
    b.init()
    block, xcd = b.tune(steps_per_trial=4)
    assert block in (64, 128, 256, 512, 1024) and xcd in (0, 1)
    info = b.launch_info()
    assert info["block"] == block and info["nodes"] == sum(nx * ny for nx, ny, *_ in SMALL[:2])
    assert info["march_rows"] in (0, 16, 32, 64)
    b.step(3)
    for k, (nx, ny, *_rest) in enumerate(SMALL[:2]):
        t = b.case_steps(k)
        assert t == 1 + 16 * (2 + 4) + 3       # 2 x (3 single-pass + 5 two-pass) trials
        r_ref, _ = oracle_run(nx, ny, rhos[k], temps[k], t)
        np.testing.assert_array_equal(b.get(k, "rho"), r_ref)


def test_fp32_batch_mass_does_not_drift():
    """FP32 batch collisions add the moment change to the populations (no per-step
    rebuild), so each closed periodic case keeps its mass over thousands of steps."""
    b, rhos, temps = make_batch(SMALL, M.FP32)
    b.init()
    masses = []
    for k, (nx, ny, *_) in enumerate(SMALL):
        j, i = np.divmod(np.arange(nx * ny), nx)
        sites = (i < nx - 1) & (j < ny - 1)
        masses.append((sites, float(np.asarray(b.get(k, "rho"), np.float64)[sites].sum())))
    b.step(4000)
    for k, (sites, m0) in enumerate(masses):
        m = float(np.asarray(b.get(k, "rho"), np.float64)[sites].sum())
        print("  case %d mass rel change %.2e" % (k, m / m0 - 1.0), flush=True)
        assert abs(m / m0 - 1.0) < 1e-5


@pytest.mark.parametrize("precision", [M.FP64, M.FP32, M.FP16S, M.FP16N])
def test_single_pass_equals_two_pass(precision):
    """The single-pass kernel (rows of up to four equal-size cases per workgroup, psi in an
    LDS ring) does each node's arithmetic exactly as the two-pass kernels: identical
    states, for uneven cases, several cases per item and row counts that do not divide
    the case height."""
    cases = SMALL + [(16, 16, 0.8, 4.5, False), (16, 16, 0.75, 3.0, True), (16, 16, 0.9, 5.0, True),
                     (16, 16, 0.65, 4.0, False), (16, 16, 0.7, 3.5, True)]
    two, _, _ = make_batch(cases, precision)
    two.set_launch(128, 1)
    two.init()
    two.step(60)
    for rows in (5, 64):
        one, _, _ = make_batch(cases, precision)
        one.set_march(rows)
        one.init()
        one.step(60)
        for k in range(len(cases)):
            np.testing.assert_array_equal(one.get(k, "rho"), two.get(k, "rho"))
            np.testing.assert_array_equal(one.get(k, "p"), two.get(k, "p"))


@pytest.mark.parametrize("precision", [M.FP64, M.FP32, M.FP16N])
def test_persistent_launch_equals_per_step_launches(precision):
    """One persistent cooperative launch per run of steps (the grid meets at a barrier
    between steps) against one launch per step: identical states after an odd number of
    plain steps (the level buffers end swapped), after further graph-length runs, and
    through run_to_steady; four cases per work item, and (one row per item) about 1360
    work items, more than the resident workgroups, so each workgroup takes several."""
    cases = SMALL + [(32, 32, 0.65 + 0.25 * (k % 11) / 10, 5.0 + (k % 4), k % 2 == 0) for k in range(160)]
    runs = []
    for persistent in (False, True):
        b, _, _ = make_batch(cases, precision)
        b.set_march(1)
        if persistent:
            b.set_persistent(True)
        b.init()
        b.step(37)
        after_plain = [b.get(k, "rho") for k in range(len(cases))]
        nconv, steps, delta, state = b.run_to_steady(1e-12, check_every=40, max_steps=160)
        runs.append((after_plain, [b.get(k, "rho") for k in range(len(cases))],
                     [b.get(k, "p") for k in range(len(cases))], steps, state))
    (p0, r0, q0, s0, st0), (p1, r1, q1, s1, st1) = runs
    np.testing.assert_array_equal(s0, s1)
    np.testing.assert_array_equal(st0, st1)
    for k in range(len(cases)):
        np.testing.assert_array_equal(p1[k], p0[k])
        np.testing.assert_array_equal(r1[k], r0[k])
        np.testing.assert_array_equal(q1[k], q0[k])

"""The hot kernels across their tile seams, through the C ABI.

* k_fused_march (the isothermal D3Q15 single pass) works on 64 x 4 columns of sites and
  z chunks: a 130 x 12 x 40 periodic bubble has three x tiles (the last one 2 sites wide),
  three column rows in y and five 8-plane z chunks, so ring sites cross tile and chunk
  boundaries everywhere.  It must follow the two-pass fused kernels (same per-node
  arithmetic; the compiler may fuse a multiply-add differently in the two kernels, so the
  bar is FP32 rounding) and track the FP64 oracle; FP16 storage likewise against its two
  passes, and the variants that stage each plane's pulls in LDS (FP32) or issue them a
  plane early (FP32, FP16S) likewise.
* k_fused_tmarch (the thermal D3Q15 single pass: xuMRT + myMRT with heat transfer)
  keeps rho, T and U on a two-site ring and psi on a one-site ring in LDS; on a
  130 x 10 x 24 periodic thermal bubble (three x tiles, a partial y tile, z chunks) it
  must follow the three fused passes to FP32 rounding, with FP32 and FP16 storage,
  be unchanged by a mid-run read-back, and track the FP64 oracle.
* k_batch_march (the D2Q9 ensemble single pass) gives each 256-lane workgroup whole rows
  of up to four equal-size cases: widths 64 (four cases per workgroup), 96 and 128 (two),
  192 and 256 (one) are run together, FP64, bit for bit against the oracle; cases wider
  than 256 nodes take the two-pass kernels, also bit for bit.
"""
import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import pair, sphere_fields
from test_gpu_batch import make_batch, oracle_run
from test_gpu_parity import TC, d3

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


SEAM = d3(nx=130, ny=12, nz=40, btype=5, g_model=None, write_interval=10 ** 6)
SEAM_FIELDS = dict(T=0.8 * TC, r_in=0.03, r_out=0.16, radius=5.0, centre=(63.0, 5.5, 19.5))


VARIANT_ENV = {"staged": "MPLB_MARCH_LDS", "ahead": "MPLB_MARCH_PF"}


@pytest.mark.parametrize("precision,variant", [(M.FP32, ""), (M.FP32, "staged"), (M.FP32, "ahead"), (M.FP16S, ""),
                                               (M.FP16S, "ahead"), (M.FP16N, "")],
                         ids=["fp32", "fp32-staged", "fp32-ahead", "fp16s", "fp16s-ahead", "fp16n"])
def test_march_across_tiles_and_chunks_equals_two_pass(precision, variant, monkeypatch):
    """staged: the variant that loads each plane's pulls straight into LDS
    (MPLB_MARCH_LDS=1); ahead: the variant that issues each plane's pulls one plane
    early (MPLB_MARCH_PF=1).  For both, a read-back half way must leave the trajectory
    unchanged."""
    steps = 120
    print("seam case 130x12x40 precision", precision, variant, flush=True)
    fields = sphere_fields(SEAM, seed=21, **SEAM_FIELDS)
    monkeypatch.setenv("MPLB_FUSED_KERNEL", "twopass")
    two, _ = pair(SEAM, fields, precision=precision, ht=False, fused=True)
    monkeypatch.delenv("MPLB_FUSED_KERNEL")
    if variant:
        monkeypatch.setenv(VARIANT_ENV[variant], "1")
    one, o = pair(SEAM, fields, precision=precision, ht=False, fused=True)
    assert one.get_info().fused_active == 1
    assert two.step(steps) == steps
    if variant:
        assert one.step(steps // 2) == steps // 2
        one.get("rho")
        assert one.step(steps - steps // 2) == steps - steps // 2
    else:
        assert one.step(steps) == steps
    # rho and f relative to their largest value, U absolute (lattice units: a wrong
    # population moves U by ~1e-3, a rounding difference by ~1e-7)
    tol = {"rho": 1e-5, "f": 1e-5, "U": 1e-6} if precision == M.FP32 else {"rho": 1e-4, "f": 1e-4, "U": 1e-5}
    for f in ("rho", "U", "f"):
        a, b = one.get(f), two.get(f)
        e = rel(a, b) if f != "U" else float(np.max(np.abs(a - b)))
        print("  %s single vs two pass: exact %s, diff %.2e" % (f, np.array_equal(a, b), e), flush=True)
        assert e < tol[f], f
    if precision == M.FP32:
        o.run(steps)
        errs = {f: rel(one.get(f), o.get(f)) for f in ("rho", "f")}
        print("  vs FP64 oracle:", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
        assert errs["rho"] < 2e-5 and errs["f"] < 2e-5


TSEAM = d3(nx=130, ny=10, nz=24, btype=5, write_interval=10 ** 6)
TSEAM_FIELDS = dict(T=0.85 * TC, r_in=0.03, r_out=0.16, radius=4.5, centre=(64.0, 4.5, 11.5))


@pytest.mark.parametrize("precision", [M.FP32, M.FP16S, M.FP16N])
def test_thermal_march_tracks_three_passes(precision, monkeypatch):
    """Same per-node arithmetic as the three passes; the compiler may still fuse a
    multiply-add differently in the two kernels, so the comparison allows FP32 rounding
    (1e-5 relative; FP16 storage turns a one-ulp FP32 difference into an FP16 rounding
    step, 1e-3).  A read-back half way must leave the single pass's trajectory exactly
    as without it."""
    steps = 80
    print("thermal seam case 130x10x24 precision", precision, flush=True)
    fields = sphere_fields(TSEAM, seed=23, **TSEAM_FIELDS)
    monkeypatch.setenv("MPLB_THERMAL_MARCH", "0")
    three, _ = pair(TSEAM, fields, precision=precision, ht=True, fused=True)
    monkeypatch.setenv("MPLB_THERMAL_MARCH", "1")
    one, o = pair(TSEAM, fields, precision=precision, ht=True, fused=True)
    two, _ = pair(TSEAM, fields, precision=precision, ht=True, fused=True)
    assert one.get_info().fused_active == 1
    assert three.step(steps) == steps and two.step(steps) == steps
    assert one.step(steps // 2) == steps // 2
    one.get("rho")
    assert one.step(steps - steps // 2) == steps - steps // 2
    tol = 1e-5 if precision == M.FP32 else 1e-3
    for f in ("rho", "U", "T", "f", "g"):
        a, b, c = one.get(f), three.get(f), two.get(f)
        assert np.array_equal(a, c), f
        e = rel(a, b) if f != "U" else float(np.max(np.abs(a - b)))
        print("  %s single vs three passes: exact %s, diff %.2e" % (f, np.array_equal(a, b), e), flush=True)
        assert e < (tol if f != "U" else tol / 10), f
    if precision == M.FP32:
        o.run(steps)
        errs = {f: rel(one.get(f), o.get(f)) for f in ("rho", "T", "f", "g")}
        print("  vs FP64 oracle:", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
        assert errs["rho"] < 2e-5 and errs["T"] < 2e-5 and errs["f"] < 2e-5 and errs["g"] < 2e-5


def test_batch_march_at_every_ensemble_width_fp64():
    steps = 200
    widths = [(64, 4), (96, 2), (128, 2), (192, 1), (256, 1)]
    cases = []
    for L, n in widths:
        for k in range(n):
            tr = 0.6 + 0.08 * ((k + len(cases)) % 4)
            cases.append((L, L, tr, L / 5.0 + 1.5 * k, (k + L) % 2 == 0))
    b, rhos, temps = make_batch(cases, M.FP64)
    info = b.launch_info()
    print("batch widths 64..256:", len(cases), "cases, march rows", info["march_rows"], flush=True)
    assert info["march_rows"] > 0          # the single-pass kernel serves all of them
    b.init()
    b.step(steps - 1)
    for k, (nx, ny, *_r) in enumerate(cases):
        print("  case %d (%d x %d)" % (k, nx, ny), flush=True)
        r_ref, p_ref = oracle_run(nx, ny, rhos[k], temps[k], steps)
        np.testing.assert_array_equal(b.get(k, "rho"), r_ref)
        np.testing.assert_array_equal(b.get(k, "p"), p_ref)


def test_batch_wider_than_a_workgroup_takes_two_passes_fp64():
    steps = 200
    cases = [(300, 40, 0.7, 12.0, True), (260, 24, 0.85, 8.0, False)]
    b, rhos, temps = make_batch(cases, M.FP64)
    assert b.launch_info()["march_rows"] == 0
    b.init()
    b.step(steps - 1)
    for k, (nx, ny, *_r) in enumerate(cases):
        print("  wide case %d (%d x %d)" % (k, nx, ny), flush=True)
        r_ref, p_ref = oracle_run(nx, ny, rhos[k], temps[k], steps)
        np.testing.assert_array_equal(b.get(k, "rho"), r_ref)
        np.testing.assert_array_equal(b.get(k, "p"), p_ref)

"""Fused structured-box kernels (FP32 and FP16-deviation storage) against the FP64
oracle and against the per-routine FP32 kernels, through the C ABI."""
import dataclasses

import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import make_case_dir, pair, sphere_fields
from multiphaselb_amd.initfields import maxwell, noise, sphere
from test_gpu_parity import CASES, TC, d3

pytestmark = pytest.mark.gpu

FUSED_CASES = ["d2q9_droplet_nht", "d2q9_bubble_thermal", "d2q9_column_fixedT", "d2q9_column_liTempMRT",
               "d2q9_surface_tension", "d3q15_bubble_nht", "d3q15_thermal_column", "d3q15_surface_tension",
               "d2q9_column_two_frdt", "d3q15_column_frdt_nopert", "d3q15_bubble_thermal",
               "d2q9_carnahan_starling_keys_first", "d3q15_peng_robinson_keys_first"]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("name", FUSED_CASES)
def test_fused_fp32_tracks_oracle(name):
    spec, fk, ht, frozen, steps = CASES[name]
    print("case", name, flush=True)
    fields = sphere_fields(spec, seed=7, **fk)
    c, o = pair(spec, fields, precision=M.FP32, ht=ht, frozen=frozen, fused=True)
    assert c.get_info().fused_active == 1
    assert c.step(steps) == steps
    o.run(steps)
    errs = {f: rel(c.get(f), o.get(f)) for f in ("rho", "T", "f")}
    du = np.max(np.abs(c.get("U") - o.get("U"))) / max(np.max(np.abs(o.get("U"))), 1e-12)
    print("  fp32 fused vs oracle:", {k: "%.2e" % v for k, v in errs.items()}, "U %.2e" % du, flush=True)
    assert errs["rho"] < 2e-5 and errs["T"] < 2e-5 and errs["f"] < 2e-5
    assert du < 2e-2


@pytest.mark.parametrize("precision,tol", [(M.FP16S, 2e-2), (M.FP16N, 2e-4)], ids=["fp16s", "fp16n"])
@pytest.mark.parametrize("name", ["d3q15_bubble_nht", "d3q15_thermal_column", "d2q9_column_fixedT"])
def test_fused_fp16_storage_error(name, precision, tol):
    """FP16S keeps f - w rho_ref with one reference density: the light phase's deviations
    are of the order of the density jump, ~1e-2 of rho after the run.  FP16N keeps each
    population relative to its own node's density, so both phases keep the FP16
    mantissa: ~1e-5."""
    spec, fk, ht, frozen, steps = CASES[name]
    fields = sphere_fields(spec, seed=7, **fk)
    c, o = pair(spec, fields, precision=precision, ht=ht, frozen=frozen, fused=True)
    c.step(steps)
    o.run(steps)
    e = rel(c.get("rho"), o.get("rho"))
    print("  fp16 storage (precision %d) rho err vs FP64 oracle %.2e" % (precision, e), flush=True)
    assert np.isfinite(c.get("f")).all()
    assert e < tol


def test_fused_readback_does_not_perturb_the_run():
    """Reading fields mid-run (materialisation) leaves the trajectory bit-identical to an
    uninterrupted run (both runs take the same write steps)."""
    spec, fk, ht, frozen, steps = CASES["d3q15_thermal_column"]
    fields = sphere_fields(spec, seed=5, **fk)
    a, _ = pair(spec, fields, precision=M.FP32, ht=ht, fused=True)
    b, _ = pair(spec, fields, precision=M.FP32, ht=ht, fused=True)
    a.step(steps)
    for _ in range(4):
        b.step(steps // 4)
        b.get("rho")
        b.get("f")
    for f in ("rho", "U", "T", "f", "g"):
        assert np.array_equal(a.get(f), b.get(f)), f


@pytest.mark.parametrize("precision", [M.FP32, M.FP16S, M.FP16N])
def test_march_matches_two_pass(precision, monkeypatch):
    """The single-pass kernel follows the two-pass kernels on a periodic D3Q15 bubble, with
    FP32 and FP16 storage; psi is computed with the same arithmetic in both, so the
    trajectories agree to FP32
    contraction differences: 1e-5 relative on rho and f, and 1e-6 absolute on U (lattice
    units; a rounding-level difference moves U by ~1e-7, a wrong population by ~1e-3,
    while max |U| is itself small, so a bound relative to it would be fragile).  With FP16
    storage a one-ulp FP32 difference can change a stored FP16 value, so the bars are ten
    times wider."""
    spec, fk, ht, frozen, steps = CASES["d3q15_bubble_nht"]
    fields = sphere_fields(spec, seed=11, **fk)
    monkeypatch.setenv("MPLB_FUSED_KERNEL", "twopass")
    two, _ = pair(spec, fields, precision=precision, ht=False, fused=True)
    monkeypatch.delenv("MPLB_FUSED_KERNEL")
    one, _ = pair(spec, fields, precision=precision, ht=False, fused=True)
    two.step(steps)
    one.step(steps)
    widen = 1 if precision == M.FP32 else 10
    for f in ("rho", "f"):
        e = rel(one.get(f), two.get(f))
        print("  %s rel diff %.2e" % (f, e), flush=True)
        assert e < 1e-5 * widen
    du = float(np.max(np.abs(one.get("U") - two.get("U"))))
    print("  U abs diff %.2e (max |U| %.2e)" % (du, float(np.max(np.abs(two.get("U"))))), flush=True)
    assert du < 1e-6 * widen


@pytest.mark.parametrize("kernel", ["single-pass", "twopass"])
def test_fp32_mass_does_not_drift(kernel, monkeypatch):
    """The FP32 collision adds its moment change to the pulled populations instead of
    rebuilding them from the post-collision moments, so the mass of a closed periodic box
    stays put to ~1e-6: a rebuild rounds every population afresh and drifts by about 5e-8 of the
    mass per step (2e-4 over these 4000 steps)."""
    if kernel == "twopass":
        monkeypatch.setenv("MPLB_FUSED_KERNEL", "twopass")
    n = 24
    spec = d3(nx=n, ny=n, nz=n, btype=5, end=10 ** 8, write_interval=10 ** 8)
    c = M.Case(make_case_dir(spec), precision=M.FP32, heat_transfer=False, fused=True)
    rv, rl = maxwell(0.8 * TC, a=0.5, b=4.0)
    cen = ((n - 1) / 2.0,) * 3
    rho = (sphere((n, n, n), cen, 6.0, rv, rl, 5.0) + noise((n, n, n), 1e-4, 3)).ravel()
    c.set("rho", rho)
    c.set("U", np.zeros((c.N, 3)))
    c.set("T", np.full(c.N, 0.8 * TC))
    c.init()
    idx = np.indices((n, n, n)).reshape(3, -1)
    sites = (idx < n - 1).all(axis=0)           # the duplicate end nodes repeat the first
    m0 = float(np.asarray(c.get("rho"), np.float64)[sites].sum())
    for chunk in range(4):
        c.step(1000)
        m = float(np.asarray(c.get("rho"), np.float64)[sites].sum())
        print("  %s: after %d steps mass rel change %.2e" % (kernel, 1000 * (chunk + 1), m / m0 - 1.0), flush=True)
        assert abs(m / m0 - 1.0) < 1e-5


@pytest.mark.parametrize("name", ["d3q15_bubble_nht", "d2q9_bubble_thermal"])
def test_fused_write_interval_leaves_the_trajectory_unchanged(name):
    """Without gravity rho_0 enters no collision, so write steps (pressure, rho_0
    average, NaN check) change nothing later steps read: write intervals of 7 and of the
    case's own give bit-identical FP32 trajectories."""
    spec, fk, ht, frozen, steps = CASES[name]
    fields = sphere_fields(spec, seed=5, **fk)
    a, _ = pair(spec, fields, precision=M.FP32, ht=ht, fused=True)
    b, _ = pair(dataclasses.replace(spec, write_interval=7), fields, precision=M.FP32, ht=ht, fused=True)
    a.step(steps)
    b.step(steps)
    for f in ("rho", "U", "T", "f"):
        assert np.array_equal(a.get(f), b.get(f)), f

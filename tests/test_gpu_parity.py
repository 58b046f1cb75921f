"""GPU parity: the HIP library (FP64 mode) against the CPU oracle on identical seeded
inputs, through the C ABI.  Each case runs the reference's full time loop (collision,
streaming, boundaries, macroscopic fields, force, heat source, write-step pressure and
rho_0 averaging) and compares every field.  Bar: 1e-12 relative (the FP64 kernels keep
the reference's operation order, so most cases are bit-identical)."""
import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import max_rel, pair, sphere_fields
from multiphaselb_amd.casefiles import BC, CaseSpec
from multiphaselb_amd.initfields import critical_temperature, maxwell

pytestmark = pytest.mark.gpu

TC = 8 * 0.5 / (27 * 4.0)            # van der Waals critical temperature for a=0.5, b=4, R=1
LAM9 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0)
LAM15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)
GLAM9 = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0)
GLAM15 = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
RL, RV = 0.16, 0.03
TOL = 1e-12
TC_CS = critical_temperature("cs", a=1.0, b=4.0)
TC_PR = critical_temperature("pr", a=2.0 / 49.0, b=2.0 / 21.0)
CS_85 = maxwell(0.85 * TC_CS, "cs", a=1.0, b=4.0)
PR_85 = maxwell(0.85 * TC_PR, "pr", a=2.0 / 49.0, b=2.0 / 21.0, w=0.344)


def d2(**kw):
    base = dict(lbmodel=0, f_model=0, f_Lambda=LAM9, g_model=1, g_Lambda=GLAM9)
    base.update(kw)
    return CaseSpec(**base)


def d3(**kw):
    base = dict(lbmodel=1, f_model=2, f_Lambda=LAM15, g_model=1, g_Lambda=GLAM15)
    base.update(kw)
    return CaseSpec(**base)


# name -> (spec, field builder kwargs, ht, frozen, steps)
CASES = {
    "d2q9_droplet_nht": (d2(nx=32, ny=32, btype=3, write_interval=100),
                         dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=8.0), False, False, 300),
    "d2q9_bubble_thermal": (d2(nx=24, ny=24, btype=3, write_interval=50),
                            dict(T=0.85 * TC, r_in=RV, r_out=RL, radius=6.0), True, False, 200),
    "d2q9_column_fixedT": (d2(nx=5, ny=40, btype=1, g=(0.0, -1e-6, 0.0), rho0=1.0, write_interval=100,
                              f_bc={"Y0": BC("bounceBack"), "Y1": BC("bounceBack")},
                              g_bc={"Y0": BC("fixedT", (0.95 * TC,)), "Y1": BC("fixedT", (0.8 * TC,))}),
                           dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=0, column=14.0), True, False, 300),
    "d2q9_column_liTempMRT": (d2(nx=5, ny=40, btype=1, g_model=3, g=(0.0, -1e-6, 0.0), write_interval=100,
                                 f_bc={"Y0": BC("bounceBack"), "Y1": BC("bounceBack")},
                                 g_bc={"Y0": BC("fixedT", (0.95 * TC,)), "Y1": BC("fixedT", (0.8 * TC,))}),
                              dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=0, column=14.0), True, False, 300),
    "d2q9_surface_tension": (d2(nx=24, ny=24, btype=3, f_st=1, f_kappa=0.3, write_interval=100),
                             dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=6.0), False, False, 200),
    "d2q9_channel_fixedU_outflow": (d2(nx=16, ny=16, btype=0, write_interval=50,
                                       f_bc={"X0": BC("fixedU", (0.01, 0.0, 0.0)), "X1": BC("outflow"),
                                             "Y0": BC("bounceBack"), "Y1": BC("bounceBack")}),
                                    dict(T=0.85 * TC, r_in=RL, r_out=RL, radius=4.0), False, False, 150),
    "d2q9_frozen_thermal": (d2(nx=16, ny=16, btype=3, write_interval=40),
                            dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=5.0), True, True, 80),
    "d2q9_carnahan_starling": (d2(nx=16, ny=16, btype=3, eos=1, a=0.02, b=0.4, write_interval=20),
                               dict(T=0.02, r_in=0.5, r_out=0.2, radius=4.0), False, False, 30),
    "d2q9_peng_robinson": (d2(nx=16, ny=16, btype=3, eos=2, a=0.02, b=0.4, w_pr=0.3, write_interval=20),
                           dict(T=0.02, r_in=0.5, r_out=0.2, radius=4.0), False, False, 30),
    # the EOS block with R, g, rho_0 ... ahead of the model name: the first-substring
    # lookups then read them as written (gravity and buoyancy reference exercised)
    "d2q9_carnahan_starling_keys_first": (d2(nx=16, ny=20, btype=3, eos=1, a=1.0, b=4.0, g=(0.0, -2e-6, 0.0), rho0=1.0,
                                             eos_keys_first=True, write_interval=20),
                                          dict(T=0.85 * TC_CS, r_in=CS_85[1], r_out=CS_85[0], radius=4.5), False,
                                          False, 60),
    "d3q15_peng_robinson_keys_first": (d3(nx=10, ny=10, nz=10, btype=5, eos=2, a=2.0 / 49.0, b=2.0 / 21.0,
                                          w_pr=0.344, R=1.0, g=(0.0, 0.0, -1e-5), rho0=1.0, eos_keys_first=True,
                                          write_interval=20),
                                       dict(T=0.85 * TC_PR, r_in=PR_85[0], r_out=PR_85[1], radius=3.0), False,
                                       False, 60),
    "d3q15_bubble_nht": (d3(nx=12, ny=12, nz=12, btype=5, write_interval=50),
                         dict(T=0.85 * TC, r_in=RV, r_out=RL, radius=3.5), False, False, 100),
    "d3q15_bubble_thermal": (d3(nx=12, ny=12, nz=12, btype=5, write_interval=50),
                             dict(T=0.85 * TC, r_in=RV, r_out=RL, radius=3.5), True, False, 100),
    "d3q15_thermal_column": (d3(nx=6, ny=6, nz=16, btype=3, g=(0.0, 0.0, -1e-6), rho0=1.0, write_interval=40,
                                f_bc={"Z0": BC("bounceBack"), "Z1": BC("bounceBack")},
                                g_bc={"Z0": BC("fixedRandomT", (0.95 * TC,), pert=0.5, duration=50),
                                      "Z1": BC("fixedT", (0.8 * TC,))}),
                             dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=0, column=5.0), True, False, 120),
    # two fixedRandomT lists that stop drawing at different steps (fixedRandomT.c:79 is per list)
    "d2q9_column_two_frdt": (d2(nx=5, ny=40, btype=1, g=(0.0, -1e-6, 0.0), rho0=1.0, write_interval=100,
                                f_bc={"Y0": BC("bounceBack"), "Y1": BC("bounceBack")},
                                g_bc={"Y0": BC("fixedRandomT", (0.95 * TC,), pert=0.5, duration=7),
                                      "Y1": BC("fixedRandomT", (0.8 * TC,), pert=1.0, duration=40)}),
                             dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=0, column=14.0), True, False, 120),
    # a fixedRandomT list without pert never draws, next to one that does
    "d3q15_column_frdt_nopert": (d3(nx=6, ny=6, nz=16, btype=3, g=(0.0, 0.0, -1e-6), rho0=1.0, write_interval=40,
                                    f_bc={"Z0": BC("bounceBack"), "Z1": BC("bounceBack")},
                                    g_bc={"Z0": BC("fixedRandomT", (0.95 * TC,), pert=0.5, duration=50),
                                          "Z1": BC("fixedRandomT", (0.8 * TC,))}),
                                 dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=0, column=5.0), True, False, 80),
    "d3q15_surface_tension": (d3(nx=10, ny=10, nz=10, btype=5, f_st=1, f_kappa=0.3, write_interval=50),
                              dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=3.0), False, False, 60),
    "d3q15_closed_box": (d3(nx=8, ny=8, nz=8, btype=0, write_interval=30,
                            f_bc={p: BC("bounceBack") for p in ("X0", "X1", "Y0", "Y1", "Z0", "Z1")},
                            g_bc={p: BC("bounceBack") for p in ("X0", "X1", "Y0", "Y1", "Z0", "Z1")}),
                         dict(T=0.85 * TC, r_in=RL, r_out=RV, radius=2.5), True, False, 60),
}

FIELDS = ("rho", "U", "T", "f", "g", "Fi", "Ft", "ss", "p")


@pytest.mark.parametrize("name", list(CASES))
def test_full_run_matches_oracle(name):
    spec, fk, ht, frozen, steps = CASES[name]
    print("case", name, flush=True)
    fields = sphere_fields(spec, seed=7, **fk)
    c, o = pair(spec, fields, precision=M.FP64, ht=ht, frozen=frozen)
    done = c.step(steps)
    n = o.run(steps)
    assert done == n == steps, (done, n)
    assert o.error()[0] == 0, o.error()
    worst = 0.0
    for f in FIELDS:
        a, b = c.get(f), o.get(f)
        if f == "p" and spec.write_interval > steps:
            continue
        r = max_rel(a, b)
        exact = bool(np.array_equal(a, b))
        print("  %-3s max_rel %.3e exact %s" % (f, r, exact), flush=True)
        worst = max(worst, r)
        assert np.isfinite(a).all(), f
    assert c.get_info().rho0 == pytest.approx(o.rho0, rel=1e-15, abs=0)
    assert worst <= TOL, worst


ROUTINES = [("collision", 0), ("stream", 0), ("update_boundaries", 0), ("collision", 1), ("stream", 1),
            ("update_boundaries", 1), ("macro_density", None), ("heat_source", None),
            ("macro_temperature", None), ("inter_force", None), ("macro_velocity", None),
            ("macro_pressure", None)]


@pytest.mark.parametrize("name", ["d2q9_column_fixedT", "d3q15_thermal_column", "d2q9_channel_fixedU_outflow",
                                  "d2q9_surface_tension"])
def test_each_routine_matches_oracle(name):
    """Routine by routine from the same state, so a mismatch names its routine."""
    spec, fk, ht, frozen, steps = CASES[name]
    fields = sphere_fields(spec, seed=3, **fk)
    c, o = pair(spec, fields, precision=M.FP64, ht=ht, frozen=frozen)
    c.step(5)
    o.run(5)
    for rname, arg in ROUTINES:
        args = () if arg is None else (arg,)
        getattr(c, rname)(*args)
        o.call(rname, *args)
        for f in FIELDS:
            r = max_rel(c.get(f), o.get(f))
            assert r <= TOL, (rname, arg, f, r)


def test_fp32_tracks_fp64_oracle():
    spec, fk, ht, frozen, steps = CASES["d2q9_droplet_nht"]
    fields = sphere_fields(spec, seed=7, **fk)
    c, o = pair(spec, fields, precision=M.FP32, ht=ht, frozen=frozen)
    c.step(steps)
    o.run(steps)
    rho, rho_o = c.get("rho"), o.get("rho")
    err = np.max(np.abs(rho - rho_o)) / np.max(np.abs(rho_o))
    print("fp32 rho rel err %.3e" % err)
    assert err < 1e-5


def test_nan_at_write_step_is_reported():
    """writeScalarField.c:16-24: a NaN field at a write step is an error, not a crash."""
    spec, fk, ht, frozen, steps = CASES["d2q9_droplet_nht"]
    fields = sphere_fields(spec, seed=7, **fk)
    rho = fields[0].copy()
    rho[10] = -0.01     # p - rho cs2 > 0 there: potential() takes the root of a negative number
    c, o = pair(spec, (rho, fields[1], fields[2]), precision=M.FP64, ht=ht, frozen=frozen)
    with pytest.raises(M.MplbError) as e:
        c.step(steps)
    assert e.value.code == -5
    assert o.run(steps) == -1


def _maxwell(tr):
    from multiphaselb_amd.initfields import maxwell
    return maxwell(tr * TC, a=0.5, b=4.0)


# The benchmark configurations' own shapes, 1000 steps in FP64: the D2Q9 liMRT droplet
# at full size (128^2, R = 30, T = 0.85 Tc), the D3Q15 bubble (xuMRT scheme, T = 0.8 Tc) and
# the D3Q15 thermal column with gravity, bounceBack floor and ceiling, fixedRandomT floor
# and fixedT ceiling (xuMRT + myMRT) at reduced sizes, all from Maxwell densities.
LONG = {
    "c0_droplet_128": (d2(nx=128, ny=128, btype=3, write_interval=500),
                       dict(T=0.85 * TC, r_in=_maxwell(0.85)[1], r_out=_maxwell(0.85)[0], radius=30.0), False),
    "c1_bubble_40": (d3(nx=40, ny=40, nz=40, btype=5, write_interval=500),
                     dict(T=0.8 * TC, r_in=_maxwell(0.8)[0], r_out=_maxwell(0.8)[1], radius=6.25), False),
    "c2_column_16x16x48": (d3(nx=16, ny=16, nz=48, btype=3, g=(0.0, 0.0, -1e-5), write_interval=500,
                              f_bc={"Z0": BC("bounceBack"), "Z1": BC("bounceBack")},
                              g_bc={"Z0": BC("fixedRandomT", (0.95 * TC,), pert=0.5, duration=10 ** 6),
                                    "Z1": BC("fixedT", (0.8 * TC,))}),
                           dict(T=0.8 * TC, r_in=_maxwell(0.8)[1], r_out=_maxwell(0.8)[0], radius=0, column=16.0),
                           True),
}


@pytest.mark.parametrize("name", list(LONG))
def test_config_shapes_1000_steps_match_oracle(name):
    spec, fk, ht = LONG[name]
    print("case", name, flush=True)
    fields = sphere_fields(spec, seed=11, **fk)
    c, o = pair(spec, fields, precision=M.FP64, ht=ht)
    assert c.step(1000) == 1000 and o.run(1000) == 1000
    assert o.error()[0] == 0, o.error()
    for f in ("rho", "U", "T", "f", "g", "p"):
        a, b = c.get(f), o.get(f)
        r = max_rel(a, b)
        print("  %-3s max_rel %.3e exact %s" % (f, r, bool(np.array_equal(a, b))), flush=True)
        assert np.isfinite(a).all() and r <= TOL, (f, r)



def _config0(a, b):
    from multiphaselb_amd.initfields import maxwell, vdw_critical
    T = 0.85 * vdw_critical(a, b)[1]
    rv, rl = maxwell(T, a=a, b=b)
    spec = d2(nx=128, ny=128, btype=3, a=a, b=b, g_model=None, write_interval=5000)
    return spec, sphere_fields(spec, T=T, r_in=rl, r_out=rv, radius=30.0, seed=20240601), rv, rl


def test_config0_droplet_20000_steps_match_oracle():
    """The D2Q9 static droplet configuration at its own size and length: 128^2 fully
    periodic, van der Waals at T = 0.85 Tc, radius 30 with a tanh profile of width 5
    between the Maxwell densities, seeded +-1e-4 noise, zero velocity, G = -1, s_nu = 1,
    liMRT (Li force scheme), 20000 FP64 steps with write steps every 5000 (pressure, rho_0
    re-average): every field within 1e-12 of the oracle.  a = 0.5, b = 4 (the reference
    examples' constants): with a = 9/49, b = 2/21 the potential is undefined at the
    vapour density (next test)."""
    spec, fields, rv, rl = _config0(0.5, 4.0)
    c, o = pair(spec, fields, precision=M.FP64, ht=False)
    print("case c0 128^2, Maxwell densities %.6f / %.6f" % (rv, rl), flush=True)
    assert c.step(20000) == 20000 and o.run(20000) == 20000
    assert o.error()[0] == 0, o.error()
    for f in ("rho", "U", "f", "p"):
        x, y = c.get(f), o.get(f)
        r = max_rel(x, y)
        print("  %-3s max_rel %.3e exact %s" % (f, r, bool(np.array_equal(x, y))), flush=True)
        assert np.isfinite(x).all() and r <= TOL, (f, r)
    # the droplet is still there: liquid inside, vapour outside
    rho = c.get("rho").reshape(128, 128)
    assert rho[64, 64] > 0.9 * rl and rho[2, 2] < 1.2 * rv


def test_config0_literal_constants_are_undefined_as_in_the_reference():
    """With a = 9/49, b = 2/21, R = 1 at T = 0.85 Tc the Maxwell vapour density (1.119)
    lies where p - rho cs2 > 0, so psi = sqrt(2 (p - rho cs2) / G) (potential.c) is NaN
    there: after one step the library and the oracle hold NaN at exactly the same nodes
    and identical values elsewhere, and the write step refuses the field with the
    reference's message."""
    spec, fields, rv, rl = _config0(9.0 / 49.0, 2.0 / 21.0)
    c, o = pair(spec, fields, precision=M.FP64, ht=False)
    assert c.step(1) == 1 and o.run(1) == 1
    x, y = c.get("rho"), o.get("rho")
    nan = np.isnan(y)
    print("NaN nodes after one step: %d of %d" % (int(nan.sum()), nan.size), flush=True)
    assert nan.any() and not nan.all()
    np.testing.assert_array_equal(np.isnan(x), nan)
    np.testing.assert_array_equal(x[~nan], y[~nan])
    with pytest.raises(M.MplbError) as e:
        c.step(5000)
    assert "NaN" in str(e.value)

"""Case-level helpers of runs to steady state and of the warm start, through the C ABI:
the radial density writer against its host restatement, the mass-matching radius
search, and the device density-change reduction against host differences."""
import numpy as np
import pytest

import multiphaselb_amd as M
from multiphaselb_amd import warmstart as W
from _cases import make_case_dir, sphere_fields
from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def d3_case(precision=M.FP64, fused=False):
    spec, fk, ht, frozen, steps = CASES["d3q15_bubble_nht"]
    c = M.Case(make_case_dir(spec), precision=precision, heat_transfer=False, fused=fused)
    return spec, fk, c


@pytest.mark.parametrize("precision", [M.FP64, M.FP32])
def test_radial_density_matches_host(precision):
    spec, fk, c = d3_case(precision)
    n = spec.nx
    cen = ((n - 1) / 2.0, (n - 1) / 2.0, (n - 1) / 2.0)
    y = 0.5 * (1 - np.tanh(W.DK / 2.5))
    prof = W.profile_density(y, 0.03, 0.16)
    fine = W.fine_profile(prof)
    r = W.node_radius((n, n, n), cen)
    sites = W.distinct_sites((n, n, n))            # periodicXYZ: duplicates left out
    for R in (2.5, 3.7, 5.0):
        mass = c.set_radial_density(fine, W.FINE_D0, W.FINE_DD, cen, R)
        host = W._interp(fine, r - R)
        if precision == M.FP32:
            host = host.astype(np.float32).astype(np.float64)
        got = c.get("rho")
        np.testing.assert_allclose(got, host, rtol=1e-15, atol=0)
        assert abs(mass - host[sites].sum()) <= 1e-12 * host.sum()
    # the radius search reproduces a target mass
    target = float(W._interp(fine, r - 4.2)[sites].sum())
    R = W.raster_case(c, prof, cen, target, 3.0, int(sites.sum()))
    assert abs(R - 4.2) < 1e-6


@pytest.mark.parametrize("fused", [False, True])
def test_density_change_matches_host_difference(fused):
    spec, fk, c = d3_case(M.FP64 if not fused else M.FP32, fused=fused)
    rho, U, T = sphere_fields(spec, seed=3, **fk)
    c.set("rho", rho)
    c.set("U", U)
    c.set("T", T)
    c.init()
    assert np.isnan(c.density_change())
    before = c.get("rho")
    c.step(7)
    d = c.density_change()
    after = c.get("rho")
    assert d == np.max(np.abs(after - before))
    assert d > 0
    c.step(3)
    d2 = c.density_change()
    assert d2 == np.max(np.abs(c.get("rho") - after))


def test_radial_density_survives_field_transfers_on_the_fused_path():
    """After steps on the fused path, a radial density written into the case must be
    what mplb_init starts from, even when U and T are set afterwards."""
    spec, fk, c = d3_case(M.FP32, fused=True)
    rho, U, T = sphere_fields(spec, seed=3, **fk)
    c.set("rho", rho)
    c.set("U", U)
    c.set("T", T)
    c.init()
    c.step(20)
    n = spec.nx
    cen = ((n - 1) / 2.0,) * 3
    prof = W.profile_density(0.5 * (1 - np.tanh(W.DK / 2.5)), 0.03, 0.16)
    fine = W.fine_profile(prof)
    c.set_radial_density(fine, W.FINE_D0, W.FINE_DD, cen, 3.3)
    c.set("U", np.zeros_like(U))
    c.set("T", T)
    expect = W._interp(fine, W.node_radius((n, n, n), cen) - 3.3).astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(c.get("rho"), expect)
    c.init()
    np.testing.assert_array_equal(c.get("rho"), expect)


def test_uniform_raster_carries_the_mass():
    """A dissolved-state proposal is written as the uniform density mass / sites."""
    spec, fk, c = d3_case(M.FP64)
    n = spec.nx
    cen = ((n - 1) / 2.0,) * 3
    sites = W.distinct_sites((n, n, n))
    mass = 0.137 * sites.sum() * (1.0 + 1e-3)
    assert W.raster_case(c, np.full(W.NPROF, 0.1), cen, mass, 3.0, int(sites.sum()), uniform=True) is None
    got = c.get("rho")
    assert np.all(got == got[0])
    assert abs(got[sites].sum() / mass - 1.0) < 1e-14

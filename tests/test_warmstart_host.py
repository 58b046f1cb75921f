"""Host side of the learned warm start: profile extraction from steady fields (interface
kept or dissolved), the network target round trip, the proposal range check and the
mass-matching rasteriser, on synthetic fields whose answers are known by construction."""
import numpy as np
import pytest

from multiphaselb_amd import warmstart as W

L = 48
CEN = ((L - 1) / 2.0, (L - 1) / 2.0)
RV, RL = 0.02, 0.16


def tanh_field(R, droplet, width=4.0):
    r = W.node_radius((L, L), CEN)
    t = np.tanh(2.0 * (r - R) / width)
    y = 0.5 * (1.0 - t) if droplet else 0.5 * (1.0 + t)
    return RV + (RL - RV) * y, r


@pytest.mark.parametrize("droplet", [True, False])
def test_extract_profile_finds_radius_and_shape(droplet):
    rho, r = tanh_field(11.3, droplet)
    y, R = W.extract_profile(rho, r, RV, RL, droplet)
    # shells of width 0.05 average nodes at slightly different radii: the crossing and the
    # profile are exact to that resolution
    assert abs(R - 11.3) < 0.05
    want = 0.5 * (1.0 - np.tanh(2.0 * W.DK / 4.0)) if droplet else 0.5 * (1.0 + np.tanh(2.0 * W.DK / 4.0))
    inner = np.abs(W.DK) <= 6.0
    assert np.max(np.abs(y[inner] - want[inner])) < 0.03


def test_steady_profile_of_a_dissolved_case_is_flat_at_the_mean():
    sites = W.distinct_sites((L, L))
    rng = np.random.default_rng(3)
    rho = np.full(L * L, 0.151) + 1e-6 * rng.standard_normal(L * L)
    r = W.node_radius((L, L), CEN)
    y, kept = W.steady_profile(rho, r, sites, RV, RL, droplet=False)
    assert not kept
    assert np.all(y == y[0])
    assert abs(RV + (RL - RV) * y[0] - rho[sites].mean()) < 1e-15
    # and a field that kept its interface gives the interface profile
    rho2, _ = tanh_field(9.0, True)
    y2, kept2 = W.steady_profile(rho2, r, sites, RV, RL, droplet=True)
    assert kept2 and y2[0] > 0.9 and y2[-1] < 0.1


def test_steady_profile_rejects_a_field_that_is_neither():
    sites = W.distinct_sites((L, L))
    r = W.node_radius((L, L), CEN)
    # two density levels split along x: no radial interface, not uniform
    rho = np.where(np.arange(L * L) % L < L // 2, RL, RV).astype(np.float64)
    rho[(r < 2.0)] = rho[r > r.max() - 2.0].mean()
    with pytest.raises(ValueError):
        W.steady_profile(rho, r, sites, RV, RL, droplet=True)


@pytest.mark.parametrize("droplet", [True, False])
def test_target_round_trip(droplet):
    y = np.linspace(0.0, 1.0, W.NPROF) ** 2
    o = W.target_from_profile(y, droplet)
    np.testing.assert_allclose(W.profile_from_output(o, droplet), y, rtol=0, atol=1e-15)
    # the cold profile itself is the zero correction
    assert np.all(W.target_from_profile(W.cold_profile(droplet), droplet) == 0.0)


def test_proposal_range_check():
    ok = np.linspace(RV, RL, W.NPROF)
    assert W.proposal_ok(ok, RV, RL)
    assert not W.proposal_ok(np.where(np.arange(W.NPROF) == 3, -1e-3, ok), RV, RL)
    assert not W.proposal_ok(np.where(np.arange(W.NPROF) == 3, np.nan, ok), RV, RL)
    assert not W.proposal_ok(ok * 1.3, RV, RL)


def test_raster_matches_mass_for_interface_and_flat_profiles():
    sites = W.distinct_sites((L, L))
    r = W.node_radius((L, L), CEN)
    prof = W.profile_density(0.5 * (1.0 - np.tanh(2.0 * W.DK / 4.0)), RV, RL)
    target = float(W._interp(W.fine_profile(prof), r - 10.7)[sites].sum())
    rho, R = W.raster_host(prof, r, target, 9.0, sites)
    assert abs(R - 10.7) < 1e-6
    assert abs(rho[sites].sum() - target) <= 1e-10 * target
    # a flat proposal cannot move an interface: the mass goes on as a uniform offset
    flat = np.full(W.NPROF, 0.1)
    rho2, _ = W.raster_host(flat, r, 0.12 * sites.sum(), 9.0, sites)
    np.testing.assert_allclose(rho2, 0.12, rtol=1e-12)


def test_flat_proposals_are_recognised():
    flat = np.full(W.NPROF, 0.15) + 0.001 * np.sin(np.arange(W.NPROF))
    assert W.is_flat(flat, RV, RL)
    step = W.profile_density(0.5 * (1.0 - np.tanh(2.0 * W.DK / 4.0)), RV, RL)
    assert not W.is_flat(step, RV, RL)


class _Fixed:
    """Stand-in network returning a fixed output row."""

    def __init__(self, o):
        self.o = np.asarray(o, np.float32)

    def forward(self, x):
        return np.tile(self.o, (len(x), 1))


def test_flat_targets_and_two_network_proposal():
    y_flat = np.full(W.NPROF, 0.9)
    y_step = 0.5 * (1.0 + np.tanh(2.0 * W.DK / 4.0))
    x_b = W.features(0.8, 20.0, 128, False, 3)
    X = np.stack([x_b, x_b])
    Y = np.stack([np.concatenate([W.target_from_profile(y, False), np.zeros(W.NPROF)]) for y in (y_flat, y_step)])
    np.testing.assert_array_equal(W.flat_targets(X, Y), [True, False])
    out = lambda y: np.concatenate([W.target_from_profile(y, False), np.zeros(W.NPROF)])
    # the first network says "dissolves": its flat profile is the proposal
    prof, flat = W.propose(_Fixed(out(y_flat)), _Fixed(out(y_step)), x_b, False, RV, RL)
    assert flat and np.allclose(prof, W.profile_density(y_flat, RV, RL), atol=1e-6)
    # otherwise the interface network's profile is used
    y_other = 0.5 * (1.0 + np.tanh(2.0 * W.DK / 6.0))
    prof, flat = W.propose(_Fixed(out(y_step)), _Fixed(out(y_other)), x_b, False, RV, RL)
    assert not flat and np.allclose(prof, W.profile_density(y_other, RV, RL), atol=1e-6)


def test_plateau_fit_recovers_a_smooth_law_and_needs_enough_samples():
    """The plateau model is least squares on s^i and s^i / R (s = sqrt(1 - T/Tc)): a law
    in that span comes back exactly (to rounding), a flag with too few samples gets no
    fit (the network's own plateau is then kept)."""
    rng = np.random.default_rng(5)
    n = 60
    tr = 0.6 + 0.35 * rng.random(n)
    R = 15.0 + 15.0 * rng.random(n)
    s = np.sqrt(1.0 - tr)
    law = 1.0 - 0.08 * s ** 2 + 0.01 * s ** 3 + 0.02 * s / R
    droplet = np.zeros(n, bool)
    droplet[:10] = True                      # fewer than PLATEAU_MIN_SAMPLES droplets
    fit, rms = W.fit_plateau(tr, R, droplet, law)
    assert fit[True] is None and rms[True] is None
    assert rms[False] < 1e-12
    for t, rr in ((0.65, 17.0), (0.8, 25.0), (0.93, 29.0)):
        ss = np.sqrt(1.0 - t)
        assert abs(W.plateau_at(fit, t, rr, False) - (1.0 - 0.08 * ss ** 2 + 0.01 * ss ** 3 + 0.02 * ss / rr)) < 1e-10
    assert W.plateau_at(fit, 0.7, 20.0, True) is None


def test_plateau_profile_moves_only_the_plateaus():
    """plateau_profile puts the fitted values on the profile's ends (the smooth windows
    reach 1 - 2e-6 there) and leaves the interface band (|d| < 3, window weight <= 0.02)
    as proposed up to 2% of the plateau change; without fits the profile is unchanged."""
    rng = np.random.default_rng(6)
    n = 40
    tr, R = 0.6 + 0.3 * rng.random(n), 16.0 + 14.0 * rng.random(n)
    out_fit, _ = W.fit_plateau(tr, R, np.zeros(n, bool), np.full(n, 0.99))
    in_fit, _ = W.fit_plateau(tr, R, np.zeros(n, bool), np.full(n, -0.01))
    y = W.cold_profile(False)
    yp = W.plateau_profile(y, out_fit, in_fit, 0.75, 22.0, False)
    assert abs(yp[-1] - 0.99) < 1e-7 and abs(yp[0] - (-0.01)) < 1e-7
    band = np.abs(W.DK) < 3.0
    assert np.max(np.abs(yp[band] - y[band])) < 0.02 * 0.01
    assert np.array_equal(W.plateau_profile(y, {}, {}, 0.75, 22.0, False), y)

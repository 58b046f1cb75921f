"""Learned warm start for static droplet/bubble cases.

A small network (Mlp: BF16 matrix-core products, FP32 accumulation, trained on the GPU
by the library's own forward/backward/Adam kernels) maps the case parameters to the
radial profile of the steady density around the interface.  The proposal is written
onto the lattice with the radius that gives its distinct sites the same mass as the
case's cold initial field, and the solver then runs from it to its own steady criterion:
the network only chooses the starting point, the solver decides the result, and a warm
run is checked against the cold run of the same case.

Profile: 64 samples of the density at signed distances DK from the interface (d < 0
inside; 0.25 apart for |d| < 6, 1.36 apart out to |d| = 16), as
y = (rho - rho_v) / (rho_l - rho_v) with the Maxwell densities of the case's
temperature.  The network predicts the correction to the cold start's own profile (tanh
of width 5 between the Maxwell densities), o = (y - y_cold) / RES_SCALE, so its
BF16-level relative error lands on a correction of a few percent rather than on the
whole density jump.  A second block of 64 outputs carries the temperature profile as
100 (T / T0 - 1) (zero for isothermal cases).  For the lattice the 64 samples are
interpolated (monotone cubic) onto a uniform 0.05 grid; beyond |d| = 16 the end values
continue.

A droplet or bubble that dissolves in its closed box ends as a uniform field at the
box's mean density; its target is that flat profile, so a network trained on every
steady state also learns which cases keep an interface.  The interface profile itself
comes from a second network trained on the kept interfaces alone (propose).  A proposal
outside the physical density range is not used (proposal_ok): the run then starts cold.
"""
import numpy as np
from scipy.interpolate import PchipInterpolator

from .initfields import maxwell, vdw_critical

NPROF = 64
DK = np.concatenate([np.linspace(-16.0, -6.5, 8), np.linspace(-5.875, 5.875, 48), np.linspace(6.5, 16.0, 8)])
FINE_D0, FINE_DD = -16.0, 0.05
FINE_N = 641
N_IN = 9
N_OUT = 2 * NPROF
WIDTHS = [N_IN, 256, 256, 256, N_OUT]
RES_SCALE = 0.05
COLD_WIDTH = 5.0


def features(tr, R, L, droplet, dim, a=0.5, b=4.0, G=-1.0, tau=1.0):
    """Network input: T/Tc, the EOS constants a and b, G, tau, radius/L, droplet or
    bubble, the lattice dimension and 8/R (the interface curvature scale), each shifted
    and scaled to about [-1, 1] over the generator's ranges."""
    return np.array([(tr - 0.775) / 0.175, a / 0.5 - 1.0, b / 4.0 - 1.0, G + 1.0, tau - 1.0,
                     (R / L - 0.15) / 0.1, 1.0 if droplet else -1.0, 2.0 * (dim - 2.5), 8.0 / R - 0.6],
                    dtype=np.float32)


def node_radius(shape, centre):
    """Distance of every stored node (node-major, x fastest) from the centre."""
    idx = np.indices(shape[::-1]).reshape(len(shape), -1)[::-1]      # x, y(, z) per node
    return np.sqrt(sum((idx[d] - centre[d]) ** 2 for d in range(len(shape))))


def distinct_sites(shape):
    """Mask of the distinct sites of a fully periodic box with duplicate end nodes
    (node-major, x fastest): the last node along each axis repeats the first."""
    idx = np.indices(shape[::-1]).reshape(len(shape), -1)[::-1]
    m = np.ones(idx.shape[1], bool)
    for d, n in enumerate(shape):
        m &= idx[d] < n - 1
    return m


def _shell_profile(rho, r, width):
    """Density averaged over shells of the given width, at the shells' mean radius."""
    edges = np.arange(0.0, r.max() + width, width)
    cnt, _ = np.histogram(r, edges)
    tot, _ = np.histogram(r, edges, weights=rho)
    rs, _ = np.histogram(r, edges, weights=r)
    ok = cnt > 0
    return rs[ok] / cnt[ok], tot[ok] / cnt[ok]


def extract_profile(rho, r, rho_v, rho_l, droplet):
    """Target profile of one steady field: the interface radius where the shell-averaged
    density crosses the mid density, then the shell-averaged density at distances DK
    from it, normalised (y).  Raises ValueError when the field has no interface left
    (the droplet or bubble dissolved)."""
    rm, pm = _shell_profile(rho, r, 0.05)
    rin, rout = pm[rm < rm[0] + 1.5].mean(), pm[rm > rm[-1] - 2.0].mean()
    if abs(rin - rout) < 0.25 * (rho_l - rho_v):
        raise ValueError("no interface left")
    mid = 0.5 * (rin + rout)
    s = np.sign(pm - mid)
    cross = np.nonzero(s[:-1] != s[1:])[0]
    if cross.size == 0:
        raise ValueError("no interface found")
    i = cross[0]
    R = rm[i] + (mid - pm[i]) * (rm[i + 1] - rm[i]) / (pm[i + 1] - pm[i])
    y = np.interp(R + DK, rm, pm)
    return (y - rho_v) / (rho_l - rho_v), R


def steady_profile(rho, r, sites, rho_v, rho_l, droplet, flat_tol=0.05):
    """Target profile of one steady field, whether it kept its interface or not: the
    interface profile of extract_profile, or, for a droplet or bubble that dissolved, the
    flat profile of the field's mean density over the distinct sites (the uniform state a
    closed box ends in).  Returns (y, kept); raises ValueError for a field that is neither
    (no interface, yet not uniform within flat_tol of the density jump)."""
    try:
        y, _ = extract_profile(rho, r, rho_v, rho_l, droplet)
        return y, True
    except ValueError:
        pass
    span = float(np.max(rho) - np.min(rho))
    if not np.isfinite(span) or span > flat_tol * (rho_l - rho_v):
        raise ValueError("field is neither an interface nor uniform")
    mean = float(np.mean(rho[sites]))
    return np.full(NPROF, (mean - rho_v) / (rho_l - rho_v)), False


def proposal_ok(prof, rho_v, rho_l):
    """A proposal the solver may start from: finite, and inside [rho_v / 4, 5 rho_l / 4]
    (a density at or below zero has no pseudopotential)."""
    prof = np.asarray(prof, np.float64)
    return bool(np.isfinite(prof).all() and prof.min() > 0.25 * rho_v and prof.max() < 1.25 * rho_l)


def cold_profile(droplet):
    """y at DK of the cold start: 0.5 (in + out) + 0.5 (out - in) tanh(2 d / W)."""
    t = np.tanh(2.0 * DK / COLD_WIDTH)
    return 0.5 * (1.0 - t) if droplet else 0.5 * (1.0 + t)


def target_from_profile(y, droplet):
    """Network target of a steady profile y: the scaled correction to the cold profile."""
    return (np.asarray(y, np.float64) - cold_profile(droplet)) / RES_SCALE


def profile_from_output(o, droplet):
    return cold_profile(droplet) + RES_SCALE * np.asarray(o, np.float64)


def profile_density(y, rho_v, rho_l):
    return rho_v + (rho_l - rho_v) * np.asarray(y, np.float64)


def fine_profile(prof):
    """The 64 samples at DK interpolated (monotone cubic) onto the uniform fine grid."""
    d = FINE_D0 + FINE_DD * np.arange(FINE_N)
    return PchipInterpolator(DK, np.asarray(prof, np.float64))(d)


def _interp(fine, d):
    t = np.clip((d - FINE_D0) / FINE_DD, 0.0, FINE_N - 1.0)
    q = np.minimum(t.astype(np.int64), FINE_N - 2)
    w = t - q
    return fine[q] + w * (fine[q + 1] - fine[q])


def _secant(mass_of, target, R0, tol=1e-12, iters=40):
    Ra, Rb = R0, R0 * 1.02 + 0.05
    ma, mb = mass_of(Ra) - target, mass_of(Rb) - target
    for _ in range(iters):
        if mb == ma:
            break
        Rc = Rb - mb * (Rb - Ra) / (mb - ma)
        Ra, ma = Rb, mb
        Rb, mb = Rc, mass_of(Rc) - target
        if abs(mb) <= tol * abs(target):
            break
    return Rb, mb


FLAT_TOL = 0.1


def is_flat(prof, rho_v, rho_l):
    """A proposal without an interface: its density range is below FLAT_TOL of the jump
    (the network's prediction that the droplet or bubble dissolves)."""
    return float(np.ptp(np.asarray(prof, np.float64))) < FLAT_TOL * (rho_l - rho_v)


def raster_host(prof, r, mass, R0, sites):
    """Density on the nodes at distances r from the centre whose distinct sites (mask)
    carry the mass `mass`: rho = profile(r - R), R from the secant method (numpy).
    When moving the interface cannot give that mass (a flat proposal, or a radius that
    leaves the box), the remainder is spread uniformly over the sites."""
    fine = fine_profile(prof)
    R, err = _secant(lambda R: float(_interp(fine, r[sites] - R).sum()), mass, R0)
    if not np.isfinite(R) or not 0.25 * R0 <= R <= 4.0 * R0 or abs(err) > 1e-10 * abs(mass):
        R = R if np.isfinite(R) and 0.25 * R0 <= R <= 4.0 * R0 else R0
        fine = fine + (mass - float(_interp(fine, r[sites] - R).sum())) / float(sites.sum())
    return _interp(fine, r - R), R


def raster_case(case, prof, centre, mass, R0, nsites, uniform=False):
    """Same on a case's device density (mplb_set_radial_density, which sums the distinct
    sites of the case's periodic box; nsites of them); returns R.  uniform: the proposal
    is a dissolved state (is_flat), written as the uniform density mass / nsites (its
    residual ripple would only add a transient for the solver to damp); returns None."""
    if uniform:
        case.set_radial_density(np.full(FINE_N, mass / nsites), FINE_D0, FINE_DD, centre, R0)
        return None
    fine = fine_profile(prof)
    R, err = _secant(lambda R: case.set_radial_density(fine, FINE_D0, FINE_DD, centre, R), mass, R0)
    if not np.isfinite(R) or not 0.25 * R0 <= R <= 4.0 * R0 or abs(err) > 1e-10 * abs(mass):
        R = R if np.isfinite(R) and 0.25 * R0 <= R <= 4.0 * R0 else R0
        fine = fine + (mass - case.set_radial_density(fine, FINE_D0, FINE_DD, centre, R)) / float(nsites)
    case.set_radial_density(fine, FINE_D0, FINE_DD, centre, R)
    return R


def flat_targets(X, Y, tol=1e-3):
    """Mask of the training samples whose target is a dissolved (flat) profile."""
    droplet = np.asarray(X)[:, 6] > 0
    cold = np.where(droplet[:, None], cold_profile(True)[None, :], cold_profile(False)[None, :])
    y = cold + RES_SCALE * np.asarray(Y, np.float64)[:, :NPROF]
    return np.ptp(y, axis=1) < tol


def propose(net_all, net_kept, x, droplet, rho_v, rho_l):
    """Density profile proposed for one case (features x): the network trained on every
    steady state decides whether the case dissolves (a flat proposal, returned as is with
    flat = True); otherwise the network trained on the interface profiles alone gives the
    profile (its fit is not diluted by the flat targets; a list of such networks gives
    their mean).  Returns (profile, flat)."""
    x = np.asarray(x, np.float32)[None, :]
    prof = profile_density(profile_from_output(net_all.forward(x)[0, :NPROF], droplet), rho_v, rho_l)
    if is_flat(prof, rho_v, rho_l):
        return prof, True
    if net_kept is None:
        return prof, False
    # several interface networks (trained from different seeds): their mean output
    nets = net_kept if isinstance(net_kept, (list, tuple)) else [net_kept]
    o = np.mean([n.forward(x)[0, :NPROF] for n in nets], axis=0)
    return profile_density(profile_from_output(o, droplet), rho_v, rho_l), False


def init_params(widths, seed):
    """Glorot-uniform weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(len(widths) - 1):
        lim = np.sqrt(6.0 / (widths[l] + widths[l + 1]))
        parts += [rng.uniform(-lim, lim, widths[l] * widths[l + 1]), np.zeros(widths[l + 1])]
    return np.concatenate(parts).astype(np.float32)


def train(net, X, Y, iters=4000, seed=0, schedule=((0.4, 2e-3), (0.35, 6e-4), (0.25, 1.5e-4))):
    """Adam over random minibatches, in stages of decreasing learning rate (fraction of
    the iterations, rate); returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    for frac, lr in schedule:
        n = max(1, int(round(frac * iters)))
        idx = rng.integers(0, X.shape[0], n * net.batch).astype(np.int32)
        hist.append(net.train(X, Y, idx, lr=lr))
    return np.concatenate(hist)


# The far plateau (the phase outside the interface, which fills most of a closed box)
# decides how long a warm start takes: on 128^3 bubbles an error of 3e-4 of the density
# jump there turns the ~1600 steps of the exact steady profile into 5000-7000 (the
# interface has to move to restore the mass balance), while the same error inside, or a
# 3e-3 error in the interface shape, costs a few hundred.  The network fits the profile
# to ~3e-4, so a second model trained on the same solver runs gives the plateau: least
# squares on s^i (i <= 5) and s^i / R (i <= 3), s = sqrt(1 - T/Tc) (coexistence
# densities go as powers of s near the critical point) and R the steady interface radius
# (the Laplace shift goes as 1 / R), one fit for droplets and one for bubbles, on the 3D
# steady states that kept their interface.  On 154 steady 128^3 bubbles its
# leave-one-out error is 4e-5 of the jump (powers of T/Tc instead of s: 2e-4).
# The inside plateau gets the same kind of fit (it matters less: the phase inside holds
# little of the box's mass).
PLATEAU_POWERS = [(i, 0) for i in range(6)] + [(i, 1) for i in range(4)]
PLATEAU_MIN_SAMPLES = 24
WIN_OUT = 0.5 * (1.0 + np.tanh((DK - 6.0) / 1.5))     # 1 on the outside plateau, 0 inside
WIN_IN = 0.5 * (1.0 - np.tanh((DK + 6.0) / 1.5))      # 1 on the inside plateau, 0 outside


def _plateau_basis(tr, R):
    sq = np.sqrt(np.maximum(1.0 - np.asarray(tr, np.float64), 0.0))
    inv = 1.0 / np.asarray(R, np.float64)
    return np.stack([sq ** i * inv ** j for i, j in PLATEAU_POWERS], axis=-1)


def fit_plateau(tr, R, droplet, y_out):
    """Least-squares fit of a plateau value y_out (normalised, as the profile) of steady
    3D interfaces; returns {droplet flag: coefficients or None (too few samples)} and the
    fit's rms residual per flag."""
    tr, R, y_out = (np.asarray(v, np.float64) for v in (tr, R, y_out))
    droplet = np.asarray(droplet, bool)
    fit, rms = {}, {}
    for flag in (True, False):
        m = droplet == flag
        if m.sum() < PLATEAU_MIN_SAMPLES:
            fit[flag], rms[flag] = None, None
            continue
        A = _plateau_basis(tr[m], R[m])
        coef, *_ = np.linalg.lstsq(A, y_out[m], rcond=None)
        fit[flag] = coef
        rms[flag] = float(np.sqrt(np.mean((A @ coef - y_out[m]) ** 2)))
    return fit, rms


def plateau_at(fit, tr, R, droplet):
    """The fitted plateau (normalised) of a steady interface at radius R, or None."""
    coef = fit.get(bool(droplet)) if fit else None
    if coef is None:
        return None
    return float(_plateau_basis(tr, R) @ coef)


def plateau_profile(y, fit, fit_in, tr, R, droplet):
    """Profile y (normalised) with its outside plateau, and its inside plateau when fit_in
    has a fit for this flag, replaced by the fitted values at interface radius R."""
    y = np.asarray(y, np.float64)
    out = plateau_at(fit, tr, R, droplet)
    if out is not None:
        y = y + (out - y[-1]) * WIN_OUT
    inn = plateau_at(fit_in, tr, R, droplet)
    if inn is not None:
        y = y + (inn - y[0]) * WIN_IN
    return y


def raster_case_plateau(case, y, fit, tr, droplet, rho_v, rho_l, centre, mass, R0, nsites, fit_in=None):
    """raster_case for a proposed interface profile y (normalised) whose outside plateau
    (fit) and inside plateau (fit_in, if given) are replaced by the fitted ones at the
    interface radius R, R chosen so that the distinct sites carry the cold field's mass
    (secant on the device raster).  Falls back to raster_case with y itself when there is
    no outside fit for this flag; returns R."""
    y = np.asarray(y, np.float64)
    if plateau_at(fit, tr, R0, droplet) is None:
        return raster_case(case, profile_density(y, rho_v, rho_l), centre, mass, R0, nsites)

    def fine_at(R):
        return fine_profile(profile_density(plateau_profile(y, fit, fit_in, tr, R, droplet), rho_v, rho_l))

    R, err = _secant(lambda R: case.set_radial_density(fine_at(R), FINE_D0, FINE_DD, centre, R), mass, R0)
    if not np.isfinite(R) or not 0.25 * R0 <= R <= 4.0 * R0 or abs(err) > 1e-10 * abs(mass):
        return raster_case(case, profile_density(y, rho_v, rho_l), centre, mass, R0, nsites)
    case.set_radial_density(fine_at(R), FINE_D0, FINE_DD, centre, R)
    return R


def case_profile_densities(tr, a=0.5, b=4.0):
    rv, rl = maxwell(tr * vdw_critical(a, b)[1], a=a, b=b)
    return rv, rl

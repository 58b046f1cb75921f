"""Learned warm start for static droplet/bubble cases.

A small network (Mlp: BF16 matrix-core products, FP32 accumulation, trained on the GPU
by the library's own forward/backward/Adam kernels) maps the case parameters to the
radial profile of the steady density around the interface.  The proposal is written
onto the lattice with the radius that gives it the same total mass as the case's cold
initial field, and the solver then runs from it to its own steady criterion: the
network only chooses the starting point, the solver decides the result, and a warm run
is checked against the cold run of the same case.

Profile: 64 samples of the density at signed distance d_k = -15.75 + 0.5 k from the
interface (d < 0 inside), as y_k = (rho - rho_v) / (rho_l - rho_v) with the Maxwell
densities of the case's temperature; a second block of 64 outputs carries the
temperature profile as 100 (T / T0 - 1) (zero for isothermal cases).  Beyond
|d| = 15.75 the end values continue.
"""
import numpy as np

from .initfields import maxwell, vdw_critical

NPROF = 64
D0, DD = -15.75, 0.5
N_IN = 9
N_OUT = 2 * NPROF
WIDTHS = [N_IN, 256, 256, 256, N_OUT]


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


def extract_profile(rho, r, rho_v, rho_l, droplet):
    """Target profile of one steady field: the interface radius where the radially
    averaged density crosses the mid density, then the density averaged over shells of
    width 0.5 centred at d_k; returned normalised (y_k)."""
    rin = rho[r < max(2.0, r.min() + 1.5)].mean()
    rout = rho[r > r.max() - 2.0].mean()
    mid = 0.5 * (rin + rout)
    edges = np.arange(0.0, r.max() + 0.5, 0.5)
    cnt, _ = np.histogram(r, edges)
    tot, _ = np.histogram(r, edges, weights=rho)
    ok = cnt > 0
    rc = 0.5 * (edges[1:] + edges[:-1])[ok]
    prof = tot[ok] / cnt[ok]
    s = np.sign(prof - mid)
    cross = np.nonzero(s[:-1] != s[1:])[0]
    if cross.size == 0:
        raise ValueError("no interface found")
    i = cross[0]
    R = rc[i] + (mid - prof[i]) * (rc[i + 1] - rc[i]) / (prof[i + 1] - prof[i])
    d = r - R
    dk = D0 + DD * np.arange(NPROF)
    k = np.floor((d - (D0 - 0.5 * DD)) / DD).astype(np.int64)
    sel = (k >= 0) & (k < NPROF)
    c2 = np.bincount(k[sel], minlength=NPROF)
    t2 = np.bincount(k[sel], weights=rho[sel], minlength=NPROF)
    y = np.empty(NPROF)
    have = c2 > 0
    y[have] = t2[have] / c2[have]
    if not have.all():                 # shells without nodes (small radii): interpolate
        y[~have] = np.interp(dk[~have], dk[have], y[have])
    return (y - rho_v) / (rho_l - rho_v), R


def profile_density(y, rho_v, rho_l):
    return rho_v + (rho_l - rho_v) * np.asarray(y, np.float64)


def _interp(prof, d):
    t = np.clip((d - D0) / DD, 0.0, NPROF - 1.0)
    q = np.minimum(t.astype(np.int64), NPROF - 2)
    w = t - q
    return prof[q] + w * (prof[q + 1] - prof[q])


def _secant(mass_of, target, R0, tol=1e-9, iters=30):
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


def raster_host(prof, r, mass, R0):
    """Density on the nodes at distances r from the centre with the mass `mass`:
    rho = prof(r - R) with R solved by the secant method (host, numpy)."""
    R, _ = _secant(lambda R: float(_interp(prof, r - R).sum()), mass, R0)
    return _interp(prof, r - R), R


def raster_case(case, prof, centre, mass, R0):
    """Same on a case's device density (mplb_set_radial_density); returns R."""
    R, _ = _secant(lambda R: case.set_radial_density(prof, D0, DD, centre, R), mass, R0)
    case.set_radial_density(prof, D0, DD, centre, R)
    return R


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


def case_profile_densities(tr, a=0.5, b=4.0):
    rv, rl = maxwell(tr * vdw_critical(a, b)[1], a=a, b=b)
    return rv, rl

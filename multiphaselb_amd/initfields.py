"""Initial fields for benchmark and parity cases.

Replaces the reference's setInitialFields generators for the synthetic workloads:
the reference's random field is seeded from the clock (randomScalarField.c:40) and has
no Maxwell construction, so runs could not be repeated.  Here:
  * maxwell()      coexistence densities of an EOS at temperature T from the equal
                   pressure and equal area (in specific volume) conditions, FP64;
  * noise()        uniform +-amp noise from splitmix64 (seed), drawn on the periodic
                   sites and copied onto the duplicate end planes so the two stored
                   copies of a site start identical;
  * sphere()/pool() the reference's tanh profile
                   0.5 (in + out) + 0.5 (out - in) tanh(2 (r - R) / W)
                   (sphereBoundedScalarField.c:184) around a centre or along the
                   vertical axis.
"""
import numpy as np
from scipy import optimize


def p_eos(rho, T, eos="vdw", a=0.5, b=4.0, R=1.0, M=1.0, w=0.344):
    """Same forms as EOS/p_eos.c:7-64."""
    rho = np.asarray(rho, float)
    if eos == "vdw":
        return rho * R * T / (M - rho * b) - a * rho * rho / (M * M)
    if eos == "cs":
        al = b * rho / 4.0
        be = 1 - al
        return rho * R * T * ((1 + al + al * al - al ** 3) / be ** 3) - a * rho * rho
    if eos == "pr":
        tc = 0.170151343 * a / (b * R)
        phi = 1 + (0.37464 + 1.5422 * w - 0.26992 * w * w) * (1 - np.sqrt(T / tc))
        return rho * R * T / (1 - b * rho) - a * phi * phi * rho * rho / (1 + 2 * b * rho - (b * rho) ** 2)
    if eos == "ideal":
        return rho * R * T
    raise ValueError(eos)


def vdw_critical(a=0.5, b=4.0, R=1.0, M=1.0):
    """rho_c = M / (3 b), T_c = 8 a / (27 b R) (with the reference's M scaling)."""
    return M / (3.0 * b), 8.0 * a / (27.0 * b * R * M)


def maxwell(T, eos="vdw", **kw):
    """(rho_vapour, rho_liquid) at temperature T below the critical point."""
    if eos != "vdw":
        return _maxwell_numeric(T, eos, **kw)
    a, b, R, M = kw.get("a", 0.5), kw.get("b", 4.0), kw.get("R", 1.0), kw.get("M", 1.0)
    rc, tc = vdw_critical(a, b, R, M)
    if not 0 < T < tc:
        raise ValueError("T must be below T_c = %g" % tc)
    p = lambda r: p_eos(r, T, "vdw", a=a, b=b, R=R, M=M)
    dp = lambda r: R * T * M / (M - r * b) ** 2 - 2 * a * r / (M * M)
    rmax = (M / b) * (1 - 1e-13)
    lo = optimize.brentq(dp, 1e-12 * rc, rc)      # vapour-side spinodal
    hi = optimize.brentq(dp, rc, rmax)            # liquid-side spinodal

    def roots(ps):
        rv = optimize.brentq(lambda r: p(r) - ps, 0.0, lo, xtol=1e-300, rtol=1e-15)
        rl = optimize.brentq(lambda r: p(r) - ps, hi, rmax, xtol=1e-16, rtol=1e-15)
        return rv, rl

    def prim(r, ps):
        # antiderivative of (p(r) - ps) / r^2 for the van der Waals p
        return R * T / M * np.log(r / (M - b * r)) - a * r / (M * M) + ps / r

    def area(ps):
        # equal areas in specific volume v = 1/rho: integral of (p - ps) dv = 0
        rv, rl = roots(ps)
        return prim(rl, ps) - prim(rv, ps)

    pmin = max(p(hi) * (1 + 1e-12), p(lo) * 1e-10)
    pmax = p(lo) * (1 - 1e-12)
    ps = optimize.brentq(area, pmin, pmax, xtol=1e-18, rtol=1e-15)
    rv, rl = roots(ps)
    return float(rv), float(rl)


def _maxwell_numeric(T, eos, **kw):
    """Equal pressures and equal areas in specific volume for any EOS of p_eos: the
    spinodal densities from a scan of dp/drho, the saturation pressure by bisection on
    the area integral of (p - ps) / rho^2 (scipy quad)."""
    from scipy import integrate

    kw = dict(kw)
    b = kw.get("b", 4.0)
    rmax = {"cs": 4.0 / b, "pr": 1.0 / b}.get(eos, kw.get("M", 1.0) / b) * (1 - 1e-9)
    p = lambda r: float(p_eos(r, T, eos, **kw))
    h = 1e-7 * rmax
    dp = lambda r: (p(r + h) - p(r - h)) / (2 * h)
    grid = np.linspace(rmax * 1e-6, rmax * (1 - 1e-6), 4001)
    d = np.array([dp(r) for r in grid])
    sign = np.nonzero(np.sign(d[:-1]) != np.sign(d[1:]))[0]
    if len(sign) < 2:
        raise ValueError("T is not below the critical temperature of this EOS")
    lo = optimize.brentq(dp, grid[sign[0]], grid[sign[0] + 1])        # vapour-side spinodal
    hi = optimize.brentq(dp, grid[sign[1]], grid[sign[1] + 1])        # liquid-side spinodal

    def roots(ps):
        rv = optimize.brentq(lambda r: p(r) - ps, 1e-300, lo, xtol=1e-300, rtol=1e-15)
        rl = optimize.brentq(lambda r: p(r) - ps, hi, rmax, xtol=1e-16, rtol=1e-15)
        return rv, rl

    def area(ps):
        rv, rl = roots(ps)
        return integrate.quad(lambda r: (p(r) - ps) / (r * r), rv, rl, limit=400, epsabs=0, epsrel=1e-13)[0]

    pmin = max(p(hi) * (1 + 1e-12), p(lo) * 1e-10)
    pmax = p(lo) * (1 - 1e-12)
    ps = optimize.brentq(area, pmin, pmax, xtol=1e-18, rtol=1e-15)
    rv, rl = roots(ps)
    return float(rv), float(rl)


def critical_temperature(eos="vdw", a=0.5, b=4.0, R=1.0, M=1.0, w=0.344):
    """T_c: 8a/(27bR) (van der Waals, with the reference's M scaling), 0.3773 a/(bR)
    (Carnahan-Starling, the value the pseudopotential literature uses), and for
    Peng-Robinson the 0.170151343 a/(bR) of EOS/p_eos.c."""
    if eos == "vdw":
        return vdw_critical(a, b, R, M)[1]
    if eos == "cs":
        return 0.3773 * a / (b * R)
    if eos == "pr":
        return 0.170151343 * a / (b * R)
    raise ValueError(eos)


def splitmix64(seed, n):
    """n uniform doubles in [0, 1) from the splitmix64 sequence started at seed."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def noise(shape, amp, seed, periodic=(True, True, True)):
    """+-amp noise on a (nz, ny, nx) lattice; periodic axes copy plane 0 onto the
    duplicate last plane."""
    nz, ny, nx = shape
    uz = nz - 1 if (periodic[2] and nz > 1) else nz
    uy = ny - 1 if periodic[1] else ny
    ux = nx - 1 if periodic[0] else nx
    core = amp * (2.0 * splitmix64(seed, uz * uy * ux) - 1.0)
    core = core.reshape(uz, uy, ux)
    out = np.empty(shape)
    out[:uz, :uy, :ux] = core
    if ux < nx:
        out[:uz, :uy, ux] = core[:, :, 0]
    if uy < ny:
        out[:uz, uy, :] = out[:uz, 0, :]
    if uz < nz:
        out[uz, :, :] = out[0, :, :]
    return out


def sphere(shape, centre, radius, inside, outside, width):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((i - centre[0]) ** 2 + (j - centre[1]) ** 2 + (k - centre[2]) ** 2)
    return 0.5 * (inside + outside) + 0.5 * (outside - inside) * np.tanh(2.0 * (r - radius) / width)


def pool(shape, height, below, above, width, axis=0):
    """Layer of density `below` under `height` along axis 0 (z) of (nz, ny, nx)."""
    nz, ny, nx = shape
    h = np.arange(shape[axis]).astype(float)
    prof = 0.5 * (below + above) + 0.5 * (above - below) * np.tanh(2.0 * (h - height) / width)
    s = [1, 1, 1]
    s[axis] = shape[axis]
    return np.broadcast_to(prof.reshape(s), shape).copy()


def ensemble(n, seed=7, sizes=(64, 96, 128, 192, 256), tr_range=(0.6, 0.95), a=0.5, b=4.0,
             width=5.0, amp=1e-4, noise_seed=20240601):
    """Parameters of n independent D2Q9 droplet/bubble cases on L x L periodic boxes:
    L uniform over `sizes`, radius uniform in [6, L/4], T/Tc uniform in tr_range, droplet
    or bubble with probability 1/2 each, all from one splitmix64 stream (seed).  Each
    case's noise uses seed noise_seed + case index."""
    u = splitmix64(seed, 4 * n).reshape(n, 4)
    tc = vdw_critical(a, b)[1]
    out = []
    for k in range(n):
        L = int(sizes[min(int(u[k, 0] * len(sizes)), len(sizes) - 1)])
        R = 6.0 + u[k, 1] * (L / 4.0 - 6.0)
        tr = tr_range[0] + u[k, 2] * (tr_range[1] - tr_range[0])
        rv, rl = maxwell(tr * tc, a=a, b=b)
        out.append(dict(index=k, L=L, radius=R, tr=tr, T=tr * tc, droplet=bool(u[k, 3] < 0.5),
                        rho_v=rv, rho_l=rl, width=width, amp=amp, noise_seed=noise_seed + k))
    return out


def ensemble_density(case):
    """Initial density (node-major, i fastest) of one ensemble case: tanh profile
    between the coexistence densities around the box centre, plus noise."""
    L = case["L"]
    inside, outside = (case["rho_l"], case["rho_v"]) if case["droplet"] else (case["rho_v"], case["rho_l"])
    c = (L - 1) / 2.0
    rho = sphere((1, L, L), (c, c, 0.0), case["radius"], inside, outside, case["width"])
    rho = rho + noise((1, L, L), case["amp"], case["noise_seed"], (True, True, False))
    return rho.ravel()

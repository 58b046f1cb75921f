"""Pins of the CPU oracle that need no reference run: closed forms, conservation laws
and fixed points, chosen so that a plausible restatement mistake (a dropped term, a
wrong sign or index, a transposed table) fails one of them."""
import numpy as np
import pytest

import _oracle as O
from _cases import oracle_params
from multiphaselb_amd.casefiles import BC, CaseSpec
from multiphaselb_amd.initfields import maxwell, p_eos, vdw_critical

LAM9 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 0.8, 0.8)
LAM15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)
GLAM9 = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0)
GLAM15 = (0.0,) + (1.0,) * 2 + (1.2, 1.0) * 3 + (1.0,) * 6


@pytest.mark.parametrize("model", [O.D2Q9, O.D3Q15])
def test_lattice_isotropy_and_inverse(model):
    t = O.lattice_tables(model)
    c, w, om = t["vel"].astype(float), t["w"], t["omega"]
    assert abs(om.sum() - 1.0) < 1e-15
    d = t["d"]
    second = np.einsum("k,ki,kj->ij", om, c, c)[:d, :d]
    assert np.allclose(second, t["cs2"] * np.eye(d), atol=1e-15)
    fourth = np.einsum("k,ki,kj,kl,km->ijlm", om, c, c, c, c)[:d, :d, :d, :d]
    I = np.eye(d)
    iso = t["cs2"] ** 2 * (np.einsum("ij,lm->ijlm", I, I) + np.einsum("il,jm->ijlm", I, I) +
                           np.einsum("im,jl->ijlm", I, I))
    if model == O.D2Q9:
        assert np.allclose(fourth, iso, atol=1e-15)
    else:  # D3Q15 is isotropic to fourth order only for the diagonal terms it needs
        assert np.allclose(np.einsum("iijj->ij", fourth), np.einsum("iijj->ij", iso), atol=1e-15)
    assert np.allclose(np.einsum("k,ki,kj->ij", w, c, c)[:d, :d], np.eye(d), atol=1e-15)
    assert np.abs(t["M"] @ t["invM"] - np.eye(t["Q"])).max() < 1e-15
    for k in range(t["Q"]):
        assert np.array_equal(c[t["rev"][k]], -c[k])
    # rows of M are the moments the collision relaxes: density, momentum
    assert np.array_equal(t["M"][0], np.ones(t["Q"]))
    assert np.array_equal(t["M"][3], c[:, 0])


def test_eos_closed_forms():
    p = O.make_params(eos=O.VDW, a=0.5, b=4.0, R=1.0, M=1.0)
    rc, tc = vdw_critical(0.5, 4.0)
    L = O.lib()
    pe = lambda r, T: L.orc_p_eos(p, r, T)
    h = 1e-4 * rc
    d1 = (pe(rc + h, tc) - pe(rc - h, tc)) / (2 * h)
    d2 = (pe(rc + h, tc) - 2 * pe(rc, tc) + pe(rc - h, tc)) / h ** 2
    assert abs(d1) < 1e-7 and abs(d2) < 1e-3           # inflection point at (rho_c, T_c)
    assert pe(rc, tc) == pytest.approx(0.5 / (27 * 16.0), rel=1e-12)   # p_c = a / (27 b^2)
    q = O.make_params(eos=O.IDEAL, R=0.7)
    assert L.orc_p_eos(q, 0.3, 0.2) == 0.3 * 0.7 * 0.2
    cs = O.make_params(eos=O.CARSTAR, a=0.3, b=0.2, R=1.0)
    r = 1e-6
    assert L.orc_p_eos(cs, r, 0.5) == pytest.approx(r * 0.5, rel=1e-6)  # dilute limit
    # the oracle and the host initialiser agree on every EOS
    for eos, name in ((O.VDW, "vdw"), (O.CARSTAR, "cs"), (O.PENGROB, "pr"), (O.IDEAL, "ideal")):
        pp = O.make_params(eos=eos, a=0.3, b=0.5, w_pr=0.3, R=1.0)
        for rho, T in ((0.1, 0.05), (0.6, 0.04)):
            assert L.orc_p_eos(pp, rho, T) == pytest.approx(float(p_eos(rho, T, name, a=0.3, b=0.5, w=0.3)), rel=1e-12)


def test_maxwell_construction_invariants():
    rc, tc = vdw_critical()
    for tr in (0.6, 0.8, 0.9, 0.95):
        T = tr * tc
        rv, rl = maxwell(T)
        ps = float(p_eos(rl, T))
        assert float(p_eos(rv, T)) == pytest.approx(ps, rel=1e-10)
        v = np.linspace(1.0 / rl, 1.0 / rv, 400001)          # equal areas in specific volume
        integrand = p_eos(1.0 / v, T) - ps
        area = np.trapezoid(integrand, v)
        assert abs(area) < 1e-7 * np.trapezoid(np.abs(integrand), v)
    rv, rl = maxwell(0.9 * tc)
    assert (rv / rc, rl / rc) == pytest.approx((0.4257, 1.6573), abs=2e-4)   # tabulated reduced VdW values


def _random_case(model, f_model, nx=6, ny=6, nz=6, seed=0, g=(0.0, 0.0, 0.0)):
    spec = CaseSpec(lbmodel=model, nx=nx, ny=ny, nz=nz if model else 1, btype=3 if model == 0 else 5,
                    f_model=f_model, f_Lambda=LAM9 if model == 0 else LAM15, f_sigma=0.1,
                    g_model=1, g_Lambda=GLAM9 if model == 0 else GLAM15, g=g)
    o = O.Oracle(oracle_params(spec, ht=True))
    rng = np.random.default_rng(seed)
    Q = o.Q
    t = O.lattice_tables(model)
    f = t["omega"] * (0.1 + 0.02 * rng.random((o.N, 1))) * (1 + 0.05 * rng.standard_normal((o.N, Q)))
    o.set("f", f)
    o.set("g", t["omega"] * 0.03 * (1 + 0.05 * rng.standard_normal((o.N, Q))))
    rho = f.sum(1)
    o.set("rho", rho)
    o.set("T", np.full(o.N, 0.03))
    Ft = 1e-4 * rng.standard_normal((o.N, 3))
    if model == 0:
        Ft[:, 2] = 0
    o.set("Ft", Ft)
    o.set("Fi", Ft * 0.9)
    o.set("ss", 1e-5 * rng.standard_normal(o.N))
    return o, t


@pytest.mark.parametrize("model,f_model", [(O.D2Q9, O.LIMRT), (O.D3Q15, O.XUMRT)])
def test_collision_conserves_mass_and_adds_the_force(model, f_model):
    o, t = _random_case(model, f_model, g=(0.0, 0.0, -1e-4) if model else (0.0, -1e-4, 0.0))
    o.call("macro_velocity")                    # U = (sum c f + F/2) / rho
    f0 = o.get("f")
    rho = o.get("rho")
    o.call("collision", 0)
    f1 = o.get("f")
    c = t["vel"].astype(float)
    assert np.allclose(f1.sum(1), f0.sum(1), rtol=0, atol=1e-16)
    if f_model == O.LIMRT:
        F = o.get("Ft")
    else:
        F = o.get("Fi") + (rho - o.rho0)[:, None] * np.array([0.0, 0.0, -1e-4])
    dj = f1 @ c - f0 @ c
    assert np.allclose(dj, F, rtol=0, atol=1e-16)


@pytest.mark.parametrize("model", [O.D2Q9, O.D3Q15])
def test_thermal_collision_adds_the_source(model):
    o, t = _random_case(model, O.LIMRT if model == 0 else O.XUMRT)
    g0 = o.get("g")
    o.call("collision", 1)                       # myMRT, Lambda_0 = 0
    g1 = o.get("g")
    assert np.allclose(g1.sum(1) - g0.sum(1), o.get("ss"), rtol=0, atol=1e-15)


@pytest.mark.parametrize("model,f_model", [(O.D2Q9, O.LIMRT), (O.D3Q15, O.XUMRT)])
def test_rest_equilibrium_is_a_fixed_point(model, f_model):
    spec = CaseSpec(lbmodel=model, nx=7, ny=7, nz=7 if model else 1, btype=0,
                    f_model=f_model, f_Lambda=LAM9 if model == 0 else LAM15, g_model=1,
                    g_Lambda=GLAM9 if model == 0 else GLAM15)
    from multiphaselb_amd.casefiles import BC
    for p in spec.patches():
        spec.f_bc[p] = BC("bounceBack")
        spec.g_bc[p] = BC("bounceBack")
    o = O.Oracle(oracle_params(spec, ht=True))
    o.set("rho", np.full(o.N, 0.12))
    o.set("T", np.full(o.N, 0.03))
    o.set("U", np.zeros((o.N, 3)))
    o.init()
    F = o.get("Fi")
    assert np.all(F == 0)                        # uniform density: no interaction force
    assert np.all(o.get("ss") == 0)              # uniform fields: no heat source
    f0, g0 = o.get("f"), o.get("g")
    assert o.run(3) == 3
    assert np.allclose(o.get("f"), f0, rtol=0, atol=1e-15)
    assert np.allclose(o.get("g"), g0, rtol=0, atol=1e-15)


def test_periodic_stream_is_a_permutation():
    for model, f_model in ((O.D2Q9, O.LIMRT), (O.D3Q15, O.XUMRT)):
        o, t = _random_case(model, f_model, nx=7, ny=5, nz=6)
        nx, ny, nz = 7, 5, (6 if model else 1)
        f0 = o.get("f").reshape(nz, ny, nx, -1)
        # the two stored copies of a periodic site must hold the same values
        f0[:, :, nx - 1] = f0[:, :, 0]
        f0[:, ny - 1] = f0[:, 0]
        if model:
            f0[nz - 1] = f0[0]
        f0 = f0.reshape(-1, o.Q)
        o.set("f", f0)
        o.call("stream", 0)
        f1 = o.get("f")
        # each population moved by +c_k on the period-(n-1) lattice of unique sites
        uniq = lambda a: a.reshape(nz, ny, nx, -1)[: max(nz - 1, 1), : ny - 1, : nx - 1]
        a0, a1 = uniq(f0), uniq(f1)
        for k in range(o.Q):
            cx, cy, cz = t["vel"][k]
            shifted = np.roll(a0[..., k], shift=(cz, cy, cx) if model else (0, cy, cx), axis=(0, 1, 2))
            assert np.array_equal(shifted, a1[..., k])
        assert np.isclose(a1.sum(), a0.sum(), rtol=1e-15)


def test_interaction_force_zero_on_uniform_and_mirror_antisymmetric():
    spec = CaseSpec(lbmodel=0, nx=12, ny=9, btype=3, f_model=0, f_Lambda=LAM9, g_model=None)
    o = O.Oracle(oracle_params(spec, ht=False))
    o.set("T", np.full(o.N, 0.03))
    o.set("rho", np.full(o.N, 0.1))
    o.call("inter_force")
    assert np.all(o.get("Fi") == 0)
    x = np.arange(12) % 11
    prof = 0.1 + 0.04 * np.sin(2 * np.pi * x / 11.0)
    rho = np.tile(prof, 9)
    o.set("rho", rho)
    o.call("inter_force")
    F1 = o.get("Fi").reshape(9, 12, 3)
    mir = np.tile(prof[(11 - x) % 11], 9)
    o.set("rho", mir)
    o.call("inter_force")
    F2 = o.get("Fi").reshape(9, 12, 3)
    xi = np.arange(11)
    assert np.allclose(F2[:, (11 - xi) % 11, 0], -F1[:, xi, 0], rtol=1e-12, atol=1e-20)
    assert np.allclose(F1[..., 1], 0, atol=1e-12 * np.abs(F1).max())
    # smooth field: F ~ -G psi grad psi (pseudopotential weights sum to the identity)
    o.set("rho", rho)
    o.call("inter_force")
    p = O.make_params(eos=O.VDW, a=0.5, b=4.0)
    psi = np.sqrt(2 * (np.array([O.lib().orc_p_eos(p, r, 0.03) for r in prof]) - prof / 3) / -1.0)
    grad = (np.roll(psi, -1) - np.roll(psi, 1)) / 2
    grad[0] = (psi[1] - psi[10]) / 2
    grad[10] = (psi[0] - psi[9]) / 2
    expect = psi[:11] * grad[:11]                 # -G psi grad psi with G = -1
    got = o.get("Fi").reshape(9, 12, 3)[4, :11, 0]
    assert np.allclose(got, expect, rtol=0.05, atol=1e-3 * np.abs(expect).max())


def test_heat_source_divergence_term_closed_form():
    """U = (b x, 0, 0), uniform rho and T: S = b T (1 - dp/dT / (rho Cv)) on interior nodes."""
    spec = CaseSpec(lbmodel=0, nx=9, ny=7, btype=1, f_model=0, f_Lambda=LAM9, g_model=1,
                    g_Lambda=GLAM9)
    o = O.Oracle(oracle_params(spec, ht=True))
    b, rho, T = 1e-4, 0.12, 0.03
    x = np.tile(np.arange(9.0), 7)
    U = np.zeros((o.N, 3))
    U[:, 0] = b * x
    o.set("U", U)
    o.set("rho", np.full(o.N, rho))
    o.set("T", np.full(o.N, T))
    o.call("heat_source")
    p = O.make_params(eos=O.VDW, a=0.5, b=4.0)
    L = O.lib()
    dpdT = (L.orc_p_eos(p, rho, 1.1 * T) - L.orc_p_eos(p, rho, 0.9 * T)) / (0.2 * T)
    expect = b * T * (1.0 - dpdT / rho)
    ss = o.get("ss").reshape(7, 9)
    assert np.allclose(ss[1:-1, 1:-1], expect, rtol=1e-9)


@pytest.mark.parametrize("lbmodel", [0, 1])
def test_heat_source_conduction_term_closed_form(lbmodel):
    """U = 0, T = T0 + a x and rho = r0 + b x: on interior nodes S = -lambda a g / Cv, with
    g = (1/rho(x+1) - 1/rho(x-1)) / 2 and the myMRT conductivity lambda = rho Cv (1/L3 - 1/2)
    (4 + 3 a1 + 2 a2) / 6 (D2Q9) or (6 + 11 a1 + a2) / 9 (D3Q15); the divergence term is
    zero.  Pins the conductivity coefficients and the orientation of both gradients."""
    a1, a2, cv = 0.3, 0.7, 1.7
    if lbmodel == 0:
        spec = CaseSpec(lbmodel=0, nx=9, ny=7, btype=1, f_model=0, f_Lambda=LAM9, g_model=1, g_Lambda=GLAM9,
                        g_alpha1=a1, g_alpha2=a2, Cv=cv)
        shape, lam3, fac = (7, 9), GLAM9[3], (4.0 + 3.0 * a1 + 2.0 * a2) / 6.0
    else:
        spec = CaseSpec(lbmodel=1, nx=9, ny=6, nz=5, btype=0, f_model=2, f_Lambda=LAM15, g_model=1,
                        g_Lambda=GLAM15, g_alpha1=a1, g_alpha2=a2, Cv=cv)
        shape, lam3, fac = (5, 6, 9), GLAM15[3], (6.0 + 11.0 * a1 + a2) / 9.0
    o = O.Oracle(oracle_params(spec, ht=True))
    x = np.indices(shape)[-1].ravel().astype(float)
    T0, a, r0, b = 0.03, 2e-4, 0.1, 3e-3
    rho = r0 + b * x
    o.set("U", np.zeros((o.N, 3)))
    o.set("rho", rho)
    o.set("T", T0 + a * x)
    o.call("heat_source")
    lam = rho * cv * (1.0 / lam3 - 0.5) * fac
    g = 0.5 * (1.0 / (r0 + b * (x + 1)) - 1.0 / (r0 + b * (x - 1)))
    expect = (-lam * a * g / cv).reshape(shape)
    ss = o.get("ss").reshape(shape)
    inner = (slice(1, -1),) * len(shape)
    assert np.allclose(ss[inner], expect[inner], rtol=1e-12, atol=0)
    assert np.all(expect[inner] > 0)


def test_fixed_u_gives_the_wall_equilibrium_next_to_an_equilibrium_neighbour():
    """fixedU rebuilds each missing population as f_eq(rho_nb, U_w) + (f_nb - f_eq(rho_nb,
    u_nb)) (fixedU.c): with uniform density at rest (so the neighbour holds exactly its
    equilibrium and no force acts) the missing populations of the wall nodes become the
    textbook D2Q9 equilibrium w_k rho (1 + 3 c.U + 9/2 (c.U)^2 - 3/2 U^2) of the wall
    velocity, and every other population stays as it was."""
    Uw = (0.01, 0.002, 0.0)
    spec = CaseSpec(lbmodel=0, nx=10, ny=10, btype=0, f_model=0, f_Lambda=LAM9, g_model=None,
                    f_bc={"X0": BC("fixedU", Uw), "X1": BC("bounceBack"), "Y0": BC("bounceBack"),
                          "Y1": BC("bounceBack")})
    o = O.Oracle(oracle_params(spec, ht=False))
    rho0 = 0.13
    o.set("rho", np.full(o.N, rho0))
    o.set("U", np.zeros((o.N, 3)))
    o.set("T", np.full(o.N, 0.03))
    o.init()
    before = o.get("f").reshape(10, 10, 9).copy()
    o.call("update_boundaries", 0)
    after = o.get("f").reshape(10, 10, 9)
    t = O.lattice_tables(O.D2Q9)
    c, w = t["vel"][:, :2], t["omega"]
    cu = c @ np.array(Uw[:2])
    feq = w * rho0 * (1.0 + 3.0 * cu + 4.5 * cu ** 2 - 1.5 * np.dot(Uw, Uw))
    missing = c[:, 0] == 1                          # pulled from x - c, outside at x = 0
    for j in range(1, 9):                           # X0 nodes off the corners
        np.testing.assert_allclose(after[j, 0, missing], feq[missing], rtol=1e-13, atol=0)
        np.testing.assert_array_equal(after[j, 0, ~missing], before[j, 0, ~missing])
    np.testing.assert_array_equal(after[1:-1, 1:-1], before[1:-1, 1:-1])


def test_generic_box_lists_that_leave_the_lattice_are_refused():
    """genericBoundary's (nx-1)*ny row index points past the lattice when nx > ny: the
    oracle refuses the box, as the library does, instead of writing past its arrays."""
    spec = CaseSpec(lbmodel=0, nx=10, ny=8, btype=0, f_model=0, f_Lambda=LAM9, g_model=None)
    o = O.Oracle(oracle_params(spec, ht=False))
    code, msg = o.error()
    assert code != 0 and "leaves the lattice" in msg


def test_pressure_closed_forms():
    """liMRTPressure: p = rho cs2 + G psi^2 / 2 + G psi lap(psi) / 12 + 2 G^2 sigma |grad psi|^2.
    (i) A uniform field gives the equation of state, p = p_eos(rho, T), on every node.
    (ii) A density chosen so that psi = psi0 + bx x + by y is linear gives lap = 0 and
    grad = (bx, by) on every node of the walled box, the one-sided first and second
    differences at edges and corners included (both are exact for a linear psi):
    p = rho / 3 - psi^2 / 2 + 2 sigma (bx^2 + by^2) (G = -1).  psi is inverted from the
    van der Waals closed form, independently of the oracle."""
    from scipy.optimize import brentq
    sigma = 0.125
    spec = CaseSpec(lbmodel=0, nx=12, ny=12, btype=0, f_model=0, f_Lambda=LAM9, f_sigma=sigma, g_model=None,
                    write_interval=1)
    o = O.Oracle(oracle_params(spec, ht=False))
    tc = vdw_critical(0.5, 4.0)[1]
    T = 0.8 * tc
    o.set("T", np.full(o.N, T))
    o.set("U", np.zeros((o.N, 3)))
    o.set("rho", np.full(o.N, 0.1))
    o.call("macro_pressure")
    assert np.allclose(o.get("p"), p_eos(0.1, T, a=0.5, b=4.0), rtol=1e-13, atol=0)
    psi_of = lambda r: np.sqrt(2.0 * (p_eos(r, T, a=0.5, b=4.0) - r / 3.0) / -1.0)   # G = -1
    y, x = (v.ravel().astype(float) for v in np.indices((12, 12)))
    for bx, by in ((4e-3, 0.0), (2e-3, 1.5e-3)):
        psi = psi_of(0.012) + bx * x + by * y
        rho = np.array([brentq(lambda r: psi_of(r) - s, 1e-4, 0.06, xtol=1e-16, rtol=1e-15) for s in psi])
        o.set("rho", rho)
        o.call("macro_pressure")
        expect = rho / 3.0 - 0.5 * psi ** 2 + 2.0 * sigma * (bx ** 2 + by ** 2)
        np.testing.assert_allclose(o.get("p"), expect, rtol=1e-9, atol=0)


def test_pressure_one_sided_laplacian_at_walls():
    """liMRTPressure.c:16-134: where a node lacks its x (or y) neighbour on one side, the
    Laplacian takes the second difference one node inwards, psi(c) - 2 psi(b) + psi(x).
    For psi quadratic in x and y every such difference is the exact second derivative, so
    with sigma = 0 (no gradient term) every node of a walled box, edges and corners
    included, has p = rho / 3 - psi^2 / 2 - psi (2 cx + 2 cy) / 12 (G = -1).  A one-sided
    stencil that reused the node itself or pointed the wrong way would not give 2 c."""
    from scipy.optimize import brentq
    spec = CaseSpec(lbmodel=0, nx=12, ny=12, btype=0, f_model=0, f_Lambda=LAM9, f_sigma=0.0, g_model=None,
                    write_interval=1)
    o = O.Oracle(oracle_params(spec, ht=False))
    tc = vdw_critical(0.5, 4.0)[1]
    T = 0.8 * tc
    o.set("T", np.full(o.N, T))
    o.set("U", np.zeros((o.N, 3)))
    psi_of = lambda r: np.sqrt(2.0 * (p_eos(r, T, a=0.5, b=4.0) - r / 3.0) / -1.0)   # G = -1
    y, x = (v.ravel().astype(float) for v in np.indices((12, 12)))
    cx, cy = 2e-4, 1.5e-4
    psi = psi_of(0.012) + 1e-3 * x + cx * x ** 2 + 5e-4 * y + cy * y ** 2
    rho = np.array([brentq(lambda r: psi_of(r) - s, 1e-4, 0.06, xtol=1e-16, rtol=1e-15) for s in psi])
    o.set("rho", rho)
    o.call("macro_pressure")
    expect = rho / 3.0 - 0.5 * psi ** 2 - psi * (2.0 * cx + 2.0 * cy) / 12.0
    np.testing.assert_allclose(o.get("p"), expect, rtol=1e-9, atol=0)
    # the Laplacian term is resolved: dropping it moves the wall nodes far outside the bound
    assert np.max(np.abs(o.get("p") - (rho / 3.0 - 0.5 * psi ** 2)) / expect) > 1e-6


@pytest.mark.parametrize("force", [0.0, 2e-4])
def test_outflow_is_the_implicit_upwind_convective_condition(force):
    """outflow.c is the convective condition df/dt + U df/dn = 0 taken implicit-upwind on
    the outlet face: each population entering from outside (pulled from x - c with
    c_x = -1 at x = nx - 1) becomes (a + U b) / (1 + U), a the node's own value, b the
    inner neighbour's and U that neighbour's outward velocity (j + F / 2) / rho; the other
    populations keep their values.  The neighbours hold the textbook D2Q9 equilibrium of a
    known velocity, so U is known in closed form."""
    n = 10
    spec = CaseSpec(lbmodel=0, nx=n, ny=n, btype=0, f_model=0, f_Lambda=LAM9, g_model=None,
                    f_bc={"X0": BC("bounceBack"), "X1": BC("outflow"), "Y0": BC("bounceBack"),
                          "Y1": BC("bounceBack")})
    o = O.Oracle(oracle_params(spec, ht=False))
    rho0, u = 0.13, 0.04
    o.set("rho", np.full(o.N, rho0))
    o.set("U", np.zeros((o.N, 3)))
    o.set("T", np.full(o.N, 0.03))
    o.init()
    t = O.lattice_tables(O.D2Q9)
    c, w = t["vel"][:, :2], t["omega"]
    cu = c[:, 0] * u
    feq = w * rho0 * (1.0 + 3.0 * cu + 4.5 * cu ** 2 - 1.5 * u * u)
    f = np.tile(feq, (n, n, 1))
    rng = np.random.default_rng(11)
    f[:, n - 1, :] = feq * (1.0 + 0.3 * rng.random((n, 9)))      # outlet nodes: arbitrary values
    o.set("f", f.reshape(-1, 9))
    Fi = np.zeros((n, n, 3))
    Fi[:, n - 2, 0] = force                                        # on the inner neighbours
    o.set("Fi", Fi.reshape(-1, 3))
    before = o.get("f").reshape(n, n, 9).copy()
    o.call("update_boundaries", 0)
    after = o.get("f").reshape(n, n, 9)
    U = u + 0.5 * force / rho0
    entering = c[:, 0] == -1
    for j in range(1, n - 1):                                      # outlet nodes off the corners
        a, b = before[j, n - 1], before[j, n - 2]
        np.testing.assert_allclose(after[j, n - 1, entering], ((a + U * b) / (1.0 + U))[entering], rtol=1e-13, atol=0)
        np.testing.assert_array_equal(after[j, n - 1, ~entering], a[~entering])
    np.testing.assert_array_equal(after[1:-1, 1:-2], before[1:-1, 1:-2])


@pytest.mark.parametrize("lbmodel,lam3,a1,a2", [(0, 1.2, 0.25, 0.25), (0, 1.0, -1.0, 1.0),
                                                 (1, 1.2, 0.25, 0.25), (1, 1.0, -0.5, 0.5)])
def test_mymrt_temperature_diffuses_at_the_heat_source_conductivity(lbmodel, lam3, a1, a2):
    """myMRT (myMRTCollision.c:35-212, with its off-diagonal relaxation entries) is a
    diffusion scheme: with the flow frozen, uniform density and zero velocity the heat
    source vanishes and a sine temperature profile of wavenumber k decays as
    exp(-kappa k^2 t), where kappa = (1/Lambda3 - 1/2)(4 + 3 a1 + 2 a2)/6 (D2Q9) or
    (1/Lambda3 - 1/2)(6 + 11 a1 + a2)/9 (D3Q15), the conductivity the reference's
    heatSource.c assumes.  The measured rate agrees to lattice discretisation error
    (< 4e-3 here); dropping the off-diagonal entry gives 0.81 of it, flipping its sign
    0.62."""
    if lbmodel == 0:
        n, reps = 128, 4
        gl = (0.0, 1.0, 1.0, lam3, 1.0, lam3, 1.0, 1.0, 1.0)
        spec = CaseSpec(lbmodel=0, nx=n, ny=4, btype=3, f_model=0, g_model=1, g_Lambda=gl, g_alpha1=a1,
                        g_alpha2=a2, write_interval=10 ** 6)
        kappa = (1.0 / lam3 - 0.5) * (4.0 + 3.0 * a1 + 2.0 * a2) / 6.0
    else:
        n, reps = 96, 16
        gl = (0.0, 1.0, 1.0, lam3, 1.0, lam3, 1.0, lam3, 1.0) + (1.0,) * 6
        lam15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)
        spec = CaseSpec(lbmodel=1, nx=n, ny=4, nz=4, btype=5, f_model=2, f_Lambda=lam15, g_model=1, g_Lambda=gl,
                        g_alpha1=a1, g_alpha2=a2, write_interval=10 ** 6)
        kappa = (1.0 / lam3 - 0.5) * (6.0 + 11.0 * a1 + a2) / 9.0
    o = O.Oracle(oracle_params(spec, ht=True, frozen=True))
    P = n - 1                                   # duplicate-node period
    x = np.tile(np.arange(n), reps).astype(float)
    T0, eps = 0.03, 1e-4
    o.set("rho", np.full(o.N, 0.1))
    o.set("U", np.zeros((o.N, 3)))
    o.set("T", T0 + eps * np.sin(2.0 * np.pi * x / P))
    o.init()
    ts, amp = [], []
    for k in range(10):
        assert o.run(200) == 200
        T = o.get("T")[:P]
        ts.append(200.0 * (k + 1))
        amp.append(2.0 / P * np.sum((T - T0) * np.sin(2.0 * np.pi * np.arange(P) / P)))
    rate = -np.polyfit(ts[1:], np.log(amp[1:]), 1)[0]
    assert np.all(o.get("U") == 0.0)
    assert abs(rate / (kappa * (2.0 * np.pi / P) ** 2) - 1.0) < 4e-3


@pytest.mark.parametrize("lam3,lam4,u", [(1.2, 1.0, 0.05), (1.0, 1.5, 0.05), (1.2, 1.0, 0.0)])
def test_litempmrt_advects_temperature_at_the_corrected_speed(lam3, lam4, u):
    """liTempMRTCollision.c:75-77 adds (1 - Lambda3/2) Lambda4 n4 to n3 using the whole
    moment n4 (the form with the non-equilibrium part n4 - n4_eq is commented out at :71-73).
    For a uniform field the moment fixed point is then n4 = -T u, n3 = T u (1 - (1 -
    Lambda3/2) Lambda4 / Lambda3): in a frozen uniform flow a sine temperature profile
    travels at that fraction of u (0.667 u for Lambda = (1.2, 1.0), 0.25 u for (1.0, 1.5);
    the commented-out form would give u) and decays at the D2Q9 advection-diffusion rate
    cs2 (1/Lambda3 - 1/2) k^2 (cs2 = 1/3, from e_eq = -2T)."""
    n = 128
    gl = (0.0, 1.0, 1.0, lam3, lam4, lam3, lam4, 1.0, 1.0)
    spec = CaseSpec(lbmodel=0, nx=n, ny=4, btype=3, f_model=0, g_model=3, g_Lambda=gl, write_interval=10 ** 6)
    o = O.Oracle(oracle_params(spec, ht=True, frozen=True))
    P = n - 1
    k = 2.0 * np.pi / P
    x = np.tile(np.arange(n), 4).astype(float)
    U = np.zeros((o.N, 3))
    U[:, 0] = u
    T0 = 0.03
    o.set("rho", np.full(o.N, 0.1))
    o.set("U", U)
    o.set("T", T0 + 1e-4 * np.sin(k * x))
    o.init()
    xs = np.arange(P)
    ts, amp, phase = [], [], []
    for it in range(10):
        assert o.run(200) == 200
        T = o.get("T")[:P] - T0
        a, b = np.sum(T * np.sin(k * xs)), np.sum(T * np.cos(k * xs))
        ts.append(200.0 * (it + 1))
        amp.append(2.0 / P * np.hypot(a, b))
        phase.append(np.arctan2(b, a))               # T ~ sin(k (x - c t)): phase = -k c t
    rate = -np.polyfit(ts[1:], np.log(amp[1:]), 1)[0]
    np.testing.assert_allclose(rate, (1.0 / lam3 - 0.5) / 3.0 * k * k, rtol=1e-2)
    speed = -np.polyfit(ts[1:], np.unwrap(phase[1:]), 1)[0] / k
    expect = u * (1.0 - (1.0 - lam3 / 2.0) * lam4 / lam3)
    assert abs(speed - expect) <= 5e-3 * u + 1e-9, (speed, expect)


def test_fixed_u_is_the_non_equilibrium_extrapolation_with_a_forced_neighbour():
    """fixedU.c (liMRT) is the non-equilibrium extrapolation of Guo, Zheng and Shi (2002):
    each population k missing at a wall node x becomes f_eq_k(rho_n, U_w) + (f_k(n) -
    f_eq_k(rho_n, u_n)), n = x + c_k the inner node it would have come from reversed,
    rho_n and u_n = (sum c f(n) + F(n) / 2) / rho_n from that node's current populations
    and force.  Here the inner nodes hold arbitrary non-equilibrium populations and a
    force, and the expectation uses the textbook D2Q9 equilibrium."""
    n = 10
    Uw = np.array([0.02, -0.005, 0.0])
    spec = CaseSpec(lbmodel=0, nx=n, ny=n, btype=0, f_model=0, f_Lambda=LAM9, g_model=None,
                    f_bc={"X0": BC("fixedU", tuple(Uw)), "X1": BC("bounceBack"), "Y0": BC("bounceBack"),
                          "Y1": BC("bounceBack")})
    o = O.Oracle(oracle_params(spec, ht=False))
    o.set("rho", np.full(o.N, 0.13))
    o.set("U", np.zeros((o.N, 3)))
    o.set("T", np.full(o.N, 0.03))
    o.init()
    t = O.lattice_tables(O.D2Q9)
    c, w = t["vel"][:, :2].astype(float), t["omega"]

    def feq(rho, u):
        cu = c @ u[:2]
        return w * rho * (1.0 + 3.0 * cu + 4.5 * cu ** 2 - 1.5 * np.dot(u[:2], u[:2]))

    rng = np.random.default_rng(5)
    f = o.get("f").reshape(n, n, 9) * (1.0 + 0.2 * rng.random((n, n, 9)))
    Fi = np.zeros((n, n, 3))
    Fi[:, :, :2] = 1e-3 * rng.standard_normal((n, n, 2))
    o.set("f", f.reshape(-1, 9))
    o.set("Fi", Fi.reshape(-1, 3))
    before = o.get("f").reshape(n, n, 9).copy()
    o.call("update_boundaries", 0)
    after = o.get("f").reshape(n, n, 9)
    for j in range(2, n - 2):                   # X0 nodes whose inner neighbours are interior
        for k in np.nonzero(c[:, 0] == 1)[0]:     # pulled from x - c: outside at x = 0
            jn, inn = j + int(c[k, 1]), 1
            fn = before[jn, inn]
            rho_n = fn.sum()
            u_n = (c.T @ fn + 0.5 * Fi[jn, inn, :2]) / rho_n
            want = feq(rho_n, Uw)[k] + fn[k] - feq(rho_n, u_n)[k]
            np.testing.assert_allclose(after[j, 0, k], want, rtol=1e-13, atol=0)
        others = c[:, 0] != 1
        np.testing.assert_array_equal(after[j, 0, others], before[j, 0, others])


def test_outflow_leaves_outlet_corners_unchanged():
    """At an outlet corner outflow.c's normal is the float sum of two axis directions,
    (1, 1)/sqrt(2) (or (1, -1)/sqrt(2)); converting its components to int truncates them
    to 0, so the matching lattice direction is the rest one and the node is its own
    neighbour: (a + U a) / (1 + U) = a leaves every population as it was, up to the
    rounding of that expression.  (A normal rounded to the diagonal would mix in the
    diagonal neighbour, a change of order 10% here.)  The Y patches carry no condition
    here, so only outflow touches the corners."""
    n = 10
    spec = CaseSpec(lbmodel=0, nx=n, ny=n, btype=0, f_model=0, f_Lambda=LAM9, g_model=None,
                    f_bc={"X0": BC("bounceBack"), "X1": BC("outflow"), "Y0": BC("periodic"),
                          "Y1": BC("periodic")})
    o = O.Oracle(oracle_params(spec, ht=False))
    o.set("rho", np.full(o.N, 0.13))
    U = np.zeros((o.N, 3))
    U[:, 0] = 0.03
    o.set("U", U)
    o.set("T", np.full(o.N, 0.03))
    o.init()
    rng = np.random.default_rng(9)
    f = o.get("f").reshape(n, n, 9) * (1.0 + 0.3 * rng.random((n, n, 9)))
    o.set("f", f.reshape(-1, 9))
    before = o.get("f").reshape(n, n, 9).copy()
    o.call("update_boundaries", 0)
    after = o.get("f").reshape(n, n, 9)
    for j in (0, n - 1):
        np.testing.assert_allclose(after[j, n - 1], before[j, n - 1], rtol=1e-15, atol=0)
    assert not np.array_equal(after[1:-1, n - 1], before[1:-1, n - 1])    # the face itself is updated


def test_heat_source_makes_the_steady_flux_uniform_across_a_density_profile():
    """The energy equation behind heatSource.c is rho Cv dT/dt = div(lambda grad T) with
    lambda = rho Cv kappa: myMRT diffuses at kappa and the source adds -lambda grad T .
    grad(1/rho) / Cv = kappa grad T . grad(rho) / rho.  With the flow frozen, a static
    density profile rho(y) and fixed temperatures on two walls, the steady state therefore
    carries a uniform flux rho dT/dy (while rho varies threefold).  Measured to within 2%
    (lattice discretisation across the interface); the source with its sign flipped in the
    oracle gives a 170% spread."""
    nx, ny = 4, 48
    Th, Tc = 0.032, 0.028
    spec = CaseSpec(lbmodel=0, nx=nx, ny=ny, btype=1, f_model=0, g_model=1, write_interval=10 ** 6,
                    f_bc={"Y0": BC("bounceBack"), "Y1": BC("bounceBack")},
                    g_bc={"Y0": BC("fixedT", (Th,)), "Y1": BC("fixedT", (Tc,))})
    o = O.Oracle(oracle_params(spec, ht=True, frozen=True))
    y = np.repeat(np.arange(ny), nx).astype(float)
    rho = 0.10 + 0.05 * np.tanh((y - 24.0) / 6.0)
    o.set("rho", rho)
    o.set("U", np.zeros((o.N, 3)))
    o.set("T", Th + (Tc - Th) * y / (ny - 1))
    o.init()
    assert o.run(20000) == 20000
    T = o.get("T").reshape(ny, nx)
    assert np.all(T == T[:, :1])                      # uniform along the periodic x
    r = rho.reshape(ny, nx)[:, 0]
    q = (r[1:-1] * (T[2:, 0] - T[:-2, 0]) / 2.0)[3:-3]   # central flux away from the walls
    assert q.mean() < 0                                # heat flows from the hot wall
    assert np.ptp(q) / abs(q.mean()) < 2e-2

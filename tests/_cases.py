"""Shared case builders for the parity tests: one CaseSpec drives both the library
(through a case directory in the reference's format) and the CPU oracle."""
import tempfile

import numpy as np

import _oracle as O
from multiphaselb_amd.casefiles import BC, CaseSpec, write_case

BC_IDS = {"bounceBack": O.BBACK, "fixedT": O.FT, "fixedRandomT": O.FRDT, "fixedGradT": O.FGT,
          "fixedU": O.FU, "outflow": O.OF, "periodic": O.PERIODIC}


def parsed_eos_quirks(spec):
    """Values the reference's first-substring dictionary lookups produce for keys that
    also occur inside the EOS model name (dictIO/singleEntry.c:14)."""
    R, g = spec.R, tuple(spec.g)
    if spec.eos_keys_first:         # the keys come before the model name
        return R, g
    if spec.eos == 1:    # "Carnahan-Starling": EOS/g matches the 'g' of "Starling"
        g = (0.0, spec.a, 0.0)
    if spec.eos == 2:    # "Peng-Robinson": EOS/R matches "Robinson", EOS/g matches "Peng"
        R, g = 0.0, (0.0, 0.0, 0.0)
    return R, g


def oracle_params(spec: CaseSpec, ht=True, frozen=False, nthreads=0):
    R, g = parsed_eos_quirks(spec)
    p = O.make_params(
        lbmodel=spec.lbmodel, nx=spec.nx, ny=spec.ny, nz=spec.nz, btype=spec.btype,
        spot_radius=spec.spot_radius, spot_centre=spec.spot_centre,
        eos=spec.eos, a=spec.a, b=spec.b, w_pr=spec.w_pr, R=R, G=spec.G, M=spec.M, Cv=spec.Cv,
        rho0=spec.rho0, g=g, fext=spec.fext, f_model=spec.f_model, f_Lambda=spec.f_Lambda,
        f_sigma=spec.f_sigma, f_st=spec.f_st, f_kappa=spec.f_kappa,
        g_model=spec.g_model if spec.g_model is not None else O.MYMRT,
        g_Lambda=spec.g_Lambda, g_alpha1=spec.g_alpha1, g_alpha2=spec.g_alpha2,
        ht=1 if (ht and spec.g_model is not None) else 0, frozen=0 if frozen else 1,
        start=spec.start, end=spec.end, write_interval=spec.write_interval, nthreads=nthreads)
    for b, name in enumerate(spec.patches()):
        for fld, bcs in (("f", spec.f_bc), ("g", spec.g_bc)):
            bc = bcs.get(name, BC())
            getattr(p, fld + "_bc")[b] = BC_IDS[bc.type]
            val = getattr(p, fld + "_bcval")[b]
            if bc.value is not None:
                for j, x in enumerate(bc.value):
                    val[j] = x
            if bc.type == "fixedRandomT":
                if bc.pert is None:
                    getattr(p, fld + "_pert")[b] = 0.0
                    getattr(p, fld + "_psteps")[b] = 0
                    getattr(p, fld + "_localT0")[b] = 1
                else:
                    getattr(p, fld + "_pert")[b] = bc.pert
                    getattr(p, fld + "_psteps")[b] = int(bc.duration)
                    getattr(p, fld + "_localT0")[b] = 0
    return p


def splitmix64_uniform(seed, n):
    """Uniform doubles in [0, 1) from splitmix64 (seeded, platform independent): element
    i mixes seed + (i + 1) * 0x9E3779B97F4A7C15, as the sequential generator would."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def coords(spec):
    nz = spec.nz if spec.lbmodel == 1 else 1
    k, j, i = np.meshgrid(np.arange(nz), np.arange(spec.ny), np.arange(spec.nx), indexing="ij")
    return i.ravel().astype(float), j.ravel().astype(float), k.ravel().astype(float)


def sphere_fields(spec, T, r_in, r_out, radius, width=5.0, centre=None, noise=1e-4, seed=1, column=None):
    """tanh profile between two densities (sphereBoundedScalarField.c:184 form),
    plus seeded noise; returns rho, U, T arrays of the stored lattice."""
    i, j, k = coords(spec)
    if centre is None:
        centre = ((spec.nx - 1) / 2.0, (spec.ny - 1) / 2.0, (spec.nz - 1) / 2.0 if spec.lbmodel == 1 else 0.0)
    if column is not None:       # liquid pool below height column (z in 3D, y in 2D)
        h = k if spec.lbmodel == 1 else j
        rho = 0.5 * (r_in + r_out) + 0.5 * (r_out - r_in) * np.tanh(2.0 * (h - column) / width)
    else:
        r = np.sqrt((i - centre[0]) ** 2 + (j - centre[1]) ** 2 + (k - centre[2]) ** 2)
        rho = 0.5 * (r_in + r_out) + 0.5 * (r_out - r_in) * np.tanh(2.0 * (r - radius) / width)
    if noise:
        rho = rho + noise * (2.0 * splitmix64_uniform(seed, rho.size) - 1.0)
    N = rho.size
    return rho, np.zeros((N, 3)), np.full(N, float(T))


def make_case_dir(spec, root=None):
    d = tempfile.mkdtemp(prefix="mplbcase_", dir=root)
    write_case(d, spec)
    return d


def pair(spec, fields, precision=0, ht=True, frozen=False, fused=False, extensions=0):
    """(library Case, Oracle) set up from the same spec and initial fields, both initialised."""
    import multiphaselb_amd as M
    d = make_case_dir(spec)
    c = M.Case(d, precision=precision, heat_transfer=ht, frozen=frozen, fused=fused, extensions=extensions)
    o = O.Oracle(oracle_params(spec, ht=ht, frozen=frozen))
    rho, U, T = fields
    for name, v in (("rho", rho), ("U", U), ("T", T)):
        c.set(name, v)
        o.set(name, v)
    c.init()
    o.init()
    return c, o


def max_rel(a, b):
    """Largest |a - b| relative to |b|, with entries below 1e-8 max|b| measured against that floor."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if a.size == 0:
        return 0.0
    floor = 1e-8 * float(np.max(np.abs(b))) + 1e-300
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))

"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  Builds the oracle on first use if the shared object is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# MPLB_ORACLE_LIB points the tests at another build of the oracle (e.g. one with a
# deliberately planted mistake, to check that the pins catch it)
LIB = os.environ.get("MPLB_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")
REF_TABLES = os.path.join(ORACLE_DIR, "_ref", "libreftables.so")

D2Q9, D3Q15 = 0, 1
LIMRT, MYMRT, XUMRT, LITEMPMRT = 0, 1, 2, 3
IDEAL, CARSTAR, PENGROB, VDW = 0, 1, 2, 3
BBACK, FT, FRDT, FGT, FU, UHF, FEP, OF, PERIODIC = range(9)
GENERIC, PERIODIC_X, PERIODIC_Y, PERIODIC_XY, PERIODIC_Z, PERIODIC_XYZ, PERIODIC_XY_SPOT = range(7)
MAXBD = 6


class OrcParams(C.Structure):
    _fields_ = [
        ("lbmodel", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("btype", C.c_int), ("spot_radius", C.c_double), ("spot_centre", C.c_double * 2),
        ("eos", C.c_int), ("a", C.c_double), ("b", C.c_double), ("w_pr", C.c_double),
        ("R", C.c_double), ("G", C.c_double), ("M", C.c_double), ("Cv", C.c_double),
        ("rho0", C.c_double), ("g", C.c_double * 3), ("fext", C.c_double * 3),
        ("f_model", C.c_int), ("f_Lambda", C.c_double * 15), ("f_sigma", C.c_double),
        ("f_kappa", C.c_double), ("f_st", C.c_int),
        ("g_model", C.c_int), ("g_Lambda", C.c_double * 15), ("g_alpha1", C.c_double),
        ("g_alpha2", C.c_double),
        ("ht", C.c_int), ("frozen", C.c_int),
        ("f_bc", C.c_int * MAXBD), ("g_bc", C.c_int * MAXBD),
        ("f_bcval", (C.c_double * 3) * MAXBD), ("g_bcval", (C.c_double * 3) * MAXBD),
        ("g_pert", C.c_double * MAXBD), ("g_psteps", C.c_int * MAXBD),
        ("g_localT0", C.c_int * MAXBD),
        ("f_pert", C.c_double * MAXBD), ("f_psteps", C.c_int * MAXBD),
        ("f_localT0", C.c_int * MAXBD),
        ("start", C.c_int), ("end", C.c_int), ("write_interval", C.c_int),
        ("nthreads", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])
    L = C.CDLL(LIB)
    P = C.POINTER(OrcParams)
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [P]
    for name in ("orc_destroy", "orc_init", "orc_macro_density", "orc_macro_velocity",
                 "orc_macro_temperature", "orc_inter_force", "orc_heat_source",
                 "orc_macro_pressure"):
        getattr(L, name).argtypes = [C.c_void_p]
        getattr(L, name).restype = None
    for name in ("orc_collision", "orc_stream", "orc_update_boundaries", "orc_equilibrium"):
        getattr(L, name).argtypes = [C.c_void_p, C.c_int]
        getattr(L, name).restype = None
    L.orc_error.argtypes = [C.c_void_p]
    L.orc_error_msg.argtypes = [C.c_void_p]
    L.orc_error_msg.restype = C.c_char_p
    L.orc_npoints.argtypes = [C.c_void_p]
    L.orc_npoints.restype = C.c_int64
    L.orc_Q.argtypes = [C.c_void_p]
    L.orc_nbd.argtypes = [C.c_void_p]
    L.orc_bd_size.argtypes = [C.c_void_p, C.c_int]
    L.orc_get_nb.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_get_bd.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.orc_set_field.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    L.orc_get_field.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    L.orc_get_rho0.argtypes = [C.c_void_p]
    L.orc_get_rho0.restype = C.c_double
    L.orc_set_rho0.argtypes = [C.c_void_p, C.c_double]
    L.orc_current_time.argtypes = [C.c_void_p]
    L.orc_run.argtypes = [C.c_void_p, C.c_int]
    L.orc_average_rho.argtypes = [C.c_void_p]
    L.orc_average_rho.restype = C.c_double
    L.orc_lattice_tables.argtypes = [C.c_int] + [C.c_void_p] * 9
    L.orc_p_eos.argtypes = [P, C.c_double, C.c_double]
    L.orc_p_eos.restype = C.c_double
    _lib = L
    return L


def lattice_tables(model):
    Q = 9 if model == D2Q9 else 15
    q, d, cs2 = C.c_int(), C.c_int(), C.c_double()
    vel = np.zeros((Q, 3), np.int32)
    rev = np.zeros(Q, np.int32)
    omega, w = np.zeros(Q), np.zeros(Q)
    M, invM = np.zeros((Q, Q)), np.zeros((Q, Q))
    lib().orc_lattice_tables(model, C.byref(q), C.byref(d), C.byref(cs2), vel.ctypes.data,
                             rev.ctypes.data, omega.ctypes.data, w.ctypes.data, M.ctypes.data,
                             invM.ctypes.data)
    return dict(Q=q.value, d=d.value, cs2=cs2.value, vel=vel, rev=rev, omega=omega, w=w, M=M,
                invM=invM)


def make_params(**kw):
    """Defaults follow the example dictionaries (vdW, G = -1, R = M = Cv = 1)."""
    p = OrcParams()
    p.lbmodel = kw.get("lbmodel", D2Q9)
    p.nx, p.ny = kw.get("nx", 16), kw.get("ny", 16)
    p.nz = kw.get("nz", 1)
    p.btype = kw.get("btype", PERIODIC_XY)
    p.spot_radius = kw.get("spot_radius", 0.0)
    sc = kw.get("spot_centre", (0.0, 0.0))
    p.spot_centre[0], p.spot_centre[1] = sc
    p.eos = kw.get("eos", VDW)
    p.a, p.b = kw.get("a", 9.0 / 49.0), kw.get("b", 2.0 / 21.0)
    p.w_pr = kw.get("w_pr", 0.344)
    p.R, p.G, p.M, p.Cv = kw.get("R", 1.0), kw.get("G", -1.0), kw.get("M", 1.0), kw.get("Cv", 1.0)
    p.rho0 = kw.get("rho0", 0.0)
    for j, v in enumerate(kw.get("g", (0.0, 0.0, 0.0))):
        p.g[j] = v
    for j, v in enumerate(kw.get("fext", (0.0, 0.0, 0.0))):
        p.fext[j] = v
    p.f_model = kw.get("f_model", LIMRT if p.lbmodel == D2Q9 else XUMRT)
    for k, v in enumerate(kw.get("f_Lambda", ())):
        p.f_Lambda[k] = v
    p.f_sigma = kw.get("f_sigma", 0.0)
    p.f_kappa = kw.get("f_kappa", 0.0)
    p.f_st = kw.get("f_st", 0)
    p.g_model = kw.get("g_model", MYMRT)
    for k, v in enumerate(kw.get("g_Lambda", ())):
        p.g_Lambda[k] = v
    p.g_alpha1, p.g_alpha2 = kw.get("g_alpha1", 0.25), kw.get("g_alpha2", 0.25)
    p.ht = kw.get("ht", 0)
    p.frozen = kw.get("frozen", 1)
    for b, t in enumerate(kw.get("f_bc", ())):
        p.f_bc[b] = t
    for b, t in enumerate(kw.get("g_bc", ())):
        p.g_bc[b] = t
    for b, v in enumerate(kw.get("f_bcval", ())):
        for j, x in enumerate(v):
            p.f_bcval[b][j] = x
    for b, v in enumerate(kw.get("g_bcval", ())):
        for j, x in enumerate(v):
            p.g_bcval[b][j] = x
    for b, v in enumerate(kw.get("g_pert", ())):
        p.g_pert[b] = v
    for b, v in enumerate(kw.get("g_psteps", ())):
        p.g_psteps[b] = v
    for b, v in enumerate(kw.get("g_localT0", ())):
        p.g_localT0[b] = v
    p.start, p.end = kw.get("start", 0), kw.get("end", 10 ** 9)
    p.write_interval = kw.get("write_interval", 10 ** 9)
    p.nthreads = kw.get("nthreads", 0)
    return p


class Oracle:
    def __init__(self, params):
        self.p = params
        self.L = lib()
        self.h = self.L.orc_create(C.byref(params))
        self.N = self.L.orc_npoints(self.h)
        self.Q = self.L.orc_Q(self.h)

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def _size(self, name):
        if name in ("rho", "T", "p", "ss"):
            return (self.N,)
        if name in ("U", "Fi", "Ft"):
            return (self.N, 3)
        return (self.N, self.Q)

    def set(self, name, v):
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(self._size(name))
        assert self.L.orc_set_field(self.h, name.encode(), a.ctypes.data) == 0
        self._keep = a

    def get(self, name):
        a = np.zeros(self._size(name))
        assert self.L.orc_get_field(self.h, name.encode(), a.ctypes.data) == 0
        return a

    def nb(self):
        a = np.zeros((self.N, self.Q), np.int32)
        self.L.orc_get_nb(self.h, a.ctypes.data)
        return a

    def boundaries(self):
        out = []
        for b in range(self.L.orc_nbd(self.h)):
            n = self.L.orc_bd_size(self.h, b)
            a = np.zeros(n, np.int32)
            self.L.orc_get_bd(self.h, b, a.ctypes.data)
            out.append(a)
        return out

    def error(self):
        return self.L.orc_error(self.h), self.L.orc_error_msg(self.h).decode()

    @property
    def rho0(self):
        return self.L.orc_get_rho0(self.h)

    def init(self):
        self.L.orc_init(self.h)

    def run(self, n):
        return self.L.orc_run(self.h, n)

    def call(self, name, *args):
        getattr(self.L, "orc_" + name)(self.h, *args)

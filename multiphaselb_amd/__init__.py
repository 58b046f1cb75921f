"""multiphaselb_amd -- MI355X lattice update for multiphaseLB cases.

Thin ctypes layer over libmplb.so (include/mplb.h).  All compute runs in the HIP
library; this module only moves arrays and writes case dictionaries.  Importing it
does not need a GPU; opening a case does, and fails loudly when the library is not
built or no GPU is usable.
"""
import ctypes as C
import os

import numpy as np

from .casefiles import CaseSpec, write_case  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
# MPLB_LIBRARY may name another build of the same library (e.g. one compiled with
# different flags for a comparison); by default the in-tree build is loaded
LIB_PATH = os.environ.get("MPLB_LIBRARY") or os.path.join(HERE, "libmplb.so")

# FP16S: FP16 deviations from one reference density per case; FP16N: FP16 relative
# deviations from each node's own density (include/mplb.h)
FP64, FP32, FP16S, FP16N = 0, 1, 2, 3
EXT_LIMRT_D3Q15 = 0x1
EXT_NO_FIXEDGRADT_ERROR = 0x2
MODEL_NAMES = {0: "liMRT", 1: "myMRT", 2: "xuMRT", 3: "liTempMRT", -1: None}
NTIMERS = 8

ERRORS = {-1: "argument", -2: "io", -3: "model", -4: "hip", -5: "nan", -6: "state"}


class MplbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mplb error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code
        self.msg = msg


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("heat_transfer", C.c_int), ("frozen", C.c_int),
                ("fused", C.c_int), ("extensions", C.c_int), ("device", C.c_int)]


class Info(C.Structure):
    _fields_ = [("lbmodel", C.c_int), ("Q", C.c_int), ("d", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("nz", C.c_int), ("npoints", C.c_int64), ("nbd", C.c_int),
                ("f_model", C.c_int), ("g_model", C.c_int), ("start_time", C.c_int),
                ("end_time", C.c_int), ("write_interval", C.c_int), ("current_time", C.c_int),
                ("precision", C.c_int), ("fused_active", C.c_int), ("rho0", C.c_double)]


_lib = None

EXPORTS = [
    "mplb_default_opts", "mplb_case_open", "mplb_case_close", "mplb_last_error", "mplb_case_info",
    "mplb_case_param",
    "mplb_set_field", "mplb_get_field", "mplb_get_neighbours", "mplb_get_boundary", "mplb_init",
    "mplb_step", "mplb_collision", "mplb_stream", "mplb_update_boundaries", "mplb_macro_density",
    "mplb_macro_velocity", "mplb_macro_temperature", "mplb_inter_force", "mplb_heat_source",
    "mplb_equilibrium", "mplb_macro_pressure", "mplb_average_density", "mplb_time_steps",
    "mplb_timer_name", "mplb_lattice_tables", "mplb_glibc_rand",
    "mplb_batch_create", "mplb_batch_close", "mplb_batch_last_error", "mplb_batch_set_case",
    "mplb_batch_init", "mplb_batch_step", "mplb_batch_run_to_steady", "mplb_batch_get_field",
    "mplb_batch_case_steps", "mplb_batch_tune", "mplb_batch_set_launch", "mplb_batch_launch_info",
    "mplb_batch_set_march", "mplb_batch_march_rows", "mplb_batch_set_persistent",
    "mplb_density_change", "mplb_set_radial_density", "mplb_set_radial_density_noalloc",
    "mplb_write_lattice", "mplb_write_fields", "mplb_read_field",
    "mplb_mlp_create", "mplb_mlp_close", "mplb_mlp_last_error", "mplb_mlp_num_params", "mplb_mlp_set_params",
    "mplb_mlp_get_params", "mplb_mlp_forward", "mplb_mlp_gradient", "mplb_mlp_train", "mplb_mlp_tensor",
]


def library():
    """Load libmplb.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmplb.so is not built; run `python -m multiphaselb_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, cp, i, d = C.c_void_p, C.c_char_p, C.c_int, C.c_double
    sig = {
        "mplb_default_opts": ([C.POINTER(Opts)], None),
        "mplb_case_open": ([cp, C.POINTER(Opts), C.POINTER(vp)], i),
        "mplb_case_close": ([vp], i),
        "mplb_last_error": ([vp], cp),
        "mplb_case_info": ([vp, C.POINTER(Info)], i),
        "mplb_case_param": ([vp, C.c_char_p, vp, C.c_size_t], i),
        "mplb_set_field": ([vp, cp, vp, C.c_size_t], i),
        "mplb_get_field": ([vp, cp, vp, C.c_size_t], i),
        "mplb_get_neighbours": ([vp, vp, C.c_size_t], i),
        "mplb_get_boundary": ([vp, i, vp, C.c_size_t, cp, C.c_size_t], i),
        "mplb_init": ([vp], i),
        "mplb_step": ([vp, i, C.POINTER(i)], i),
        "mplb_collision": ([vp, i], i),
        "mplb_stream": ([vp, i], i),
        "mplb_update_boundaries": ([vp, i], i),
        "mplb_equilibrium": ([vp, i], i),
        "mplb_macro_density": ([vp], i),
        "mplb_macro_velocity": ([vp], i),
        "mplb_macro_temperature": ([vp], i),
        "mplb_inter_force": ([vp], i),
        "mplb_heat_source": ([vp], i),
        "mplb_macro_pressure": ([vp], i),
        "mplb_average_density": ([vp, C.POINTER(d)], i),
        "mplb_time_steps": ([vp, i, C.POINTER(d), vp, vp], i),
        "mplb_timer_name": ([i], cp),
        "mplb_lattice_tables": ([i, vp, vp, vp, vp, vp, vp], i),
        "mplb_glibc_rand": ([C.c_uint, i, vp], i),
        "mplb_batch_create": ([vp, i, i, i, C.POINTER(vp)], i),
        "mplb_batch_close": ([vp], i),
        "mplb_batch_last_error": ([vp], cp),
        "mplb_batch_set_case": ([vp, i, vp, C.c_size_t, d], i),
        "mplb_batch_init": ([vp], i),
        "mplb_batch_step": ([vp, i], i),
        "mplb_batch_run_to_steady": ([vp, d, i, C.c_int64, vp, vp, vp], i),
        "mplb_batch_get_field": ([vp, i, cp, vp, C.c_size_t], i),
        "mplb_batch_case_steps": ([vp, i], C.c_int64),
        "mplb_batch_tune": ([vp, i, C.POINTER(i), C.POINTER(i)], i),
        "mplb_batch_set_launch": ([vp, i, i], i),
        "mplb_batch_set_march": ([vp, i], i),
        "mplb_batch_march_rows": ([vp], i),
        "mplb_batch_set_persistent": ([vp, i], i),
        "mplb_batch_launch_info": ([vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(C.c_int64)], i),
        "mplb_density_change": ([vp, C.POINTER(d)], i),
        "mplb_write_lattice": ([vp, cp], i),
        "mplb_write_fields": ([vp, cp], i),
        "mplb_read_field": ([vp, cp, i, cp, i, vp, C.c_size_t], i),
        "mplb_set_radial_density": ([vp, vp, i, d, d, vp, d, C.POINTER(d)], i),
        "mplb_set_radial_density_noalloc": ([vp, vp, i, d, d, vp, d, C.POINTER(d)], i),
        "mplb_mlp_create": ([i, vp, i, i, C.POINTER(vp)], i),
        "mplb_mlp_close": ([vp], i),
        "mplb_mlp_last_error": ([vp], cp),
        "mplb_mlp_num_params": ([vp], C.c_int64),
        "mplb_mlp_set_params": ([vp, vp, C.c_size_t], i),
        "mplb_mlp_get_params": ([vp, vp, C.c_size_t], i),
        "mplb_mlp_forward": ([vp, vp, i, vp], i),
        "mplb_mlp_gradient": ([vp, vp, vp, C.POINTER(C.c_float), vp, C.c_size_t], i),
        "mplb_mlp_train": ([vp, vp, vp, i, vp, i, C.c_float, C.c_float, C.c_float, C.c_float, vp], i),
        "mplb_mlp_tensor": ([vp, cp, i, vp, C.c_size_t], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def lattice_tables(lbmodel):
    Q = 9 if lbmodel == 0 else 15
    vel = np.zeros((Q, 3), np.int32)
    rev = np.zeros(Q, np.int32)
    omega, w = np.zeros(Q), np.zeros(Q)
    M, invM = np.zeros((Q, Q)), np.zeros((Q, Q))
    library().mplb_lattice_tables(lbmodel, vel.ctypes.data, rev.ctypes.data, omega.ctypes.data,
                                  w.ctypes.data, M.ctypes.data, invM.ctypes.data)
    return dict(vel=vel, rev=rev, omega=omega, w=w, M=M, invM=invM)


def glibc_rand(seed, n):
    out = np.zeros(n, np.int32)
    library().mplb_glibc_rand(seed, n, out.ctypes.data)
    return out


class Case:
    """One case on the GPU (mplb_case_open / mplb_case_close)."""

    def __init__(self, case_dir, precision=FP64, heat_transfer=True, frozen=False, fused=False,
                 extensions=0, device=0):
        self.L = library()
        o = Opts()
        self.L.mplb_default_opts(C.byref(o))
        o.precision = precision
        o.heat_transfer = 1 if heat_transfer else 0
        o.frozen = 1 if frozen else 0
        o.fused = 1 if fused else 0
        o.extensions = extensions
        o.device = device
        h = C.c_void_p()
        rc = self.L.mplb_case_open(str(case_dir).encode(), C.byref(o), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.mplb_last_error(h).decode() if h.value else "open failed"
            self.close()
            raise MplbError(rc, msg)
        self.info = self.get_info()
        self.N = self.info.npoints
        self.Q = self.info.Q

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.mplb_case_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise MplbError(rc, self.L.mplb_last_error(self.h).decode())
        return rc

    def get_info(self):
        inf = Info()
        self._check(self.L.mplb_case_info(self.h, C.byref(inf)))
        return inf

    def param(self, key):
        """Values the dictionaries gave for key (mplb_case_param), as a float array."""
        n = self.L.mplb_case_param(self.h, key.encode(), None, 0)
        if n < 0:
            raise MplbError(n, "unknown parameter key %s" % key)
        out = np.zeros(max(n, 1))
        self.L.mplb_case_param(self.h, key.encode(), out.ctypes.data, n)
        return out[:n]

    def _shape(self, name):
        if name in ("rho", "T", "p", "ss"):
            return (self.N,)
        if name in ("U", "Fi", "Ft"):
            return (self.N, 3)
        return (self.N, self.Q)

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self._shape(name))
        self._check(self.L.mplb_set_field(self.h, name.encode(), a.ctypes.data, a.size))

    def get(self, name):
        a = np.zeros(self._shape(name))
        self._check(self.L.mplb_get_field(self.h, name.encode(), a.ctypes.data, a.size))
        return a

    def neighbours(self):
        a = np.zeros((self.N, self.Q), np.int32)
        self._check(self.L.mplb_get_neighbours(self.h, a.ctypes.data, a.size))
        return a

    def boundaries(self):
        out = []
        for b in range(self.info.nbd):
            n = self.L.mplb_get_boundary(self.h, b, None, 0, None, 0)
            a = np.zeros(n, np.int32)
            name = C.create_string_buffer(64)
            self._check(self.L.mplb_get_boundary(self.h, b, a.ctypes.data, n, name, 64))
            out.append((name.value.decode(), a))
        return out

    def init(self):
        self._check(self.L.mplb_init(self.h))

    def step(self, n=1):
        done = C.c_int(0)
        self._check(self.L.mplb_step(self.h, n, C.byref(done)))
        return done.value

    def write_lattice(self, d):
        self._check(self.L.mplb_write_lattice(self.h, str(d).encode()))

    def write_fields(self, d):
        self._check(self.L.mplb_write_fields(self.h, str(d).encode()))

    def read_field(self, d, time, name, comps):
        a = np.zeros((self.N, comps)) if comps > 1 else np.zeros(self.N)
        self._check(self.L.mplb_read_field(self.h, str(d).encode(), time, name.encode(), comps, a.ctypes.data, a.size))
        return a

    def density_change(self):
        """max |rho - rho at the previous call| (NaN on the first call after init)."""
        v = C.c_double()
        self._check(self.L.mplb_density_change(self.h, C.byref(v)))
        return v.value

    def set_radial_density(self, prof, d0, dd, centre, R, noalloc=False):
        """Writes rho = prof(|x - centre| - R) into the case; returns the total mass.
        noalloc: use the entry point that allocates no device memory."""
        pr = np.ascontiguousarray(prof, dtype=np.float64)
        cen = np.ascontiguousarray(list(centre) + [0.0] * (3 - len(centre)), dtype=np.float64)
        m = C.c_double()
        fn = self.L.mplb_set_radial_density_noalloc if noalloc else self.L.mplb_set_radial_density
        self._check(fn(self.h, pr.ctypes.data, pr.size, d0, dd, cen.ctypes.data, R, C.byref(m)))
        return m.value

    def average_density(self):
        v = C.c_double()
        self._check(self.L.mplb_average_density(self.h, C.byref(v)))
        return v.value

    def time_steps(self, n, per_kernel=True):
        tot = C.c_double()
        kms = np.zeros(NTIMERS)
        kn = np.zeros(NTIMERS, np.int32)
        done = self._check(self.L.mplb_time_steps(self.h, n, C.byref(tot),
                                                   kms.ctypes.data if per_kernel else None,
                                                   kn.ctypes.data if per_kernel else None))
        names = [self.L.mplb_timer_name(i).decode() for i in range(NTIMERS)]
        kern = {names[i]: (float(kms[i]), int(kn[i])) for i in range(NTIMERS) if kn[i] > 0}
        return done, tot.value, kern

    # fine-grained routines (field 0 = f, 1 = g)
    def collision(self, field):
        self._check(self.L.mplb_collision(self.h, field))

    def stream(self, field):
        self._check(self.L.mplb_stream(self.h, field))

    def update_boundaries(self, field):
        self._check(self.L.mplb_update_boundaries(self.h, field))

    def equilibrium(self, field):
        self._check(self.L.mplb_equilibrium(self.h, field))

    def macro_density(self):
        self._check(self.L.mplb_macro_density(self.h))

    def macro_velocity(self):
        self._check(self.L.mplb_macro_velocity(self.h))

    def macro_temperature(self):
        self._check(self.L.mplb_macro_temperature(self.h))

    def inter_force(self):
        self._check(self.L.mplb_inter_force(self.h))

    def heat_source(self):
        self._check(self.L.mplb_heat_source(self.h))

    def macro_pressure(self):
        self._check(self.L.mplb_macro_pressure(self.h))


class Batch:
    """Ensemble of D2Q9 periodic liMRT cases advanced together (mplb_batch_*).

    `cases` are Case handles (host-only ones, device=-1, are enough; one handle may be
    listed for several entries); `rho` and `T` give each entry's initial density (N
    values) and uniform temperature."""

    def __init__(self, cases, rho, T, precision=FP32, device=0):
        self.L = library()
        self.cases = list(cases)
        self.n = len(self.cases)
        arr = (C.c_void_p * self.n)(*[c.h.value for c in self.cases])
        h = C.c_void_p()
        rc = self.L.mplb_batch_create(arr, self.n, precision, device, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.mplb_batch_last_error(h).decode() if h.value else "batch create failed"
            self.close()
            raise MplbError(rc, msg)
        self.sizes = [c.N for c in self.cases]
        for k in range(self.n):
            self.set_case(k, rho[k], T[k])

    def set_case(self, k, rho, T):
        r = np.ascontiguousarray(rho, dtype=np.float64).ravel()
        self._check(self.L.mplb_batch_set_case(self.h, k, r.ctypes.data, r.size, float(T)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.mplb_batch_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise MplbError(rc, self.L.mplb_batch_last_error(self.h).decode())
        return rc

    def init(self):
        self._check(self.L.mplb_batch_init(self.h))

    def step(self, n):
        self._check(self.L.mplb_batch_step(self.h, n))

    def run_to_steady(self, threshold, check_every=100, max_steps=10 ** 6):
        steps = np.zeros(self.n, np.int64)
        delta = np.zeros(self.n)
        state = np.zeros(self.n, np.int32)
        nconv = self._check(self.L.mplb_batch_run_to_steady(self.h, threshold, check_every, max_steps,
                                                             steps.ctypes.data, delta.ctypes.data,
                                                             state.ctypes.data))
        return nconv, steps, delta, state

    def get(self, k, name="rho"):
        a = np.zeros(self.sizes[k])
        self._check(self.L.mplb_batch_get_field(self.h, k, name.encode(), a.ctypes.data, a.size))
        return a

    def case_steps(self, k):
        return int(self.L.mplb_batch_case_steps(self.h, k))

    def tune(self, steps_per_trial=20):
        b, x = C.c_int(), C.c_int()
        self._check(self.L.mplb_batch_tune(self.h, steps_per_trial, C.byref(b), C.byref(x)))
        return b.value, x.value

    def set_launch(self, block, xcd_order):
        self._check(self.L.mplb_batch_set_launch(self.h, block, 1 if xcd_order else 0))

    def set_march(self, rows):
        """Single-pass kernel with `rows` rows per work item (0: the two-pass kernels)."""
        self._check(self.L.mplb_batch_set_march(self.h, rows))

    def set_persistent(self, on=True):
        """Runs of steps as one persistent cooperative launch of the single-pass kernel."""
        self._check(self.L.mplb_batch_set_persistent(self.h, 1 if on else 0))

    def launch_info(self):
        b, x, w, n = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        self._check(self.L.mplb_batch_launch_info(self.h, C.byref(b), C.byref(x), C.byref(w), C.byref(n)))
        return dict(block=b.value, xcd_order=x.value, workgroups=w.value, nodes=n.value,
                    march_rows=self.L.mplb_batch_march_rows(self.h))


class Mlp:
    """BF16-MFMA multilayer perceptron on the GPU (mplb_mlp_*)."""

    def __init__(self, widths, batch, device=0):
        self.L = library()
        self.widths = list(widths)
        self.batch = batch
        w = np.ascontiguousarray(self.widths, dtype=np.int32)
        h = C.c_void_p()
        rc = self.L.mplb_mlp_create(len(widths) - 1, w.ctypes.data, batch, device, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.mplb_mlp_last_error(h).decode() if h.value else "network create failed"
            self.close()
            raise MplbError(rc, msg)
        self.P = int(self.L.mplb_mlp_num_params(self.h))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.mplb_mlp_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise MplbError(rc, self.L.mplb_mlp_last_error(self.h).decode())
        return rc

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.L.mplb_mlp_set_params(self.h, a.ctypes.data, a.size))

    def get_params(self):
        a = np.zeros(self.P, np.float32)
        self._check(self.L.mplb_mlp_get_params(self.h, a.ctypes.data, a.size))
        return a

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.widths[-1]), np.float32)
        self._check(self.L.mplb_mlp_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(self.P, np.float32)
        self._check(self.L.mplb_mlp_gradient(self.h, x.ctypes.data, t.ctypes.data, C.byref(loss), g.ctypes.data, g.size))
        return float(loss.value), g

    def train(self, X, T, idx, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        iters = idx.size // self.batch
        hist = np.zeros(iters, np.float32)
        self._check(self.L.mplb_mlp_train(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], idx.ctypes.data, iters,
                                          lr, beta1, beta2, eps, hist.ctypes.data))
        return hist

    def tensor(self, name, j, shape):
        a = np.zeros(shape, np.float32)
        self._check(self.L.mplb_mlp_tensor(self.h, name.encode(), j, a.ctypes.data, a.size))
        return a

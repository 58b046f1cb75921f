"""Host-side checks of the batched-ensemble API and its case generator (no GPU needed):
mplb_batch_create validates every case before it touches a device."""
import ctypes as C
import collections

import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import make_case_dir
from multiphaselb_amd.casefiles import CaseSpec
from multiphaselb_amd.initfields import ensemble, ensemble_density, maxwell, p_eos, vdw_critical


def host(**kw):
    base = dict(lbmodel=0, nx=16, ny=16, btype=3, f_model=0, a=0.5, b=4.0)
    base.update(kw)
    ht = base.pop("ht", False)
    return M.Case(make_case_dir(CaseSpec(**base)), heat_transfer=ht, device=-1)


def create(cases, precision=M.FP32):
    L = M.library()
    arr = (C.c_void_p * len(cases))(*[c.h.value for c in cases])
    h = C.c_void_p()
    rc = L.mplb_batch_create(arr, len(cases), precision, 0, C.byref(h))
    msg = L.mplb_batch_last_error(h).decode() if h.value else ""
    if h.value:
        L.mplb_batch_close(h)
    return rc, msg


@pytest.mark.parametrize("kw,needle", [
    (dict(lbmodel=1, nz=16, btype=5, f_model=2, f_Lambda=(1.0,) * 15, g_Lambda=(1.0,) * 15), "D2Q9 periodicXY"),
    (dict(btype=0), "D2Q9 periodicXY"),
    (dict(f_st=1, f_kappa=0.3), "liSurfTen"),
    (dict(rho0=1.0), "rho_0"),
    (dict(ht=True, g_model=1), "heat transfer"),
])
def test_batch_rejects_unsupported_cases(kw, needle):
    good = host()
    rc, msg = create([good, host(**kw)])
    assert rc == -3 and needle in msg, (rc, msg)


def test_batch_requires_shared_constants():
    rc, msg = create([host(), host(a=0.4)])
    assert rc == -3 and "share" in msg
    rc, msg = create([host(), host(f_sigma=0.1)])
    assert rc == -3 and "share" in msg


def test_batch_argument_errors():
    L = M.library()
    h = C.c_void_p()
    assert L.mplb_batch_create(None, 1, M.FP32, 0, C.byref(h)) == -1
    c = host()
    arr = (C.c_void_p * 1)(c.h.value)
    assert L.mplb_batch_create(arr, 1, 7, 0, C.byref(h)) == -1
    assert L.mplb_batch_create(arr, 0, M.FP32, 0, C.byref(h)) == -1
    assert L.mplb_batch_close(None) == -1


def test_ensemble_generator():
    """Sizes, radii, temperatures and shapes follow the stated distributions; the fields
    are reproducible, duplicate-consistent and start between the coexistence densities."""
    e = ensemble(1024, seed=7)
    assert e == ensemble(1024, seed=7)
    cnt = collections.Counter(c["L"] for c in e)
    assert set(cnt) == {64, 96, 128, 192, 256} and min(cnt.values()) > 150
    assert all(6.0 <= c["radius"] <= c["L"] / 4.0 for c in e)
    assert all(0.6 <= c["tr"] <= 0.95 for c in e)
    assert 450 < sum(c["droplet"] for c in e) < 574
    tc = vdw_critical(0.5, 4.0)[1]
    for c in e[:20]:
        r = ensemble_density(c).reshape(c["L"], c["L"])
        np.testing.assert_array_equal(r[:, 0], r[:, -1])
        np.testing.assert_array_equal(r[0, :], r[-1, :])
        assert r.min() > c["rho_v"] - 2e-4 and r.max() < c["rho_l"] + 2e-4
        assert abs(c["T"] - c["tr"] * tc) < 1e-15
        centre = r[(c["L"] - 1) // 2, (c["L"] - 1) // 2]
        assert abs(centre - (c["rho_l"] if c["droplet"] else c["rho_v"])) < 0.05 * c["rho_l"]
        # the potential is defined over the whole profile: p < rho cs2
        assert (p_eos(r, c["T"], a=0.5, b=4.0) < r / 3.0).all()
    rv, rl = maxwell(e[0]["T"], a=0.5, b=4.0)
    assert (rv, rl) == (e[0]["rho_v"], e[0]["rho_l"])

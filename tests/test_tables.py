"""Host-side checks of the library that need no GPU: exported symbols, lattice
constants against the reference's compiled tables, the lattice tables against the
oracle's restatement of latticeBox2D/3D, and the dictionary reader on the reference's
shipped case dictionaries."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import _oracle as O
import multiphaselb_amd as M
from _cases import make_case_dir, oracle_params
from multiphaselb_amd.casefiles import BC, CaseSpec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mplb.h")).read()
    declared = set(re.findall(r"\b(mplb_[a-z_0-9]+)\s*\(", hdr))
    lib = ctypes.CDLL(M.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(M.EXPORTS)


def _ref_tables():
    path = O.REF_TABLES
    if not os.path.exists(path):
        if os.path.isdir("/root/reference"):
            subprocess.check_call(["make", "-s", "-C", O.ORACLE_DIR, "ref"])
        else:
            pytest.skip("oracle/_ref not built and reference tree absent")
    return ctypes.CDLL(path)


def _ref_arrays(R, pre, Q):
    def mat(name):
        f = getattr(R, pre + name)
        f.restype = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
        m = f()
        return np.array([[m[i][j] for j in range(Q)] for i in range(Q)])

    def vec(name, ty=ctypes.c_double):
        f = getattr(R, pre + name)
        f.restype = ctypes.POINTER(ty)
        v = f()
        return np.array([v[i] for i in range(Q)])

    f = getattr(R, pre + "_Vel")
    f.restype = ctypes.POINTER(ctypes.POINTER(ctypes.c_int))
    v = f()
    vel = np.array([[v[i][j] for j in range(3)] for i in range(Q)])
    cs2 = getattr(R, pre + "_cs2")
    cs2.restype = ctypes.c_double
    return dict(M=mat("_MRT"), invM=mat("_MRTInv"), omega=vec("_omega"), w=vec("_ppw"),
                rev=vec("_Reverse", ctypes.c_int), vel=vel, cs2=cs2())


@pytest.mark.parametrize("model,pre,Q", [(0, "D2Q9", 9), (1, "D3Q15", 15)])
def test_lattice_constants_bit_exact(model, pre, Q):
    ref = _ref_arrays(_ref_tables(), pre, Q)
    for src in (M.lattice_tables(model), O.lattice_tables(model)):
        for key in ("M", "invM", "omega", "w"):
            assert np.array_equal(np.asarray(src[key], float).view(np.int64), ref[key].view(np.int64)), key
        assert np.array_equal(src["vel"], ref["vel"])
        assert np.array_equal(src["rev"], ref["rev"])
    assert O.lattice_tables(model)["cs2"] == ref["cs2"]


BOXES = [
    (0, 7, 9, 1, 0), (0, 8, 8, 1, 0), (0, 11, 6, 1, 1), (0, 7, 7, 1, 2), (0, 10, 9, 1, 3),
    (0, 3, 12, 1, 1), (1, 6, 5, 7, 0), (1, 5, 6, 8, 3), (1, 6, 6, 6, 4), (1, 6, 5, 7, 5),
    (1, 7, 7, 4, 5), (1, 9, 8, 5, 6),
]


@pytest.mark.parametrize("lbmodel,nx,ny,nz,btype", BOXES)
def test_lattice_tables_match_oracle(lbmodel, nx, ny, nz, btype):
    spec = CaseSpec(lbmodel=lbmodel, nx=nx, ny=ny, nz=nz, btype=btype, spot_radius=2.5,
                    spot_centre=(4.0, 3.0), f_model=0 if lbmodel == 0 else 2,
                    f_Lambda=(1.0,) * (9 if lbmodel == 0 else 15),
                    g_Lambda=(0.0,) + (1.0,) * (8 if lbmodel == 0 else 14), g_model=1)
    d = make_case_dir(spec)
    c = M.Case(d, device=-1)
    o = O.Oracle(oracle_params(spec))
    assert o.error()[0] == 0
    assert np.array_equal(c.neighbours(), o.nb())
    bl = c.boundaries()
    ob = o.boundaries()
    assert [n for n, _ in bl] == spec.patches()
    assert len(bl) == len(ob)
    for (_, a), b in zip(bl, ob):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [5, 8, 13])
def test_periodic_tables_are_duplicate_node_wrap(n):
    """Fully periodic boxes: every entry is the upstream site under period N-1.  In 3D
    the stored node is exactly the wrapped one; in 2D the corner completion
    (periodicXY.c:67-96) picks the other copy of the same site for four links."""
    for lbmodel, btype in ((0, 3), (1, 5)):
        spec = CaseSpec(lbmodel=lbmodel, nx=n, ny=n + 1, nz=n + 2 if lbmodel else 1, btype=btype,
                        f_model=0 if lbmodel == 0 else 2, f_Lambda=(1.0,) * (9 if lbmodel == 0 else 15),
                        g_model=None)
        o = O.Oracle(oracle_params(spec))
        nb = o.nb()
        t = O.lattice_tables(lbmodel)
        nx, ny, nzz = spec.nx, spec.ny, (spec.nz if lbmodel else 1)
        wrap = lambda u, m: u + (m - 1) if u < 0 else (u - (m - 1) if u > m - 1 else u)
        site = lambda v: (v % nx % (nx - 1), (v // nx) % ny % (ny - 1), (v // (nx * ny)) % max(nzz - 1, 1))
        exact_miss = 0
        for idx in range(nx * ny * nzz):
            i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
            for q in range(t["Q"]):
                c = t["vel"][q]
                ui, uj = wrap(i - c[0], nx), wrap(j - c[1], ny)
                uk = wrap(k - c[2], nzz) if lbmodel else 0
                want = ui + uj * nx + uk * nx * ny
                assert site(int(nb[idx, q])) == site(want)
                if nb[idx, q] != want:
                    exact_miss += 1
                    assert lbmodel == 0 and (i in (0, nx - 1)) and (j in (0, ny - 1))
        assert exact_miss == (4 if lbmodel == 0 else 0)


def _ref_box():
    path = os.path.join(O.ORACLE_DIR, "_ref", "librefbox.so")
    if not os.path.exists(path):
        if os.path.isdir("/root/reference"):
            subprocess.check_call(["make", "-s", "-C", O.ORACLE_DIR, "ref"])
        else:
            pytest.skip("oracle/_ref not built and reference tree absent")
    R = ctypes.CDLL(path)
    R.refbox_build.argtypes = [ctypes.c_int] * 6 + [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_int64, ctypes.POINTER(ctypes.c_long)]
    return R


def _search_table(lbmodel, nx, ny, nz):
    """The neighbour search of latticeBox2D.c:148-239 / latticeBox3D.c:166-215, restated
    from its definition: the upstream node x - c_k when it lies inside the box, else -1."""
    t = M.lattice_tables(lbmodel)
    nzz = nz if lbmodel == 1 else 1
    k, j, i = np.meshgrid(np.arange(nzz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    Q = len(t["rev"])
    nb = np.full((i.size, Q), -1, np.int32)
    for q in range(Q):
        c = t["vel"][q]
        ui, uj, uk = i - c[0], j - c[1], k - c[2]
        ok = (ui >= 0) & (ui < nx) & (uj >= 0) & (uj < ny) & (uk >= 0) & (uk < nzz)
        nb[ok, q] = (ui + uj * nx + uk * nx * ny)[ok]
    return nb


REF_BOXES = [
    (0, 7, 9, 1, 0), (0, 8, 8, 1, 0), (0, 9, 8, 1, 0), (0, 5, 12, 1, 0),
    (0, 11, 6, 1, 1), (0, 3, 12, 1, 1), (0, 8, 8, 1, 1), (0, 6, 13, 1, 1),
    (0, 7, 7, 1, 2), (0, 5, 9, 1, 2), (0, 6, 11, 1, 2),
    (0, 10, 9, 1, 3), (0, 7, 7, 1, 3), (0, 4, 11, 1, 3), (0, 12, 5, 1, 3),
    (1, 6, 5, 7, 0), (1, 5, 6, 8, 0), (1, 4, 4, 4, 0), (1, 7, 6, 5, 0),
    (1, 5, 6, 8, 3), (1, 6, 5, 7, 3), (1, 7, 7, 4, 3),
    (1, 6, 6, 6, 4), (1, 5, 5, 5, 4), (1, 6, 4, 6, 4), (1, 7, 5, 7, 4),
    (1, 6, 5, 7, 5), (1, 9, 8, 5, 5), (1, 4, 4, 4, 5),
]


@pytest.mark.parametrize("lbmodel,nx,ny,nz,btype", REF_BOXES)
def test_box_tables_match_the_reference_routines(lbmodel, nx, ny, nz, btype):
    """The reference's own genericBoundary*/periodic*.c, compiled into oracle/_ref from
    their sources and run on the neighbour search's table, give the same completed
    neighbour table and boundary lists (contents and order) as the library and the
    oracle, bit for bit.  periodicX.c also writes outside its lists (its loop runs j from
    0 to ny-1 into entries j-1 of (ny-2)-long lists), and so does periodicZ3D.c for
    ny < nz (nz*(nx-2) entries into (nx-2)*ny-long Y lists); the guarded allocator sees
    both."""
    R = _ref_box()
    nb = _search_table(lbmodel, nx, ny, nz)
    N, Q = nb.shape
    sizes = np.zeros(6, np.int32)
    lists = np.zeros(8 * N, np.int32)
    oob = ctypes.c_long(0)
    n = R.refbox_build(3 if lbmodel == 1 else 2, btype, nx, ny, nz, Q, nb.ctypes.data, 6, sizes.ctypes.data,
                       lists.ctypes.data, lists.size, ctypes.byref(oob))
    assert n >= 0
    ref_lists = []
    at = 0
    for b in range(n):
        ref_lists.append(lists[at:at + sizes[b]].copy())
        at += sizes[b]
    assert (oob.value > 0) == ((lbmodel == 0 and btype == 1) or (lbmodel == 1 and btype == 4 and ny < nz))
    spec = CaseSpec(lbmodel=lbmodel, nx=nx, ny=ny, nz=nz, btype=btype, f_model=0 if lbmodel == 0 else 2,
                    f_Lambda=(1.0,) * (9 if lbmodel == 0 else 15),
                    g_Lambda=(0.0,) + (1.0,) * (8 if lbmodel == 0 else 14), g_model=1)
    c = M.Case(make_case_dir(spec), device=-1)
    o = O.Oracle(oracle_params(spec))
    assert o.error()[0] == 0
    assert np.array_equal(c.neighbours(), nb)
    assert np.array_equal(o.nb(), nb)
    lib_lists = [a for _, a in c.boundaries()]
    assert len(lib_lists) == n and len(o.boundaries()) == n
    for a, b, r in zip(lib_lists, o.boundaries(), ref_lists):
        assert np.array_equal(a, r)
        assert np.array_equal(b, r)


def _parse(case):
    d = os.path.join(GOLD, "dicts", case)
    return M.Case(d, device=-1)


def test_dictionary_reader_on_reference_examples():
    c = _parse("singleBubble")
    inf = c.info
    assert (inf.lbmodel, inf.nx, inf.ny, inf.nbd, inf.f_model, inf.g_model) == (0, 101, 101, 0, 0, 1)
    assert (inf.start_time, inf.end_time, inf.write_interval) == (0, 5000, 2500)
    c = _parse("VdWColumn_HT")
    inf = c.info
    assert (inf.nx, inf.ny, inf.nbd) == (3, 301, 4)
    assert [n for n, _ in c.boundaries()] == ["X0", "X1", "Y0", "Y1"]
    c = _parse("singleBubble3D")
    inf = c.info
    assert (inf.lbmodel, inf.nx, inf.ny, inf.nz, inf.nbd, inf.f_model, inf.g_model) == (1, 101, 101, 101, 0, 2, 1)
    c = _parse("VdWColumn3D")
    assert c.info.lbmodel == 1


# Every value the six shipped example dictionaries give, read off the files by hand.
_F9 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 0.8, 0.8)
_G9 = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0)
_F15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)
_G15 = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
_VDW = {"EOS/model": (3,), "EOS/a": (0.5,), "EOS/b": (4.0,), "EOS/R": (1.0,), "EOS/M": (1.0,), "EOS/G": (-1.0,),
        "EOS/rho_0": (0.0,), "EOS/Cv": (1.0,), "EOS/extForce": (0.0, 0.0, 0.0)}
_P, _BB, _FT = 8, 0, 1
SHIPPED = {
    "VdWColumn": dict(
        {**_VDW, "EOS/g": (0.0, -1.234567e-07, 0.0), "time/start": (0,), "time/end": (200000,),
         "time/writeInterval": (25000,), "f/Collision": (0,), "f/Lambda": _F9, "f/sigma": (0.125,),
         "f/SurfaceTension": (0,), "g/Collision": (1,), "g/Lambda": _G9, "g/alpha_1": (0.25,), "g/alpha_2": (0.25,)},
        bc={"f": [(_P,), (_P,), (_BB,), (_BB,)], "g": [(_P,), (_P,), (_FT, 0.03666667), (_FT, 0.03666667)]}),
    "VdWColumn3D": dict(
        {**_VDW, "EOS/g": (0.0, -1.234567e-07, 0.0), "time/start": (0,), "time/end": (200000,),
         "time/writeInterval": (25000,), "f/Collision": (2,), "f/Lambda": _F15, "f/sigma": (1.2,),
         "f/SurfaceTension": (0,), "g/Collision": (1,), "g/Lambda": _G15, "g/alpha_1": (0.25,), "g/alpha_2": (0.25,)},
        # periodicXY in 3D has the Z0 and Z1 lists only; the X and Y entries are never read
        bc={"f": [(_BB,), (_BB,)], "g": [(_FT, 0.03666667), (_FT, 0.03666667)]}),
    "VdWColumn_HT": dict(
        {**_VDW, "EOS/g": (0.0, -1.234567e-07, 0.0), "time/start": (0,), "time/end": (1000000,),
         "time/writeInterval": (100000,), "f/Collision": (0,), "f/Lambda": _F9, "f/sigma": (0.125,),
         "f/SurfaceTension": (0,), "g/Collision": (1,), "g/Lambda": (0.0, 1.0, 1.0, 0.8, 1.0, 0.8, 1.0, 1.0, 1.0),
         "g/alpha_1": (1.0,), "g/alpha_2": (1.0,)},
        bc={"f": [(_P,), (_P,), (_BB,), (_BB,)], "g": [(_P,), (_P,), (_FT, 0.033333), (_FT, 0.036667)]}),
    "maxwell": dict(
        {**_VDW, "EOS/g": (0.0, 0.0, 0.0), "time/start": (0,), "time/end": (5000,), "time/writeInterval": (500,),
         "f/Collision": (0,), "f/Lambda": _F9, "f/sigma": (0.125,), "f/SurfaceTension": (0,), "g/Collision": (1,),
         "g/Lambda": _G9, "g/alpha_1": (0.25,), "g/alpha_2": (0.25,)},
        bc={"f": [], "g": []}),
    "singleBubble": dict(
        {**_VDW, "EOS/g": (0.0, 0.0, 0.0), "time/start": (0,), "time/end": (5000,), "time/writeInterval": (2500,),
         "f/Collision": (0,), "f/Lambda": (1.0, 0.8, 1.1, 1.0, 1.1, 1.0, 1.1, 0.8, 0.8), "f/sigma": (0.125,),
         "f/SurfaceTension": (0,), "g/Collision": (1,), "g/Lambda": _G9, "g/alpha_1": (0.25,), "g/alpha_2": (0.25,)},
        bc={"f": [], "g": []}),
    "singleBubble3D": dict(
        {**_VDW, "EOS/g": (0.0, 0.0, 0.0), "time/start": (0,), "time/end": (5000,), "time/writeInterval": (100,),
         "f/Collision": (2,), "f/Lambda": _F15, "f/sigma": (0.125,), "f/SurfaceTension": (0,), "g/Collision": (1,),
         "g/Lambda": _G15, "g/alpha_1": (0.25,), "g/alpha_2": (0.25,)},
        bc={"f": [], "g": []}),
}


@pytest.mark.parametrize("name", sorted(SHIPPED))
def test_every_value_of_the_shipped_dictionaries(name):
    """The library reads every value of each of the reference's six example case
    dictionaries: EOS block, f and g model constants, times and the boundary conditions
    list by list (periodicXY boxes have no lists)."""
    want = dict(SHIPPED[name])
    bc = want.pop("bc")
    c = _parse(name)
    for key, vals in want.items():
        got = c.param(key)
        assert np.array_equal(got, np.array(vals, float)), (key, got, vals)
    for fld in ("f", "g"):
        assert c.info.nbd == len(bc[fld])
        for b, exp in enumerate(bc[fld]):
            assert c.param("%s/bc/%d/type" % (fld, b))[0] == exp[0], (fld, b)
            if len(exp) > 1:
                assert c.param("%s/bc/%d/value" % (fld, b))[0] == exp[1], (fld, b)


def test_dictionary_first_match_rules(tmp_path):
    """Keys match as substrings at their first occurrence (singleEntry.c:14)."""
    spec = CaseSpec(eos=2, a=0.5, b=4.0, w_pr=0.3, g=(0.0, -1e-5, 0.0))
    d = make_case_dir(spec, root=str(tmp_path))
    text = open(os.path.join(d, "properties", "macroProperties")).read()
    assert "Peng-Robinson" in text
    c = M.Case(d, device=-1)
    assert c.info.f_model == 0


def test_reference_errors_become_codes(tmp_path):
    spec = CaseSpec(lbmodel=1, nx=6, ny=6, nz=6, btype=5, f_model=0,
                    f_Lambda=(1.0,) * 15, g_Lambda=(0.0,) + (1.0,) * 14)
    d = make_case_dir(spec, root=str(tmp_path))
    with pytest.raises(M.MplbError) as e:
        M.Case(d, device=-1)
    assert "not implemented for D3Q15" in e.value.msg
    M.Case(d, device=-1, extensions=M.EXT_LIMRT_D3Q15)
    spec = CaseSpec(lbmodel=0, nx=6, ny=6, btype=1, f_bc={"Y0": BC("fixedGradT", (1.0,))})
    d = make_case_dir(spec, root=str(tmp_path))
    with pytest.raises(M.MplbError) as e:
        M.Case(d, device=-1)
    assert "Unrecognized boundary condition" in e.value.msg


def test_glibc_rand_matches_host_libc():
    libc = ctypes.CDLL("libc.so.6")
    for seed in (0, 1, 12345):
        libc.srand(seed)
        ref = np.array([libc.rand() for _ in range(3000)], np.int32)
        assert np.array_equal(M.glibc_rand(seed, 3000), ref)


def test_generic_2d_y1_list_outside_lattice_is_rejected(tmp_path):
    """genericBoundary.c:43 indexes Y1 with (nx-1)*ny; for nx > ny+1 that leaves the lattice."""
    spec = CaseSpec(lbmodel=0, nx=9, ny=7, btype=0)
    with pytest.raises(M.MplbError):
        M.Case(make_case_dir(spec, root=str(tmp_path)), device=-1)


@pytest.mark.parametrize("nx,ny,nz", [(5, 7, 5), (6, 7, 6)])
def test_periodic_z_shapes_the_reference_leaves_undefined_are_refused(tmp_path, nx, ny, nz):
    """periodicZ3D.c leaves Y-list entries unset for ny > nz, and for ny >= nx + 2 its
    x-edge pass reads and writes node j*ny + (nz-1)*nx*ny outside the lattice."""
    spec = CaseSpec(lbmodel=1, nx=nx, ny=ny, nz=nz, btype=4, f_model=2, f_Lambda=(1.0,) * 15, g_model=None)
    with pytest.raises(M.MplbError) as e:
        M.Case(make_case_dir(spec, root=str(tmp_path)), device=-1)
    assert "Ny <= Nz" in e.value.msg
    o = O.Oracle(oracle_params(spec))
    assert o.error()[0] != 0

"""Reference on-disk formats without a GPU: lattice tables written by mplb_write_lattice
are read back by mplb_case_open (as readLatticeMesh.c would read them) into identical
tables, the written files follow writeLatticeMesh.c's layout, and a case directory with
lattice files but an unrelated latticeProperties box takes its tables from the files."""
import os
import shutil

import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import make_case_dir
from multiphaselb_amd.casefiles import BC, CaseSpec


@pytest.mark.parametrize("spec", [
    CaseSpec(lbmodel=0, nx=9, ny=7, btype=3),
    CaseSpec(lbmodel=0, nx=7, ny=9, btype=0, f_bc={p: BC("bounceBack") for p in ("X0", "X1", "Y0", "Y1")}),
    CaseSpec(lbmodel=1, nx=6, ny=5, nz=7, btype=5, f_model=2, f_Lambda=(1.0,) * 15, g_Lambda=(1.0,) * 15),
    CaseSpec(lbmodel=1, nx=6, ny=6, nz=8, btype=3, f_model=2, f_Lambda=(1.0,) * 15, g_Lambda=(1.0,) * 15,
             f_bc={"Z0": BC("bounceBack"), "Z1": BC("bounceBack")}),
], ids=["d2q9_periodicXY", "d2q9_generic", "d3q15_periodicXYZ", "d3q15_periodicXY_walls"])
def test_lattice_files_roundtrip(spec):
    d = make_case_dir(spec)
    c = M.Case(d, device=-1, heat_transfer=False)
    nb, bds = c.neighbours(), c.boundaries()
    c.write_lattice(d)
    base = os.path.join(d, "processor0", "lattice")
    lines = open(os.path.join(base, "neighbours")).read().splitlines()
    assert lines[0] == "%d %d %d" % (c.N, 2 if spec.lbmodel == 0 else 3, c.Q)
    assert lines[1] == " ".join(str(v) for v in nb[0]) + " "
    pts = open(os.path.join(base, "points")).read().split()
    assert int(pts[0]) == c.N and pts[1:7] == ["0", "0", "0", "1", "0", "0"]
    cells = open(os.path.join(base, "vtkCells")).read().splitlines()
    ncell = (spec.nx - 1) * (spec.ny - 1) * ((spec.nz - 1) if spec.lbmodel == 1 else 1)
    assert cells[0] == "%d %d" % (ncell, 8 if spec.lbmodel == 1 else 4)
    # reopen: the tables now come from the files
    c2 = M.Case(d, device=-1, heat_transfer=False)
    np.testing.assert_array_equal(c2.neighbours(), nb)
    assert [(n, list(a)) for n, a in c2.boundaries()] == [(n, list(a)) for n, a in bds]


def test_lattice_files_take_precedence():
    spec = CaseSpec(lbmodel=0, nx=8, ny=8, btype=3)
    src = make_case_dir(spec)
    c = M.Case(src, device=-1, heat_transfer=False)
    nb = c.neighbours()
    c.write_lattice(src)
    other = make_case_dir(CaseSpec(lbmodel=0, nx=5, ny=11, btype=0))
    shutil.copytree(os.path.join(src, "processor0"), os.path.join(other, "processor0"))
    # the boundary lists of the 8x8 periodic box carry no patches
    c2 = M.Case(other, device=-1, heat_transfer=False)
    assert c2.N == 64 and c2.info.nx == 8 and c2.info.ny == 8
    np.testing.assert_array_equal(c2.neighbours(), nb)


def test_lattice_files_errors():
    spec = CaseSpec(lbmodel=0, nx=6, ny=6, btype=3)
    d = make_case_dir(spec)
    c = M.Case(d, device=-1, heat_transfer=False)
    c.write_lattice(d)
    path = os.path.join(d, "processor0", "lattice", "ghosts")
    open(path, "w").write("4\n")
    with pytest.raises(M.MplbError) as e:
        M.Case(d, device=-1, heat_transfer=False)
    assert "ghost" in str(e.value)

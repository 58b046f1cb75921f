"""Field files of a time level in the reference's asciiRaw / binaryRaw formats, written
from the device fields and read back through the reference's float path."""
import dataclasses
import os

import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import make_case_dir, sphere_fields
from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["asciiRaw", "binaryRaw"])
def test_write_and_read_fields(fmt):
    spec, fk, ht, frozen, steps = CASES["d2q9_bubble_thermal"]
    spec = dataclasses.replace(spec, data_format=fmt)
    d = make_case_dir(spec)
    c = M.Case(d, precision=M.FP64, heat_transfer=True)
    rho, U, T = sphere_fields(spec, seed=2, **fk)
    for name, v in (("rho", rho), ("U", U), ("T", T)):
        c.set(name, v)
    c.init()
    c.step(spec.write_interval)                # ends on a write step: p is current
    t = c.get_info().current_time
    c.write_fields(d)
    for name, comps in (("rho", 1), ("T", 1), ("p", 1), ("U", 3), ("f", 9), ("g", 9)):
        ref = c.get(name)
        got = c.read_field(d, t, name, comps)
        if fmt == "binaryRaw":
            expect = ref.astype(np.float32).astype(np.float64)
        else:
            expect = np.vectorize(lambda v: float(np.float32(float("%g" % v))))(ref)
        np.testing.assert_array_equal(got, expect)
    first = open(os.path.join(d, "processor0", str(t), "U"), "rb").read()
    if fmt == "asciiRaw":
        assert first.decode().splitlines()[0] == " ".join("%g" % v for v in c.get("U")[0]) + " "
    else:
        assert len(first) == 4 * 3 * c.N


def test_write_fields_nan_check():
    spec, fk, ht, frozen, steps = CASES["d2q9_droplet_nht"]
    d = make_case_dir(spec)
    c = M.Case(d, precision=M.FP64, heat_transfer=False)
    rho, U, T = sphere_fields(spec, seed=2, **fk)
    rho[5] = np.nan
    c.set("rho", rho)
    c.set("U", U)
    c.set("T", T)
    with pytest.raises(M.MplbError) as e:
        c.write_fields(d)
    assert e.value.code == -5 and "NaN" in str(e.value)


@pytest.mark.parametrize("name", ["d2q9_bubble_thermal", "d3q15_bubble_nht"])
def test_pvtu_fields(name):
    """pvtu output: fields.vtu with the box's points and cells and every field as an
    ascii DataArray (read back in double, as readScalarFromPvtu.c does with DP), the
    parallel fields.pvtu naming the fields of start/initialFields, and case.pvd listing
    the time folders."""
    spec, fk, ht, frozen, steps = CASES[name]
    spec = dataclasses.replace(spec, data_format="pvtu")
    d = make_case_dir(spec)
    c = M.Case(d, precision=M.FP64, heat_transfer=ht)
    rho, U, T = sphere_fields(spec, seed=4, **fk)
    for n, v in (("rho", rho), ("U", U), ("T", T)):
        c.set(n, v)
    c.init()
    c.step(spec.write_interval)
    t = c.get_info().current_time
    c.write_fields(d)
    Q = 9 if spec.lbmodel == 0 else 15
    for n, comps in (("rho", 1), ("T", 1), ("p", 1), ("U", 3), ("f", Q), ("g", Q)):
        got = c.read_field(d, t, n, comps)
        expect = np.vectorize(lambda v: float("%g" % v))(c.get(n))
        np.testing.assert_array_equal(got, expect.reshape(got.shape))
    vtu = open(os.path.join(d, "processor0", str(t), "fields.vtu")).read().splitlines()
    dim3 = spec.lbmodel == 1
    ncells = (spec.nx - 1) * (spec.ny - 1) * ((spec.nz - 1) if dim3 else 1)
    assert vtu[3] == '    <Piece NumberOfPoints="%d" NumberOfCells="%d">' % (c.N, ncells)
    assert vtu[6] == "          0.0 0.0 0.0"
    assert vtu[7] == "          1.0 0.0 0.0"
    conn = vtu[6 + c.N + 4].split()
    if dim3:
        nx, nxy = spec.nx, spec.nx * spec.ny
        assert conn == [str(v) for v in (0, 1, nx, nx + 1, nxy, nxy + 1, nxy + nx, nxy + nx + 1)]
    else:
        assert conn == ["0", "1", str(spec.nx), str(spec.nx + 1)]
    assert '      <PointData Scalars="rho T p" Vectors="U f g">' in vtu
    assert vtu[-1] == "</VTKFile>"
    pv = open(os.path.join(d, str(t), "fields.pvtu")).read()
    assert '<PDataArray type="Float32" Name="f" NumberOfComponents="%d"/>' % Q in pv
    assert '<Piece Source="../processor0/%d/fields.vtu"/>' % t in pv
    c.step(spec.write_interval)
    c.write_fields(d)
    pvd = open(os.path.join(d, "case.pvd")).read()
    assert pvd.count("<DataSet") == 2 and 'file="%d/fields.pvtu"' % t in pvd


@pytest.mark.parametrize("name", ["d2q9_bubble_thermal", "d3q15_bubble_nht"])
def test_ensight_fields(name):
    """ensight output (multiPhasePP.c:283-301 with IO/ensight/*): each write step appends
    its time to lbm.case and writes lattice.<field>_<index> files (U as three component
    blocks, f and g one file per population) that read back as the float32 values; a
    folder without lbm.case starts its list at time 0, as updateCaseFile.c does."""
    import _ensight as E
    spec, fk, ht, frozen, steps = CASES[name]
    spec = dataclasses.replace(spec, data_format="ensight")
    d = make_case_dir(spec)
    c = M.Case(d, precision=M.FP64, heat_transfer=ht)
    rho, U, T = sphere_fields(spec, seed=6, **fk)
    for n, v in (("rho", rho), ("U", U), ("T", T)):
        c.set(n, v)
    c.init()
    c.write_lattice(d)                             # lattice.geo + lbm.case listing time 0
    Q = c.Q
    fields = (("rho", 1), ("T", 1), ("p", 1), ("U", 3), ("f", Q), ("g", Q))
    saved, times = {}, [0]
    for w in range(2):
        c.step(spec.write_interval)
        t = c.get_info().current_time
        times.append(t)
        c.write_fields(d)
        assert E.case_times(os.path.join(d, "lbm.case")) == times
        saved[t] = {n: c.get(n).astype(np.float32).astype(np.float64) for n, _ in fields}
        print("write", w, "time", t, flush=True)
    for t, want in saved.items():
        for n, comps in fields:
            np.testing.assert_array_equal(c.read_field(d, t, n, comps), want[n].reshape(-1) if comps == 1 else want[n])
    t = times[-1]
    idx = len(times) - 1
    hdr, part, v = E.read_var(os.path.join(d, "lattice.U_%d" % idx), 3 * c.N)
    assert (hdr, part) == ("U", 1)
    np.testing.assert_array_equal(v.reshape(3, c.N).T, saved[t]["U"])
    hdr, part, v = E.read_var(os.path.join(d, "lattice.f%d_%d" % (Q - 1, idx)), c.N)
    assert hdr == "f"
    np.testing.assert_array_equal(v, saved[t]["f"][:, Q - 1])
    # a fresh folder: no case file, so the list is [0] and the files carry index 0
    d2 = make_case_dir(spec)
    c.write_fields(d2)
    assert E.case_times(os.path.join(d2, "lbm.case")) == [0]
    np.testing.assert_array_equal(c.read_field(d2, t, "rho", 1), saved[t]["rho"])
    assert os.path.exists(os.path.join(d2, "lattice.g0_0"))

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

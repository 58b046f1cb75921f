"""Independent reader of binary EnSight Gold files (the layout of the reference's
IO/ensight writers: 80-byte NUL-padded text records, int32 and float32 numbers) and of
the time list of an EnSight case file, for checking the library's ensight output."""
import re

import numpy as np


class _Buf:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def text(self):
        s = self.b[self.o:self.o + 80]
        self.o += 80
        assert len(s) == 80
        return s.split(b"\0", 1)[0].decode()

    def ints(self, n):
        v = np.frombuffer(self.b, np.int32, n, self.o)
        self.o += 4 * n
        return v

    def floats(self, n):
        v = np.frombuffer(self.b, np.float32, n, self.o)
        self.o += 4 * n
        return v

    def done(self):
        return self.o == len(self.b)


def read_geo(path):
    r = _Buf(path)
    head = [r.text() for _ in range(5)]
    assert r.text() == "part"
    part = int(r.ints(1)[0])
    desc, coord = r.text(), r.text()
    n = int(r.ints(1)[0])
    xyz = r.floats(3 * n).reshape(3, n)
    etype = r.text()
    ncells = int(r.ints(1)[0])
    npe = {"quad4": 4, "hexa8": 8}[etype]
    conn = r.ints(ncells * npe).reshape(ncells, npe)
    assert r.done()
    return dict(head=head, part=part, desc=desc, coord=coord, n=n, xyz=xyz, etype=etype, conn=conn)


def read_var(path, nvalues):
    """(header name, part, values) of one variable file."""
    r = _Buf(path)
    name, part_kw = r.text(), r.text()
    part = int(r.ints(1)[0])
    assert part_kw == "part" and r.text() == "coordinates"
    v = r.floats(nvalues)
    assert r.done()
    return name, part, v


def write_var(path, name, values):
    """One variable file as the reference's writers lay it out."""
    def rec(s):
        return s.encode().ljust(80, b"\0")
    with open(path, "wb") as fh:
        fh.write(rec(name) + rec("part") + np.int32(1).tobytes() + rec("coordinates"))
        fh.write(np.asarray(values, np.float32).tobytes())


def case_times(path):
    txt = open(path).read()
    ns = int(re.search(r"number of steps:\s+(\d+)", txt).group(1))
    vals = txt.split("time values:", 1)[1].split()
    return [int(v) for v in vals[:ns]]

"""FP64 parity at the benchmark configurations' full sizes: the library against the CPU
oracle on the 256^3 D3Q15 bubble and the 128 x 128 x 384 thermal column (gravity,
bounceBack floor and ceiling, fixedRandomT floor, fixedT ceiling).  By default each runs
40 steps (write steps included), which exercises every kernel at the configuration's
size in well under a minute of oracle time on the host cores; MPLB_FULL_PARITY_STEPS
sets a longer run (1000 and 750 steps were run this way: bit-identical, every field):

    MPLB_FULL_PARITY_STEPS=1000 python -m pytest tests/test_gpu_parity_full.py -s -q
"""
import os

import numpy as np
import pytest

import multiphaselb_amd as M
from _cases import max_rel, pair, sphere_fields
from multiphaselb_amd.casefiles import BC
from test_gpu_parity import TC, TOL, _maxwell, d3

pytestmark = pytest.mark.gpu

FULL = {
    "c1_bubble_256": (d3(nx=256, ny=256, nz=256, btype=5, write_interval=20),
                      dict(T=0.8 * TC, r_in=_maxwell(0.8)[0], r_out=_maxwell(0.8)[1], radius=40.0), False, 1000),
    "c2_column_128x128x384": (d3(nx=128, ny=128, nz=384, btype=3, g=(0.0, 0.0, -1e-5), write_interval=20,
                                 f_bc={"Z0": BC("bounceBack"), "Z1": BC("bounceBack")},
                                 g_bc={"Z0": BC("fixedRandomT", (0.95 * TC,), pert=0.5, duration=10 ** 6),
                                       "Z1": BC("fixedT", (0.8 * TC,))}),
                              dict(T=0.8 * TC, r_in=_maxwell(0.8)[1], r_out=_maxwell(0.8)[0], radius=0, column=128.0),
                              True, 750),
}


@pytest.mark.parametrize("name", list(FULL))
def test_full_size_configuration_matches_oracle(name):
    spec, fk, ht, long_steps = FULL[name]
    steps = min(long_steps, int(os.environ.get("MPLB_FULL_PARITY_STEPS", "40")))
    print("case", name, "N =", spec.nx * spec.ny * spec.nz, flush=True)
    fields = sphere_fields(spec, seed=11, **fk)
    c, o = pair(spec, fields, precision=M.FP64, ht=ht)
    done = 0
    while done < steps:
        n = min(50, steps - done)
        assert c.step(n) == n and o.run(n) == n
        done += n
        print("  %d / %d steps" % (done, steps), flush=True)
    assert o.error()[0] == 0, o.error()
    for f in ("rho", "U", "T", "f", "p") + (("g",) if ht else ()):
        a, b = c.get(f), o.get(f)
        r = max_rel(a, b)
        print("  %-3s max_rel %.3e exact %s" % (f, r, bool(np.array_equal(a, b))), flush=True)
        assert np.isfinite(a).all() and r <= TOL, (f, r)

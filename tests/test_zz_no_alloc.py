"""The allocation-free radial raster (mplb_set_radial_density_noalloc) against the
allocating entry point (mplb_set_radial_density), through the C ABI: two calls in a row,
the second on a smaller case with a shorter profile, both bit-identical in density and
mass; sizes beyond the scratch limits are refused with a message."""
import numpy as np
import pytest

import multiphaselb_amd as M
from multiphaselb_amd import warmstart as W
from multiphaselb_amd.casefiles import CaseSpec
from _cases import make_case_dir

pytestmark = pytest.mark.gpu

LAM15 = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.2)


def _case(n, precision, fused):
    spec = CaseSpec(lbmodel=1, nx=n, ny=n, nz=n, btype=5, f_model=2, f_Lambda=LAM15, g_model=None,
                    write_interval=1000)
    return M.Case(make_case_dir(spec), precision=precision, heat_transfer=False, fused=fused,
                  extensions=M.EXT_LIMRT_D3Q15)


def _profile(npts, width):
    d = np.linspace(-8.0, 8.0, npts)
    return 0.09 - 0.07 * np.tanh(2.0 * d / width), d[0], d[1] - d[0]


@pytest.mark.parametrize("precision,fused", [(M.FP64, False), (M.FP32, True)])
def test_noalloc_raster_matches_allocating_entry_point(precision, fused):
    calls = [(40, 301, 4.0, 9.3), (18, 37, 3.0, 4.1)]    # second call: smaller case, shorter profile
    for n, npts, width, R in calls:
        print("noalloc raster: n=%d nprof=%d precision=%d fused=%d" % (n, npts, precision, fused), flush=True)
        prof, d0, dd = _profile(npts, width)
        cen = ((n - 1) / 2.0 + 0.25, (n - 1) / 2.0 - 0.5, (n - 1) / 2.0)
        a, b = _case(n, precision, fused), _case(n, precision, fused)
        ma = a.set_radial_density(prof, d0, dd, cen, R)
        mb = b.set_radial_density(prof, d0, dd, cen, R, noalloc=True)
        ra, rb = a.get("rho"), b.get("rho")
        assert np.array_equal(ra, rb)
        assert ma == mb
        # the raster itself: host interpolation of the same profile
        r = W.node_radius((n, n, n), cen)
        host = np.interp(r - R, d0 + dd * np.arange(npts), prof)
        if precision == M.FP32:
            host = host.astype(np.float32).astype(np.float64)
        np.testing.assert_allclose(rb, host, rtol=1e-13, atol=0)
        sites = W.distinct_sites((n, n, n))
        assert abs(mb - host[sites].sum()) <= 1e-12 * host.sum()
        a.close()
        b.close()


def test_noalloc_raster_refuses_a_profile_beyond_its_scratch():
    c = _case(12, M.FP64, False)
    prof, d0, dd = _profile(4097, 3.0)
    with pytest.raises(M.MplbError) as e:
        c.set_radial_density(prof, d0, dd, (5.5, 5.5, 5.5), 3.0, noalloc=True)
    assert "exceeds the scratch limit" in e.value.msg
    prof, d0, dd = _profile(4096, 3.0)
    c.set_radial_density(prof, d0, dd, (5.5, 5.5, 5.5), 3.0, noalloc=True)
    c.close()

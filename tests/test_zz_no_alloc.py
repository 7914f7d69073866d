"""frt_otf_averages_noalloc (the allocation-free variant of
frt_otf_averages, for callers that cannot allocate) through the C ABI: two
calls in a row, the second on a smaller grid with fewer rays, each equal bit
for bit to frt_otf_averages on the same state (same kernels; the partials of
the first call left in the shared device array must not leak into the
second).  Also: refused while rays removed by find_avg_unu await compaction."""
import ctypes as C

import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from flexrt_imaging_reion_amd._abi import FrtOtf

pytestmark = pytest.mark.gpu

FRT_E_STATE = 5


def _print_case(*a):
    print("CASE", *a, flush=True)


def _both(rt, avg_rho):
    lib = rt._lib
    a, b = FrtOtf(), FrtOtf()
    rc = lib.frt_otf_averages_noalloc(rt._h, C.c_float(avg_rho), C.c_float(100.0), C.byref(a))
    assert rc == 0, lib.frt_last_error(rt._h)
    rc = lib.frt_otf_averages(rt._h, C.c_float(avg_rho), C.c_float(100.0), C.byref(b))
    assert rc == 0, lib.frt_last_error(rt._h)
    return bytes(a), bytes(b), a.as_dict()


def test_noalloc_otf_matches_allocating_entry_point():
    from flexrt_imaging_reion_amd.rt import FlexRT

    for n, n_src, steps in ((32, 24, 4), (16, 8, 2)):  # the second call is the smaller one
        _print_case("otf noalloc lognormal n=%d src=%d steps=%d" % (n, n_src, steps))
        wl = inputs.make_workload("lognormal64", n=n, n_src=n_src, steps=steps)
        with FlexRT(dict(wl["cfg"], precision=1)) as rt:
            inputs.setup(rt, wl)
            for _ in range(steps):
                rt.step()
            avg_rho = float(np.float32(rt.download_field("rho").astype(np.float64).mean()))
            na, al, d = _both(rt, avg_rho)
            assert na == al
            assert d["npht"] > 0 and 0.0 < d["fH1"][0] < 1.0


def test_noalloc_otf_refuses_pending_removals():
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("otf noalloc pending removals")
    wl = inputs.make_workload("lognormal64", n=16, n_src=8, steps=2)
    with FlexRT(dict(wl["cfg"], precision=0)) as rt:
        inputs.setup(rt, wl)
        rt.step()
        rt.update_rays()
        rt.find_avg_unu()
        o = FrtOtf()
        rc = rt._lib.frt_otf_averages_noalloc(rt._h, C.c_float(1e-27), C.c_float(100.0), C.byref(o))
        assert rc == FRT_E_STATE
        rt.split_rays()
        rc = rt._lib.frt_otf_averages_noalloc(rt._h, C.c_float(1e-27), C.c_float(100.0), C.byref(o))
        assert rc == 0

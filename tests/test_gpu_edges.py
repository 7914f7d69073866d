"""GPU edge cases against the CPU checker, through the C ABI.

The FP64 form is held bit-exact, as in test_gpu_parity.py, on configurations
the main parity runs do not reach: grids that are not a multiple of the
4x4x4 tile (the library then keeps the reference's row-major cell numbering
inside), one and eight frequency bins (FRT_MAX_FREQ), and a single C2Ray
pass (the cell kernel's first and last pass in one launch).  The error paths
are forced from the host with configurations that make the condition occur
(a ray buffer too small, a path_length too short), so every kernel still
exits normally; both the library and the checker must report them.
"""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from parity_util import STATE_FIELDS, bits_equal, compare_rays, counts, fields, make_pair, max_rel

pytestmark = pytest.mark.gpu

FRT_E_CAPACITY = 3
FRT_E_PATH_OVERFLOW = 4


def _print_case(*a):
    print("CASE", *a, flush=True)


def fp64_bit_exact(wl, steps):
    g, o = make_pair(wl, precision=0)
    try:
        for s in range(steps):
            sg = g.step()
            so = o.step()
            assert counts(sg) == counts(so), "step %d" % s
            assert sg["nphot_merge"] == so["nphot_merge"]
        err = compare_rays(g.rays(), o.rays())
        assert err is None, err
        fg, fo = fields(g), fields(o)
        bad = [k for k in STATE_FIELDS if not bits_equal(fg[k], fo[k])]
        assert not bad, {k: max_rel(fg[k], fo[k]) for k in bad}
    finally:
        g.close()
        o.close()


def with_bins(wl, nfreq):
    wl["cfg"]["nfreq"] = nfreq
    wl["nu_phot"], wl["spectrum"] = inputs.powerlaw_bins(nbins=nfreq)
    return wl


@pytest.mark.parametrize("n", [18, 21])
def test_fp64_grid_not_tiled(n):
    _print_case("fp64 lognormal n=%d src=8 steps=5 (row-major cells)" % n)
    wl = inputs.make_workload("lognormal64", n=n, n_src=8, steps=5)
    fp64_bit_exact(wl, 5)


@pytest.mark.parametrize("nfreq", [1, 8])
def test_fp64_frequency_bins(nfreq):
    _print_case("fp64 lognormal n=16 src=16 steps=5 nfreq=%d" % nfreq)
    wl = with_bins(inputs.make_workload("lognormal64", n=16, n_src=16, steps=5), nfreq)
    fp64_bit_exact(wl, 5)


def test_fp64_single_c2ray_pass():
    _print_case("fp64 lognormal n=16 src=16 steps=5 itercount=1")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=5)
    wl["cfg"]["itercount"] = 1
    fp64_bit_exact(wl, 5)


def test_mixed_eight_bins_single_pass():
    """Mixed form with the widest photon record (8 bins) and a single pass:
    the ionized volume and the typical cell agree with the FP64 checker."""
    _print_case("mixed lognormal n=16 src=16 steps=4 nfreq=8 itercount=1")
    wl = with_bins(inputs.make_workload("lognormal64", n=16, n_src=16, steps=4), 8)
    wl["cfg"]["itercount"] = 1
    g, o = make_pair(wl, precision=1)
    try:
        for _ in range(4):
            g.step()
            o.step()
        xg = g.download_field("nH1").astype(np.float64) / g.download_field("nH")
        xo = o.download_field("nH1").astype(np.float64) / o.download_field("nH")
        assert np.all(np.isfinite(xg))
        rel = np.abs(xg - xo) / np.abs(xo)
        assert np.median(rel) < 1e-5
        assert np.mean(rel > 1e-3) < 0.01
        assert abs(float((xg < 0.5).mean()) - float((xo < 0.5).mean())) < 1e-3
    finally:
        g.close()
        o.close()


def test_ray_capacity_error():
    """A ray buffer smaller than the rays a step produces: the split stops
    with FRT_E_CAPACITY before writing past the buffer."""
    from flexrt_imaging_reion_amd.rt import FlexRT, FlexRTError

    _print_case("capacity error lognormal n=16 src=16 capacity=300")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6)
    cfg = dict(wl["cfg"], ray_capacity=300)  # 192 cast rays fit; a step needs room for ~400
    with FlexRT(cfg) as g:
        with pytest.raises(FlexRTError) as ei:
            inputs.setup(g, wl)
            for _ in range(6):
                g.step()
        assert ei.value.code == FRT_E_CAPACITY


def test_path_overflow_error():
    """path_length 1 while rays cross cell faces within a step: the step in
    which the first ray needs two segments fails with FRT_E_PATH_OVERFLOW in
    the library and in the checker alike."""
    from flexrt_imaging_reion_amd.rt import FlexRT, FlexRTError
    from oracle.pyoracle import Oracle

    _print_case("path overflow lognormal n=16 src=16 path_length=1")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=4)
    cfg = dict(wl["cfg"], path_length=1)

    def failing_step(rt, exc):
        inputs.setup(rt, wl)
        for s in range(4):
            try:
                rt.step()
            except exc as e:
                return s, e
        return None, None

    with FlexRT(cfg) as g:
        sg, eg = failing_step(g, FlexRTError)
    o = Oracle(cfg)
    try:
        so, _ = failing_step(o, RuntimeError)
    finally:
        o.close()
    assert sg is not None and eg.code == FRT_E_PATH_OVERFLOW
    assert sg == so

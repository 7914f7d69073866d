"""GPU parity of the sub-grid I-front model (cfg["subgrid"] = 1, the
reference's production setting subgrid = ifront = equilibrium = TRUE) against
the CPU checker, FP64 form, bit for bit: every cell field, the per-cell
I-front state (fion, fion_max, tcross and its crossing bins, treion_out,
Delta_fion, the RadHydro nodes), the ray lists and the step counters.

The workload is ``subgrid32`` (32^3, 5 cMpc/h, z = 8, four sources of 1e53 -
1e54 /s, one bin at 19 eV) with placeholder RadHydro tables; the runs are
long enough that fronts cross many cells, partly ionized cells hold rays
back through the first-come count of their neutral part, and the crossing
bins of get_gamma_interp are in use."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from parity_util import STATE_FIELDS, bits_equal, compare_rays, counts, fields, make_pair

pytestmark = pytest.mark.gpu

SUBGRID_FIELDS = STATE_FIELDS + ["avg_gamma_prev"]


def _compare_state(g, o):
    fg, fo = fields(g), fields(o)
    for k in STATE_FIELDS:
        assert bits_equal(fg[k], fo[k]), "field %s" % k
    for k in ("avg_gamma_prev",):
        assert bits_equal(g.download_field(k), o.download_field(k)), "field %s" % k
    sg, so = g.download_subgrid_state(), o.download_subgrid_state()
    for k in so:
        assert bits_equal(sg[k], so[k]), "subgrid state %s" % k
    err = compare_rays(g.rays(), o.rays())
    assert err is None, err
    return so


@pytest.mark.parametrize("n,steps", [(16, 8), (24, 10), (32, 10)])
def test_subgrid_steps_bit_exact(n, steps):
    print("CASE subgrid n=%d steps=%d" % (n, steps), flush=True)
    wl = inputs.make_workload("subgrid32", n=n, steps=steps)
    g, o = make_pair(wl)
    try:
        partial_seen = crossed_bins = 0
        for k in range(steps):
            sg, so = g.step(), o.step()
            assert counts(sg) == counts(so), "step %d: %r != %r" % (k, counts(sg), counts(so))
            st = _compare_state(g, o)
            partial_seen = max(partial_seen, int(((st["fion"] > 0) & (st["fion"] < 1)).sum()))
            crossed_bins = max(crossed_bins, int(st["tcross_check"].sum()))
        # the model was exercised: partly ionized cells, crossing bins passed
        assert partial_seen > 0 and crossed_bins > 0
        assert int(o.download_field("treion_flag").sum()) > 0
    finally:
        g.close()


def test_subgrid_phases_bit_exact():
    """update_rays, find_avg_unu, merge, split and update_thermal called one
    at a time (the thermal update not fused into the last cell pass)."""
    print("CASE subgrid phases n=16", flush=True)
    wl = inputs.make_workload("subgrid32", n=16, steps=6)
    g, o = make_pair(wl)
    try:
        dt = float(np.float32(wl["cfg"]["dt"]))
        for k in range(6):
            for rt in (g, o):
                rt.set_time(rt.get_time() + dt)
                rt.update_rays()
                rt.find_avg_unu()
                rt.merge_rays()
                rt.split_rays()
                rt.update_thermal()
                rt.new_rays()
            _compare_state(g, o)
    finally:
        g.close()


def test_subgrid_refusals():
    """Combinations not built are refused: the mixed form, the adjoint, device
    checkpoints, and a step before frt_set_subgrid."""
    from flexrt_imaging_reion_amd.rt import FlexRT

    print("CASE subgrid refusals", flush=True)
    wl = inputs.make_workload("subgrid32", n=8, steps=1)
    with pytest.raises(Exception):
        FlexRT(dict(wl["cfg"], precision=1))
    with FlexRT(dict(wl["cfg"])) as g:
        wl2 = dict(wl)
        wl2.pop("subgrid")
        inputs.setup(g, wl2)
        with pytest.raises(RuntimeError):
            g.step()
    with FlexRT(dict(wl["cfg"])) as g:
        inputs.setup(g, wl)
        with pytest.raises(RuntimeError):
            g.adj_enable(True)
        with pytest.raises(RuntimeError):
            g.checkpoint()

"""Step graph and launch tuning: frt_step replayed from a captured HIP graph,
and the work-queue (persistent grid) march and deposit, give the same bits
as the plain launches -- every cell field, the ray list and the step
counters -- in both precision forms, through merging and splitting; and the
FP64 form in graph mode stays bit-exact with the CPU checker."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from parity_util import STATE_FIELDS, bits_equal, compare_rays, counts, fields, make_pair

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("CASE", *a, flush=True)


def _run(wl, precision, graph=False, **tuning):
    from flexrt_imaging_reion_amd.rt import FlexRT

    rt = FlexRT(dict(wl["cfg"], precision=precision))
    inputs.setup(rt, wl)
    if tuning:
        rt.set_tuning(**tuning)
    if graph:
        rt.set_step_graph(True)
    st = [rt.step() for _ in range(wl["steps"])]
    out = (st, rt.rays(), fields(rt))
    rt.close()
    return out


def _same(a, b):
    sa, ra, fa = a
    sb, rb, fb = b
    for s, (x, y) in enumerate(zip(sa, sb)):
        assert counts(x) == counts(y), "step %d" % s
        assert x["nphot_merge"] == y["nphot_merge"] and x["t"] == y["t"], "step %d" % s
    err = compare_rays(ra, rb)
    assert err is None, err
    bad = [k for k in STATE_FIELDS if not bits_equal(fa[k], fb[k])]
    assert not bad, bad


def _merging_workload():
    wl = inputs.make_workload("lognormal64", n=16, n_src=64, steps=12)
    wl["cfg"]["ray_per_cell_max"] = 12  # merges from the first step
    return wl


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_same_bits(precision):
    _print_case("graph vs plain precision=%d lognormal n=16 src=64 rpc=12 steps=12" % precision)
    wl = _merging_workload()
    plain = _run(wl, precision)
    assert sum(s["rays_merged_away"] for s in plain[0]) > 0
    _same(plain, _run(wl, precision, graph=True))


@pytest.mark.parametrize("precision", [0, 1])
def test_queue_kernels_same_bits(precision):
    _print_case("queue kernels precision=%d lognormal n=32 src=24 steps=8" % precision)
    wl = inputs.make_workload("lognormal64", n=32, n_src=24, steps=8)
    plain = _run(wl, precision)
    assert sum(s["rays_split"] for s in plain[0]) > 0
    _same(plain, _run(wl, precision, deposit_queue=1, march_queue=1, queue_chunk=128, blocks_per_cu=2,
                      fp64_chunk=2048))
    _same(plain, _run(wl, precision, graph=True, deposit_queue=1, march_queue=1, grid_cap=64))


def test_graph_fp64_matches_checker():
    _print_case("graph fp64 vs checker lognormal n=16 src=64 rpc=13 steps=14")
    wl = inputs.make_workload("lognormal64", n=16, n_src=64, steps=14)
    wl["cfg"]["ray_per_cell_max"] = 13
    g, o = make_pair(wl, precision=0)
    g.set_step_graph(True)
    for s in range(wl["steps"]):
        assert counts(g.step()) == counts(o.step()), "step %d" % s
    err = compare_rays(g.rays(), o.rays())
    assert err is None, err
    fg, fo = fields(g), fields(o)
    bad = [k for k in STATE_FIELDS if not bits_equal(fg[k], fo[k])]
    assert not bad, bad


def test_graph_recaptures_after_host_changes():
    """Host-side changes between replays (time, sources' rates, a checkpoint
    restore) take effect: graph and plain runs with the same changes agree."""
    _print_case("graph recapture lognormal n=16 src=16 steps=6")
    from flexrt_imaging_reion_amd.rt import FlexRT

    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6)
    res = []
    for graph in (False, True):
        rt = FlexRT(wl["cfg"])
        inputs.setup(rt, wl)
        rt.set_step_graph(graph)
        rt.step()
        ck = rt.checkpoint()
        rt.step()
        rt.step()
        rt.restore(ck)
        ck.free()
        rt.step()
        rt.set_time(rt.get_time() + 1e10)
        rt.step()
        res.append((rt.rays(), fields(rt), rt.get_time()))
        rt.close()
    assert res[0][2] == res[1][2]
    assert compare_rays(res[0][0], res[1][0]) is None
    assert all(bits_equal(res[0][1][k], res[1][1][k]) for k in STATE_FIELDS)


def test_tuning_rejects_bad_values():
    from flexrt_imaging_reion_amd.rt import FlexRT, FlexRTError

    wl = inputs.make_workload("uniform32", n=16)
    with FlexRT(wl["cfg"]) as rt:
        t = rt.tuning()
        assert t["num_cu"] > 0 and t["fp64_chunk"] == 4096
        for bad in (dict(queue_chunk=100), dict(fp64_chunk=1000), dict(deposit_queue=2), dict(blocks_per_cu=0)):
            with pytest.raises(FlexRTError):
                rt.set_tuning(**bad)

"""GPU parity of the reference's c2ray_iter = FALSE mode (cfg["c2ray_off"]):
the march deposits into u_nu segment by segment and attenuates the rays by
the HI opacity (ray_tracing.cc:1329-1355,1443-1470), then update_gamma and
update_chem's backward-difference solve (gas_funcs.cc:74-99,243-351) run
after the split.  FP64 form against the CPU checker, bit for bit: every cell
field, the solve's state (nHe3, the *_prev densities, recomb_He3), u_nu, the
ray lists and the step counters -- through splitting and merging, with the
phases called one at a time, and from a captured step graph.

In this mode rays are never removed for a low photon count, so rays whose
photons are all absorbed live on with zero photons, and merging groups of
them divides 0 by 0 (ray_tracing.cc:425-432): their r, dr, position and time
become NaN, as in the reference.  x86 and the GPU write different NaN bit
patterns (sign), so values are compared bit for bit except that a NaN must
meet a NaN."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from parity_util import STATE_FIELDS, bits_equal, counts, fields, make_pair

pytestmark = pytest.mark.gpu

EXTRA = ["nHe3", "nH2_prev", "nHe1_prev", "nHe2_prev", "nHe3_prev", "recomb_He3", "u_nu"]


def _print_case(*a):
    print("CASE", *a, flush=True)


def _same(a, b):
    """Bit for bit, except that any NaN matches any NaN."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind != "f":
        return bits_equal(a, b)
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb)) and bits_equal(a[~na], b[~nb])


def _compare(g, o, where):
    fg, fo = fields(g), fields(o)
    bad = [k for k in STATE_FIELDS if not _same(fg[k], fo[k])]
    bad += [k for k in EXTRA if not _same(g.download_field(k), o.download_field(k))]
    assert not bad, "%s: %s" % (where, bad)
    rg, ro = g.rays(), o.rays()
    bad = [k for k in ro if not _same(rg[k], ro[k])]
    assert not bad, "%s: ray fields %s" % (where, bad)


@pytest.mark.parametrize("name,n,n_src,steps,rpc", [
    ("uniform32", None, None, 6, None),
    ("lognormal64", 16, 64, 12, 12),
    ("lognormal64", 32, 24, 8, None),
])
def test_c2ray_off_steps_bit_exact(name, n, n_src, steps, rpc):
    _print_case("c2ray_off %s n=%s src=%s steps=%d rpc=%s" % (name, n, n_src, steps, rpc))
    wl = inputs.make_workload(name, n=n, n_src=n_src, steps=steps, c2ray_off=True)
    if rpc:
        wl["cfg"]["ray_per_cell_max"] = rpc  # merges from the first steps
    g, o = make_pair(wl)
    try:
        tot = {"rays_split": 0, "rays_merged_away": 0}
        for k in range(steps):
            sg, so = g.step(), o.step()
            assert counts(sg) == counts(so), "step %d: %r != %r" % (k, counts(sg), counts(so))
            for key in tot:
                tot[key] += sg[key]
            _compare(g, o, "step %d" % k)
        f_h1 = o.download_field("f_H1")
        assert f_h1.min() < 0.5  # fronts were driven
        assert np.all(o.download_field("nHe3") == 0.0)  # no photon reaches 54.4 eV
        if rpc:
            assert tot["rays_merged_away"] > 0
        if name == "lognormal64" and n == 32:
            assert tot["rays_split"] > 0
    finally:
        g.close()


def test_c2ray_off_phases_bit_exact():
    """update_rays (march and deposit), find_avg_unu (set_to_true), merge,
    split, update_thermal (update_gamma, update_chem, update_thermal) and
    new_rays called one at a time."""
    _print_case("c2ray_off phases lognormal n=16 src=16 steps=5")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=5, c2ray_off=True)
    g, o = make_pair(wl)
    try:
        dt = float(np.float32(wl["cfg"]["dt"]))
        for k in range(5):
            for rt in (g, o):
                rt.set_time(rt.get_time() + dt)
                rt.update_rays()
            assert bits_equal(g.download_field("u_nu"), o.download_field("u_nu")), "u_nu step %d" % k
            for rt in (g, o):
                rt.find_avg_unu()
                rt.merge_rays()
                rt.split_rays()
                rt.update_thermal()
                rt.new_rays()
            _compare(g, o, "step %d" % k)
    finally:
        g.close()


def test_c2ray_off_graph_same_bits():
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("c2ray_off graph lognormal n=16 src=64 rpc=12 steps=8")
    wl = inputs.make_workload("lognormal64", n=16, n_src=64, steps=8, c2ray_off=True)
    wl["cfg"]["ray_per_cell_max"] = 12
    out = []
    for graph in (False, True):
        rt = FlexRT(dict(wl["cfg"]))
        inputs.setup(rt, wl)
        if graph:
            rt.set_step_graph(True)
        st = [counts(rt.step()) for _ in range(wl["steps"])]
        out.append((st, rt.rays(), fields(rt), rt.download_field("u_nu")))
        rt.close()
    (sa, ra, fa, ua), (sb, rb, fb, ub) = out
    assert sa == sb
    assert not [k for k in ra if not _same(ra[k], rb[k])]
    assert not [k for k in STATE_FIELDS if not _same(fa[k], fb[k])]
    assert _same(ua, ub)


def test_c2ray_off_refusals():
    """Combinations not built are refused: the mixed form, the sub-grid model
    and the adjoint."""
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("c2ray_off refusals")
    wl = inputs.make_workload("lognormal64", n=8, n_src=4, steps=1, c2ray_off=True)
    with pytest.raises(Exception):
        FlexRT(dict(wl["cfg"], precision=1))
    with pytest.raises(Exception):
        FlexRT(dict(wl["cfg"], subgrid=1))
    with FlexRT(dict(wl["cfg"])) as g:
        inputs.setup(g, wl)
        with pytest.raises(RuntimeError):
            g.adj_enable(True)


def test_c2ray_off_helium_iii_bit_exact():
    """Bins at 20-80 eV: photons above 54.4 eV give Gamma_HeII, so every
    helium branch of solve_ion (gas_funcs.cc:243-296) runs -- He I, He II or
    He III the most abundant at the start of the step, set per cell -- and
    He III is created and recombines.  FP64 form against the checker, bit
    for bit, over six steps."""
    _print_case("c2ray_off helium III lognormal n=16 src=16 steps=6")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6, c2ray_off=True)
    wl["nu_phot"] = np.array([20.0, 30.0, 45.0, 60.0, 80.0]) / 4.135667e-15
    wl["spectrum"] = np.full(5, 0.2, np.float32)
    g, o = make_pair(wl)
    try:
        n3 = 16 ** 3
        rng = np.random.default_rng(11)
        nhe = o.download_field("nHe").astype(np.float64)
        frac = rng.dirichlet([0.3, 0.3, 0.3], size=n3)  # a dominant level per cell
        for j, k in enumerate(("nHe1", "nHe2", "nHe3")):
            v = (nhe * frac[:, j]).astype(np.float32)
            for rt in (g, o):
                rt.upload_field(k, v)
                rt.upload_field(k + "_prev", v)
        top = np.argmax(frac, axis=1)
        assert all((top == j).sum() > 100 for j in range(3))
        for k in range(6):
            sg, so = g.step(), o.step()
            assert counts(sg) == counts(so), "step %d: %r != %r" % (k, counts(sg), counts(so))
            _compare(g, o, "step %d" % k)
        assert float(o.download_field("f_He3").max()) > 0.5
    finally:
        g.close()

"""update_cosmo_expansion on the GPU against the CPU checker, through the C
ABI, in the driver's order (main.cc:129-216: the RT step, then
update_cosmo_expansion, then update_dt).  The expansion is exaggerated to 1%
per step so that the rescaled densities, the moved ray positions and the
face clamps change what the following steps do."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from flexrt_imaging_reion_amd.rt import FlexRT
from parity_util import STATE_FIELDS, bits_equal, compare_rays, counts, fields, make_pair, max_rel

pytestmark = pytest.mark.gpu

F32 = np.float32
FRT_E_ARG = 1


def _print_case(*a):
    print("CASE", *a, flush=True)


def drive(rt, s, a0, dt0, growth):
    """One pass of the driver loop after the step: scale factors one dt
    apart, and the time step following the cell size (update_dt)."""
    a_prev, a_cur = F32(a0 * growth ** s), F32(a0 * growth ** (s + 1))
    rt.update_cosmo_expansion(a_cur, a_prev)
    rt.set_dt(F32(dt0 * growth ** (s + 1)))


def test_fp64_with_expansion_bit_exact():
    _print_case("fp64 expansion lognormal n=16 src=16 steps=6 growth=1.01")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6)
    g, o = make_pair(wl, precision=0)
    a0, dt0 = float(wl["cfg"]["a_rt"]), float(wl["cfg"]["dt"])
    dx0 = g.cell_size()
    try:
        for s in range(6):
            sg, so = g.step(), o.step()
            assert counts(sg) == counts(so), "step %d" % s
            assert sg["nphot_merge"] == so["nphot_merge"]
            drive(g, s, a0, dt0, 1.01)
            drive(o, s, a0, dt0, 1.01)
            assert F32(g.cell_size()) == F32(o.cell_size())
        err = compare_rays(g.rays(), o.rays())
        assert err is None, err
        fg, fo = fields(g), fields(o)
        bad = [k for k in STATE_FIELDS if not bits_equal(fg[k], fo[k])]
        assert not bad, {k: max_rel(fg[k], fo[k]) for k in bad}
        assert abs(g.cell_size() / dx0 / 1.01 ** 6 - 1) < 1e-5
    finally:
        g.close()
        o.close()


def test_mixed_with_expansion():
    _print_case("mixed expansion lognormal n=16 src=16 steps=6 growth=1.01")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6)
    g, o = make_pair(wl, precision=1)
    a0, dt0 = float(wl["cfg"]["a_rt"]), float(wl["cfg"]["dt"])
    try:
        for s in range(6):
            g.step()
            o.step()
            drive(g, s, a0, dt0, 1.01)
            drive(o, s, a0, dt0, 1.01)
        assert bits_equal(g.download_field("rho"), o.download_field("rho"))
        xg = g.download_field("nH1").astype(np.float64) / g.download_field("nH")
        xo = o.download_field("nH1").astype(np.float64) / o.download_field("nH")
        rel = np.abs(xg - xo) / np.abs(xo)
        assert np.median(rel) < 1e-5
        assert np.mean(rel > 1e-3) < 0.01
        assert abs(float((xg < 0.5).mean()) - float((xo < 0.5).mean())) < 1e-3
    finally:
        g.close()
        o.close()


def test_expansion_needs_rho():
    from flexrt_imaging_reion_amd.rt import FlexRT, FlexRTError

    _print_case("expansion without rho")
    wl = inputs.make_workload("lognormal64", n=16, n_src=8, steps=1)
    with FlexRT(wl["cfg"]) as g:
        with pytest.raises(FlexRTError) as ei:
            g.update_cosmo_expansion(0.13, 0.125)
        assert ei.value.code == FRT_E_ARG


def test_fp64_hydro_step_bit_exact():
    """Three steps, a new comoving density snapshot and redshift
    (update_hydro_step: rho_prev set, ddelta_dt still 0), two steps, a second
    snapshot with a changed density (ddelta_dt now nonzero and entering the
    temperature update), three more steps: bit for bit like the checker."""
    _print_case("fp64 hydro step lognormal n=16 src=16 3+2+3 steps, two reloads")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=8)
    g, o = make_pair(wl, precision=0)
    rng = np.random.default_rng(5)
    a = float(wl["cfg"]["a_rt"])
    try:
        for s in range(3):
            sg, so = g.step(), o.step()
            assert counts(sg) == counts(so), "step %d" % s
        step = 3
        for zz, a_new, nsteps in ((6.6, 1 / 7.6, 2), (6.5, 1 / 7.5, 3)):
            rho_c = (g.download_field("rho").astype(np.float64) * a ** 3
                     * rng.uniform(0.9, 1.1, 16 ** 3)).astype(F32)
            avg = F32((rho_c.astype(np.float64) * float(F32(a_new)) ** -3).mean())
            for rt in (g, o):
                rt.upload_field("rho", rho_c)
                rt.update_hydro_step(F32(zz), F32(a_new), avg, F32(2.9e14))
                rt.get_nphot_min_max()
            a = float(F32(a_new))
            assert F32(g.cell_size()) == F32(o.cell_size())
            for k in ("ddelta_dt", "rho_prev", "rho"):
                assert bits_equal(g.download_field(k), o.download_field(k)), k
            for _ in range(nsteps):
                sg, so = g.step(), o.step()
                assert counts(sg) == counts(so), "step %d" % step
                step += 1
        assert np.count_nonzero(g.download_field("ddelta_dt")) > 0.9 * 16 ** 3
        err = compare_rays(g.rays(), o.rays())
        assert err is None, err
        fg, fo = fields(g), fields(o)
        bad = [k for k in STATE_FIELDS if not bits_equal(fg[k], fo[k])]
        assert not bad, {k: max_rel(fg[k], fo[k]) for k in bad}
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_otf_averages_match_checker(tmp_path, precision):
    """write_otf's averages reduced on the GPU against the checker's serial
    sums over the same state (for the mixed form the GPU state is copied into
    the checker first), so only the order of the double additions differs;
    and the text block written in the reference's format."""
    import math

    from flexrt_imaging_reion_amd._abi import FIELDS

    _print_case("otf averages precision=%d lognormal n=16 src=16 steps=4" % precision)
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=4)
    g, o = make_pair(wl, precision=precision)
    try:
        for _ in range(4):
            g.step()
            o.step()
        if precision:
            for k in FIELDS:
                o.upload_field(k, g.download_field(k))
            o.upload_rays(g.rays())
        for rt in (g, o):
            rt.upload_field("avg_gamma_prev", (rt.download_field("gamma_H1_tot") * F32(1.5)).astype(F32))
        avg_rho = F32(g.download_field("rho").astype(np.float64).mean())
        rg, ro = g.otf_averages(avg_rho), o.otf_averages(avg_rho)
        tol = 1e-6
        for k in rg:
            a = np.atleast_1d(np.asarray(rg[k], dtype=np.float64))
            b = np.atleast_1d(np.asarray(ro[k], dtype=np.float64))
            for x, y in zip(a, b):
                assert (math.isnan(x) and math.isnan(y)) or abs(x - y) <= tol * abs(y), (k, x, y)
        path = tmp_path / "otf.txt"
        g.write_otf(path, avg_rho)
        g.write_otf(path, avg_rho)
        text = path.read_text().split("\n")
        assert text.count("Global Averages") == 2
        assert text[0] == "Step Number    : 4"
        line = next(t for t in text if t.startswith("GH1"))
        got = [float(x) for x in line.split(":")[1].split()]
        assert got == [float("%e" % v) for v in rg["GH1"]]
    finally:
        g.close()
        o.close()


def test_gadget_coupling_bit_exact(tmp_path):
    """update_gadget_var after each of 3 FP64 steps, write_gadget_binary's
    averaged cubes and reset_gadget_var: bit for bit like the checker."""
    _print_case("fp64 gadget coupling lognormal n=16 src=16 steps=3")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=3)
    g, o = make_pair(wl, precision=0)
    try:
        for _ in range(3):
            g.step()
            o.step()
            g.update_gadget_var()
            o.update_gadget_var()
        assert g.gadget_step_time() == o.gadget_step_time()
        path = tmp_path / "small_gas_rt_z=7.000000"
        g.write_gadget_file(path)
        o.gadget_average()
        n3 = 16 ** 3
        cubes = np.frombuffer(path.read_bytes(), "<f4").reshape(2, n3)
        assert bits_equal(cubes[0], o.download_field("gadget_gamma"))
        assert bits_equal(cubes[1], o.download_field("gadget_heat"))
        assert bits_equal(g.download_field("gadget_gamma"), cubes[0])  # averaged in place
        g.reset_gadget_var()
        assert g.gadget_step_time() == 0.0 and not np.any(g.download_field("gadget_heat"))
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_dropin_order_matches_fused_step(precision):
    """The drop-in binding's order (main.cc:117-198 with the library's phases
    and the host's own update_thermal between them, flexrt_imaging_reion_amd
    .dropin): the host holds only what it pulls after find_avg_unu, runs the
    checker's update_thermal (the restatement of gas_funcs.cc:362-457) and
    pushes temp and the rates back.  Over 6 steps it matches a context that
    runs the fused frt_step, bit for bit, in counts, rays and every state
    field: the pulled list covers everything the host's thermal update reads."""
    from flexrt_imaging_reion_amd import dropin
    from oracle.pyoracle import Oracle

    _print_case("drop-in order precision=%d lognormal n=16 src=16 steps=6" % precision)
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6)
    cfg = dict(wl["cfg"], precision=precision)
    fused = FlexRT(cfg)
    lib = FlexRT(cfg)
    host = Oracle(cfg)
    try:
        for rt in (fused, lib, host):
            inputs.setup(rt, wl)
        dt = host.derived()["dt"]
        for s in range(6):
            sf = fused.step()
            dropin.step_dropin(lib, host, s, dt)
            assert lib.ray_count() == sf["rays_end"], "step %d" % s
        assert lib.get_time() == fused.get_time()
        err = compare_rays(lib.rays(), fused.rays())
        assert err is None, err
        fl, ff = fields(lib), fields(fused)
        bad = [k for k in STATE_FIELDS if not bits_equal(fl[k], ff[k])]
        assert not bad, {k: max_rel(fl[k], ff[k]) for k in bad}
        for k in ("temp", "heat_rate", "cool_rate", "IF_speed", "Inc_Flux"):
            assert bits_equal(host.download_field(k), fused.download_field(k)), k
    finally:
        fused.close()
        lib.close()
        host.close()

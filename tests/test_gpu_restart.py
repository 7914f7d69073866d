"""Restart files in the reference's formats (io_funcs.cc), through the C ABI.

The layouts are checked against the order and widths in which the reference's
readers consume the files (read_rays_binary io_funcs.cc:119-145,
read_grid_binary :81-116, read_treion :65-77), by parsing what the library
writes with numpy structured types built from those readers; then each file
is read back into a fresh context, and a run restarted from a ray file
continues bit for bit like the run that wrote it.
"""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from flexrt_imaging_reion_amd._abi import FIELDS
from parity_util import STATE_FIELDS, bits_equal, check_ray_file, compare_rays, counts, make_pair

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("CASE", *a, flush=True)


GAS_COLUMNS = ["rho", "temp", "nH1", "f_H1", "f_He1", "f_H2", "f_He2", "f_He3", "gamma_H1_tot",
               "gamma_He1_tot", "mfp", "mfp_912", "heat_rate", "cool_rate", "avg_nh1", "fion", "tcross",
               "treion_out"]  # read_grid_binary, io_funcs.cc:89-109, then ngam_nu[Nfreq]


def run(precision, steps, n=16, n_src=16):
    from flexrt_imaging_reion_amd.rt import FlexRT

    wl = inputs.make_workload("lognormal64", n=n, n_src=n_src, steps=steps)
    g = FlexRT(dict(wl["cfg"], precision=precision))
    inputs.setup(g, wl)
    for _ in range(steps):
        g.step()
    return wl, g


def test_ray_file_layout_and_round_trip(tmp_path):
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("ray file lognormal n=16 src=16 steps=4")
    wl, g = run(0, 4)
    path = tmp_path / "ray_z=07.000"
    try:
        g.write_ray_file(path)
        rays = g.rays()
        assert check_ray_file(path.read_bytes(), rays, int(wl["cfg"]["nfreq"])) is None
        with FlexRT(dict(wl["cfg"], precision=0)) as h:
            inputs.setup(h, wl)
            h.read_ray_file(path)
            assert compare_rays(h.rays(), rays) is None
    finally:
        g.close()


def test_ray_file_same_bytes_as_checker(tmp_path):
    """After the same FP64 steps, the library's ray file and the checker's
    restatement of write_rays_binary (one stream write per field) are the
    same bytes."""
    _print_case("ray file bytes vs checker lognormal n=16 src=16 steps=5")
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=5)
    g, o = make_pair(wl, precision=0)
    try:
        for _ in range(5):
            g.step()
            o.step()
        g.write_ray_file(tmp_path / "gpu")
        o.write_ray_file(tmp_path / "cpu")
        a, b = (tmp_path / "gpu").read_bytes(), (tmp_path / "cpu").read_bytes()
        assert len(a) == len(b) and a == b
    finally:
        g.close()
        o.close()


def test_gas_and_treion_files(tmp_path):
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("gas/treion files lognormal n=16 src=16 steps=4")
    wl, g = run(0, 4)
    n3 = int(wl["cfg"]["n"]) ** 3
    F = int(wl["cfg"]["nfreq"])
    rng = np.random.default_rng(4)
    extras = {"rho": rng.uniform(1e-31, 1e-29, n3).astype(np.float32),
              "mfp": rng.uniform(0, 1e24, n3).astype(np.float32)}
    flags = {"tcross_flag": (rng.uniform(size=n3) < 0.3).astype(np.uint8),
             "neutral_flag": (rng.uniform(size=n3) < 0.6).astype(np.uint8)}
    gas, tre = tmp_path / "gas_z=07.000", tmp_path / "tre_z=07.000"
    try:
        g.write_gas_file(gas, extras)
        g.write_treion_file(tre, flags)
        raw = np.frombuffer(gas.read_bytes(), "<f4")
        assert raw.size == 1 + n3 * (18 + F)
        assert raw[0] == np.float32(g.get_time())
        cols = raw[1:].reshape(n3, 18 + F)
        want = {k: g.download_field(k) for k in GAS_COLUMNS if k in FIELDS}
        for j, name in enumerate(GAS_COLUMNS):
            if name in want and name not in extras:
                assert bits_equal(cols[:, j], want[name]), name
        assert bits_equal(cols[:, 0], extras["rho"]) and bits_equal(cols[:, 10], extras["mfp"])
        assert np.all(cols[:, 7] == 0) and np.all(cols[:, 15] == 1) and np.all(cols[:, 16] == 0)
        assert bits_equal(cols[:, 11], g.download_field("mfp_912"))
        assert bits_equal(cols[:, 17], g.download_field("treion"))  # treion_out is set with treion
        assert bits_equal(cols[:, 18:], g.download_field("ngam_nu").reshape(n3, F))
        t = np.frombuffer(tre.read_bytes(), np.dtype([("treion", "<f4"), ("treion_flag", "u1"),
                                                      ("tcross_flag", "u1"), ("neutral_flag", "u1")]))
        assert t.size == n3
        assert bits_equal(t["treion"], g.download_field("treion"))
        assert np.array_equal(t["treion_flag"], (g.download_field("treion_flag") != 0).astype(np.uint8))
        assert np.array_equal(t["tcross_flag"], flags["tcross_flag"])
        assert np.array_equal(t["neutral_flag"], flags["neutral_flag"])
        # without extras: the held rho and mfp, neutral_flag TRUE (rt_funcs.cc:69)
        g.write_gas_file(gas)
        g.write_treion_file(tre)
        cols = np.frombuffer(gas.read_bytes(), "<f4")[1:].reshape(n3, 18 + F)
        assert bits_equal(cols[:, 0], g.download_field("rho"))
        assert bits_equal(cols[:, 10], g.download_field("mfp"))
        assert np.all(np.frombuffer(tre.read_bytes(), "u1").reshape(n3, 7)[:, 6] == 1)
        g.write_gas_file(gas, extras)
        g.write_treion_file(tre, flags)
        with FlexRT(dict(wl["cfg"], precision=0)) as h:
            inputs.setup(h, wl)
            got = h.read_gas_file(gas)
            fl = h.read_treion_file(tre)
            assert np.float32(h.get_time()) == np.float32(g.get_time())
            for name in list(want) + ["ngam_nu", "treion", "treion_flag"]:
                ref = extras[name] if name in extras else g.download_field(name)
                assert bits_equal(h.download_field(name), ref), name
            assert bits_equal(got["rho"], extras["rho"]) and np.all(got["fion"] == 1)
            assert np.array_equal(fl["neutral_flag"], flags["neutral_flag"])
    finally:
        g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_restart_from_ray_file_continues_bit_exact(tmp_path, precision):
    """A context given the writer's cell state and its ray file continues the
    run exactly as the writer does (the ray file holds the whole ray state)."""
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("restart precision=%d lognormal n=16 src=16 3+3 steps" % precision)
    wl, a = run(precision, 3)
    path = tmp_path / "ray_restart"
    try:
        a.write_ray_file(path)
        done = a.stats()["step"]
        with FlexRT(dict(wl["cfg"], precision=precision)) as b:
            inputs.setup(b, wl)
            for name in FIELDS:
                b.upload_field(name, a.download_field(name))
            b.read_ray_file(path)
            b.set_time(a.get_time())
            b.set_step(done)
            b.get_nphot_min_max()
            for s in range(3):
                sa, sb = a.step(), b.step()
                assert counts(sa) == counts(sb), "step %d" % s
            assert compare_rays(a.rays(), b.rays()) is None
            bad = [k for k in STATE_FIELDS if not bits_equal(a.download_field(k), b.download_field(k))]
            assert not bad, bad
    finally:
        a.close()


def test_vif_file(tmp_path):
    """write_vIF_binary's cube (io_funcs.cc:420-427): per row-major cell
    IF_speed, Inc_Flux, fion = 1, Delta_fion = 0."""
    _print_case("vIF file lognormal n=16 src=16 steps=4")
    wl, g = run(0, 4)
    try:
        path = tmp_path / "vIF_z=07.0000"
        g.write_vif_file(path)
        n3 = int(wl["cfg"]["n"]) ** 3
        v = np.frombuffer(path.read_bytes(), "<f4").reshape(n3, 4)
        assert bits_equal(v[:, 0], g.download_field("IF_speed"))
        assert bits_equal(v[:, 1], g.download_field("Inc_Flux"))
        assert np.all(v[:, 2] == 1) and np.all(v[:, 3] == 0)
        assert np.any(v[:, 0] > 0), "the workload must move ionization fronts"
    finally:
        g.close()


def test_file_errors(tmp_path):
    """The file entry points report FRT_E_IO, with nothing written to the
    device, for an unwritable path, a truncated ray file, a ray file without
    its end marker and a gas file of another grid; FRT_E_ARG for bad scale
    factors."""
    from flexrt_imaging_reion_amd.rt import FlexRTError

    FRT_E_IO = 7
    _print_case("file errors lognormal n=16 src=16 steps=2")
    wl, g = run(0, 2)
    try:
        before = g.rays()
        with pytest.raises(FlexRTError) as ei:
            g.write_ray_file(tmp_path / "no_such_dir" / "ray")
        assert ei.value.code == FRT_E_IO
        good = tmp_path / "ray"
        g.write_ray_file(good)
        raw = good.read_bytes()
        (tmp_path / "short").write_bytes(raw[:-3])
        (tmp_path / "nomark").write_bytes(raw[:-8] + (5).to_bytes(8, "little"))
        for bad in ("short", "nomark"):
            with pytest.raises(FlexRTError) as ei:
                g.read_ray_file(tmp_path / bad)
            assert ei.value.code == FRT_E_IO, bad
        assert compare_rays(g.rays(), before) is None
        (tmp_path / "gas").write_bytes(b"\0" * 64)
        with pytest.raises(FlexRTError) as ei:
            g.read_gas_file(tmp_path / "gas")
        assert ei.value.code == FRT_E_IO
        with pytest.raises(FlexRTError) as ei:
            g.update_hydro_step(7.0, -1.0, 1e-27, 1e14)
        assert ei.value.code == 1
        with pytest.raises(FlexRTError) as ei:
            g.update_cosmo_expansion(0.0, 0.125)
        assert ei.value.code == 1
    finally:
        g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_device_checkpoint_restore(precision):
    """A run returned to a device checkpoint, in its own context and in a
    fresh one, repeats the steps that followed it bit for bit."""
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("device checkpoint precision=%d lognormal n=16 src=16 3+3 steps" % precision)
    wl, a = run(precision, 3)
    try:
        ck = a.checkpoint()
        sa = [counts(a.step()) for _ in range(3)]
        ra, fa = a.rays(), {k: a.download_field(k) for k in STATE_FIELDS}
        a.restore(ck)
        sb = [counts(a.step()) for _ in range(3)]
        assert sa == sb
        assert compare_rays(a.rays(), ra) is None
        assert not [k for k in STATE_FIELDS if not bits_equal(a.download_field(k), fa[k])]
        with FlexRT(dict(wl["cfg"], precision=precision)) as b:
            inputs.setup(b, wl)
            b.restore(ck)
            sc = [counts(b.step()) for _ in range(3)]
            assert sa == sc
            assert compare_rays(b.rays(), ra) is None
            assert not [k for k in STATE_FIELDS if not bits_equal(b.download_field(k), fa[k])]
        ck.free()
    finally:
        a.close()

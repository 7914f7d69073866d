"""The checker's restatement of write_rays_binary (io_funcs.cc:274-303), on
the CPU: its file parses with the layout read_rays_binary consumes
(io_funcs.cc:119-145) and holds exactly the checker's rays, also for an empty
ray list (the end marker alone)."""
from flexrt_imaging_reion_amd import inputs
from oracle.pyoracle import Oracle
from parity_util import check_ray_file


def test_checker_ray_file_layout(tmp_path):
    wl = inputs.make_workload("lognormal64", n=16, n_src=8, steps=3)
    o = Oracle(wl["cfg"])
    try:
        inputs.setup(o, wl)
        for _ in range(3):
            o.step()
        o.write_ray_file(tmp_path / "rays")
        rays = o.rays()
        assert len(rays["cell"]) > 0
        assert check_ray_file((tmp_path / "rays").read_bytes(), rays, int(wl["cfg"]["nfreq"])) is None
    finally:
        o.close()


def test_checker_ray_file_empty(tmp_path):
    wl = inputs.make_workload("lognormal64", n=16, n_src=8, steps=1)
    o = Oracle(wl["cfg"])
    try:
        o.write_ray_file(tmp_path / "rays")
        raw = (tmp_path / "rays").read_bytes()
        assert raw == (-1).to_bytes(8, "little", signed=True)
    finally:
        o.close()


def test_checkpoint_names():
    """gas_funcs.cc:617-645: to_string(zz) (six decimals) less two characters,
    a leading 0 below z = 10, output_dir concatenated as given."""
    from flexrt_imaging_reion_amd.rt import checkpoint_paths

    p = checkpoint_paths("out/", 7.0)
    assert p == {"gas": "out/gas_z=07.0000", "ray": "out/ray_z=07.0000", "tre": "out/tre_z=07.0000"}
    assert checkpoint_paths("/d/", 12.345678)["gas"] == "/d/gas_z=12.3456"
    assert checkpoint_paths("", 9.99999)["tre"] == "tre_z=09.9999"


def test_vif_name():
    """io_funcs.cc:399-414: digits of z_rt, zero padding from zz."""
    from flexrt_imaging_reion_amd.rt import vif_path

    assert vif_path("o/", 7.0, 6.987654) == "o/vIF_z=06.9876"
    assert vif_path("o/", 10.5, 10.48) == "o/vIF_z=10.4800"

"""The benchmark configuration itself (128^3 cells, 4096 sources, 5 bins),
checked through properties that hold at any size, since the CPU checker is
too slow for it: the mixed form against the FP64 form (bit-identical to the
checker at every size the checker runs) within the mixed-precision targets,
run-to-run repeatability of the mixed form, and a ray-file round trip of the
full ray list."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from parity_util import STATE_FIELDS, bits_equal, compare_rays, counts, fields

pytestmark = pytest.mark.gpu

STEPS = 4


def _print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module")
def wl():
    return inputs.make_workload("batch128", steps=STEPS)


def run(wl, precision, steps=STEPS):
    from flexrt_imaging_reion_amd.rt import FlexRT

    g = FlexRT(dict(wl["cfg"], precision=precision))
    inputs.setup(g, wl)
    st = [g.step() for _ in range(steps)]
    return g, st


def test_full_size_mixed_against_fp64(wl):
    """x_HI median relative error < 1e-5, fewer than 0.5 % of cells beyond
    1e-3, ionized volume fraction within 1e-4 (the mixed-precision target)."""
    _print_case("full size batch128 mixed vs fp64, %d steps" % STEPS)
    gm, sm = run(wl, 1)
    gf, sf = run(wl, 0)
    try:
        xm = gm.download_field("nH1").astype(np.float64) / gm.download_field("nH")
        xf = gf.download_field("nH1").astype(np.float64) / gf.download_field("nH")
        assert np.all(np.isfinite(xm))
        rel = np.abs(xm - xf) / np.abs(xf)
        assert np.median(rel) < 1e-5
        assert np.mean(rel > 1e-3) < 5e-3
        assert abs(float((xm < 0.5).mean()) - float((xf < 0.5).mean())) < 1e-4
        # ray counts may differ only by threshold flips: well under 0.1 %
        assert abs(sm[-1]["rays_end"] - sf[-1]["rays_end"]) <= 1e-3 * sf[-1]["rays_end"]
    finally:
        gm.close()
        gf.close()


def test_full_size_mixed_repeatable(wl):
    _print_case("full size batch128 mixed twice, %d steps" % STEPS)
    a, sa = run(wl, 1)
    b, sb = run(wl, 1)
    try:
        assert [counts(s) for s in sa] == [counts(s) for s in sb]
        fa, fb = fields(a), fields(b)
        bad = [k for k in STATE_FIELDS if not bits_equal(fa[k], fb[k])]
        assert not bad, bad
    finally:
        a.close()
        b.close()


def test_full_size_ray_file_round_trip(wl, tmp_path):
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("full size batch128 ray file round trip, %d steps" % STEPS)
    g, _ = run(wl, 1)
    try:
        path = tmp_path / "ray_full"
        g.write_ray_file(path)
        rays = g.rays()
        n = len(rays["cell"])
        assert path.stat().st_size == n * (57 + 4 * int(wl["cfg"]["nfreq"])) + 8
        with FlexRT(dict(wl["cfg"], precision=1)) as h:
            inputs.setup(h, wl)
            h.read_ray_file(path)
            assert compare_rays(h.rays(), rays) is None
        path.unlink()
    finally:
        g.close()

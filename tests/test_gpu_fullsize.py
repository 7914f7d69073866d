"""The benchmark configuration itself (128^3 cells, 4096 sources, 5 bins),
checked through properties that hold at any size, since the CPU checker is
too slow for it: run-to-run repeatability of the mixed form over all 20
steps (merging starts near step 10), and a ray-file round trip of the full
ray list.  The mixed form against the FP64 form over the 20 steps is in
test_gpu_precision.py."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs
from parity_util import STATE_FIELDS, bits_equal, compare_rays, counts, fields

pytestmark = pytest.mark.gpu

STEPS = 20
FILE_STEPS = 4  # the ray-file round trip (2 GB of file at 20 steps)


def _print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module")
def wl():
    return inputs.make_workload("batch128", steps=STEPS)


def test_full_size_merges(wl):
    """The 20 steps reach the merge stage."""
    _print_case("full size batch128 mixed, %d steps, merging" % STEPS)
    g, st = run(wl, 1)
    try:
        assert sum(s["rays_merged_away"] for s in st) > 0
        assert all(np.isfinite(g.download_field(k)).all() for k in ("nH1", "temp", "gamma_H1_tot"))
    finally:
        g.close()


def run(wl, precision, steps=STEPS):
    from flexrt_imaging_reion_amd.rt import FlexRT

    g = FlexRT(dict(wl["cfg"], precision=precision))
    inputs.setup(g, wl)
    st = [g.step() for _ in range(steps)]
    return g, st


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

    _print_case("full size batch128 ray file round trip, %d steps" % FILE_STEPS)
    g, _ = run(wl, 1, FILE_STEPS)
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

"""The timed CPU baseline of bench.py: the checker resumed from a saved state
(every field, the ray list, time, step count and carried merge threshold)
continues exactly as the uninterrupted run, so single steps can be timed from
the FP64 GPU state; the OpenMP build of the checker (the reference's parallel
loops, sums in scheduling order) steps the same workload to the same counts
and, within rounding of the summation order, the same fields."""
import numpy as np

import bench
from flexrt_imaging_reion_amd import inputs
from oracle.pyoracle import Oracle, set_threads


def test_resumed_checker_continues_bit_exact():
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=6)
    snaps, stats = bench.fp64_snapshots(wl, 6, (4, 6), make=lambda: Oracle(wl["cfg"]))
    ref = Oracle(wl["cfg"])
    inputs.setup(ref, wl)
    for _ in range(6):
        ref.step()
    end = {k: ref.download_field(k) for k in ("nH1", "temp", "gamma_H1_tot")}
    ref.close()
    for s in (4, 6):
        o = bench.resume(wl, snaps[s], s, omp=False)
        try:
            st = o.step()
            for k in ("rays_end", "rays_split", "rays_dumped", "rays_terminated", "segments", "step"):
                assert st[k] == stats[s][k], (s, k)
            if s == 6:
                for k, v in end.items():
                    assert np.array_equal(o.download_field(k).view(np.uint32), v.view(np.uint32)), k
        finally:
            o.close()


def test_openmp_checker_matches_serial():
    set_threads(4)
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=5)
    runs = []
    for omp in (False, True):
        o = Oracle(wl["cfg"], omp=omp)
        inputs.setup(o, wl)
        st = [o.step() for _ in range(5)]
        runs.append(([s["rays_end"] for s in st], o.download_field("nH1").astype(np.float64)))
        o.close()
    (c0, x0), (c1, x1) = runs
    assert c0 == c1
    assert np.max(np.abs(x1 - x0) / x0) < 1e-5

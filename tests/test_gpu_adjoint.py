"""Adjoint of find_avg_unu (FP64 form) against central finite differences of
the forward itself, through the torch.autograd.Function wrapper: the
directional derivative of a weighted sum of x_HI along random directions in
the ray photons, the clumping factors and the per-cell log-density, with the
ray paths, dumps and recombination coefficients frozen (the perturbations are
small enough not to move a dump).

The forward stores n_HI in float, so a finite difference of a sum over many
cells carries ~1e-7 relative rounding per cell; the weights sit on the
cells the radiation has reached (ionized fraction above 1 %), where x_HI
moves most, which keeps that noise well below the 1 % tolerance.  The
clumping direction moves the loss least (a 1e-3 step changes it by ~6e-6 of
its value, about 50 float roundings), so it takes a 1e-2 step: measured on
this case, central differences at 3e-3 and 1e-2 agree with the adjoint to
4e-4 in both forms, at 1e-3 they scatter by 0.4 % (FP64) and 1.4 % (mixed)."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs

pytestmark = pytest.mark.gpu

H = 1e-3


def _print_case(*a):
    print("CASE", *a, flush=True)


def _state(n=16, n_src=8, steps=2, precision=0):
    from flexrt_imaging_reion_amd._abi import FIELDS
    from flexrt_imaging_reion_amd.rt import FlexRT

    wl = inputs.make_workload("lognormal64", n=n, n_src=n_src, steps=steps)
    rt = FlexRT(dict(wl["cfg"], precision=precision))
    inputs.setup(rt, wl)
    for _ in range(steps):
        rt.step()
    fields = {k: rt.download_field(k) for k in FIELDS}
    return wl, rt, fields, rt.rays(), rt.get_time()


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_against_finite_differences(precision):
    import torch

    from flexrt_imaging_reion_amd.adjoint import FindAvgUnu, FrozenStep

    _print_case("adjoint precision=%d lognormal n=16 src=8 after 4 steps" % precision)
    wl, rt, fields, rays, t = _state(steps=4, precision=precision)
    try:
        step = FrozenStep(rt, fields, rays, t)
        F = rt.nfreq
        n3 = 16 ** 3
        nph0 = np.asarray(rays["nphot"], dtype=np.float64).reshape(-1, F)
        c0 = fields["clump"].astype(np.float64)
        s0 = np.zeros(n3)
        nH0 = fields["nH"].astype(np.float64)
        rng = np.random.default_rng(3)
        x0 = step.run(nph0, c0, s0) / nH0
        ion = (1. - x0) > 0.01  # reached by the radiation: ionized fraction > 1 %
        assert 10 < ion.sum() < n3 // 2, "the state must hold ionized regions"
        w = np.where(ion, rng.standard_normal(n3), 0.)

        def loss_np(nph, c, s):
            return float(np.sum(w * step.run(nph, c, s) / (nH0 * np.exp(s))))

        tn = torch.tensor(nph0, requires_grad=True)
        tc = torch.tensor(c0, requires_grad=True)
        ts = torch.tensor(s0, requires_grad=True)
        nh1 = FindAvgUnu.apply(tn, tc, ts, step)
        L = (torch.from_numpy(w) * nh1 / (torch.from_numpy(nH0) * torch.exp(ts))).sum()
        L.backward()
        g = {"nph": tn.grad.numpy(), "c": tc.grad.numpy(), "s": ts.grad.numpy()}
        assert all(np.all(np.isfinite(v)) for v in g.values())
        dirs = {"nph": rng.standard_normal(nph0.shape) * nph0, "c": rng.standard_normal(n3) * c0,
                "s": rng.standard_normal(n3)}
        for name, v in dirs.items():
            h = 1e-2 if name == "c" else H
            args_p = {"nph": nph0, "c": c0, "s": s0}
            args_m = dict(args_p)
            args_p[name] = args_p[name] + h * v
            args_m[name] = args_m[name] - h * v
            fd = (loss_np(args_p["nph"], args_p["c"], args_p["s"]) -
                  loss_np(args_m["nph"], args_m["c"], args_m["s"])) / (2 * h)
            an = float(np.sum(g[name] * v))
            print("direction", name, "adjoint", an, "finite difference", fd, flush=True)
            assert abs(an - fd) <= 1e-2 * abs(fd), (name, an, fd)
        # source luminosities: the photons of each source's freshly cast rays scale with L
        from flexrt_imaging_reion_amd.adjoint import log_luminosity_gradient

        gl = log_luminosity_gradient(g["nph"], rays, wl["src_cells"])
        new = np.asarray(rays["r"]) == 0
        assert new.sum() == 12 * len(wl["src_cells"])  # 12 HEALPix rays per source cell
        j = int(np.argmax(np.abs(gl)))
        sel = (new & (np.asarray(rays["cell"]) == wl["src_cells"][j]))[:, None]
        fd = (loss_np(np.where(sel, nph0 * (1 + H), nph0), c0, s0) -
              loss_np(np.where(sel, nph0 * (1 - H), nph0), c0, s0)) / (2 * H)
        print("d/dlogL of source", j, "adjoint", gl[j], "finite difference", fd, flush=True)
        assert abs(gl[j] - fd) <= 1e-2 * abs(fd)
    finally:
        rt.close()


def test_adjoint_needs_a_tape():
    from flexrt_imaging_reion_amd.rt import FlexRT, FlexRTError

    _print_case("adjoint without tape")
    wl = inputs.make_workload("lognormal64", n=16, n_src=8, steps=1)
    with FlexRT(dict(wl["cfg"], precision=0)) as rt:
        inputs.setup(rt, wl)
        with pytest.raises(FlexRTError):
            rt.adj_find_avg_unu(np.zeros(16 ** 3))


def test_tape_leaves_forward_unchanged():
    """Recording the tape changes no bit of the forward."""
    from flexrt_imaging_reion_amd.adjoint import FrozenStep

    _print_case("adjoint tape on/off forward identity")
    wl, rt, fields, rays, t = _state(steps=3)
    try:
        step = FrozenStep(rt, fields, rays, t)
        args = (np.asarray(rays["nphot"]).reshape(-1, rt.nfreq), fields["clump"], np.zeros(16 ** 3))
        on = step.run(*args)
        rays_on = rt.rays()
        rt.adj_enable(False)
        off = step.run(*args)
        assert np.array_equal(on.view(np.uint64), off.view(np.uint64))
        from parity_util import compare_rays

        assert compare_rays(rt.rays(), rays_on) is None
    finally:
        rt.close()


def test_adjoint_at_the_differentiation_config():
    """The differentiation configuration (64^3, lognormal sigma_ln 0.8, 16
    sources on the densest cells, C = 2): d/dlog rho and d/dlog L of one
    source from the adjoint against finite differences, FP64 form."""
    import torch

    from flexrt_imaging_reion_amd._abi import FIELDS
    from flexrt_imaging_reion_amd.adjoint import FindAvgUnu, FrozenStep, log_luminosity_gradient
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("adjoint fp64 adjoint64 (64^3, 16 sources) after 3 steps")
    wl = inputs.make_workload("adjoint64", steps=3)
    rt = FlexRT(dict(wl["cfg"], precision=0))
    try:
        inputs.setup(rt, wl)
        for _ in range(3):
            rt.step()
        fields = {k: rt.download_field(k) for k in FIELDS}
        rays = rt.rays()
        step = FrozenStep(rt, fields, rays, rt.get_time())
        F, n3 = rt.nfreq, 64 ** 3
        nph0 = np.asarray(rays["nphot"], dtype=np.float64).reshape(-1, F)
        c0 = fields["clump"].astype(np.float64)
        s0 = np.zeros(n3)
        nH0 = fields["nH"].astype(np.float64)
        x0 = step.run(nph0, c0, s0) / nH0
        rng = np.random.default_rng(5)
        w = np.where((1. - x0) > 0.01, rng.standard_normal(n3), 0.)
        assert np.count_nonzero(w) > 100

        def loss_np(nph, s):
            return float(np.sum(w * step.run(nph, c0, s) / (nH0 * np.exp(s))))

        tn = torch.tensor(nph0, requires_grad=True)
        ts = torch.tensor(s0, requires_grad=True)
        nh1 = FindAvgUnu.apply(tn, torch.tensor(c0), ts, step)
        (torch.from_numpy(w) * nh1 / (torch.from_numpy(nH0) * torch.exp(ts))).sum().backward()
        v = rng.standard_normal(n3)
        fd = (loss_np(nph0, s0 + H * v) - loss_np(nph0, s0 - H * v)) / (2 * H)
        an = float(np.sum(ts.grad.numpy() * v))
        print("d/dlog rho adjoint", an, "finite difference", fd, flush=True)
        assert abs(an - fd) <= 1e-2 * abs(fd)
        gl = log_luminosity_gradient(tn.grad.numpy(), rays, wl["src_cells"])
        j = int(np.argmax(np.abs(gl)))
        sel = ((np.asarray(rays["r"]) == 0) & (np.asarray(rays["cell"]) == wl["src_cells"][j]))[:, None]
        fd = (loss_np(np.where(sel, nph0 * (1 + H), nph0), s0) -
              loss_np(np.where(sel, nph0 * (1 - H), nph0), s0)) / (2 * H)
        print("d/dlog L adjoint", gl[j], "finite difference", fd, flush=True)
        assert abs(gl[j] - fd) <= 1e-2 * abs(fd)
    finally:
        rt.close()


def test_mixed_adjoint_after_segment_growth():
    """Mixed form: the adjoint's per-slot buffer is its own (16-byte opacity
    pairs), sized by the step's segment count.  A context whose ray capacity
    is just above the live rays reserves 2 segments per ray, so the 32^3
    single-source case (about 4 crossings per moved ray) has more segments
    than half the segment buffer and grows it; its gradient must equal, bit
    for bit, that of a context with the default (large) capacity, where the
    capacity changes only array strides."""
    from flexrt_imaging_reion_amd.rt import FlexRT

    _print_case("mixed adjoint uniform32 3 steps, small vs default capacity")
    wl = inputs.make_workload("uniform32", steps=3)
    cfg = dict(wl["cfg"], precision=1)
    big = FlexRT(cfg)
    try:
        inputs.setup(big, wl)
        peak = big.ray_count()
        for _ in range(3):
            st = big.step()
            peak = max(peak, st["rays_start"], st["rays_end"])
        small_cap = int(peak * 1.1) + 12 * len(wl["src_cells"]) * 4 + 64
        small = FlexRT(dict(cfg, ray_capacity=small_cap))
        try:
            inputs.setup(small, wl)
            for _ in range(3):
                small.step()
            out = []
            for rt in (big, small):
                rt.adj_enable(True)
                rt.update_rays()
                rt.find_avg_unu()
                out.append(rt)
            nseg = small.stats()["segments"]
            print("segments", nseg, "small capacity", small_cap, flush=True)
            assert nseg > small_cap, "the case must overflow half of the reserved segment buffer"
            a = np.random.default_rng(2).standard_normal(32 ** 3)
            ga, gb = big.adj_find_avg_unu(a), small.adj_find_avg_unu(a)
            for k in ga:
                assert np.all(np.isfinite(ga[k])), k
                assert np.array_equal(ga[k].view(np.uint64), gb[k].view(np.uint64)), k
        finally:
            small.close()
    finally:
        big.close()

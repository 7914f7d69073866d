"""Adjoint across steps (frt_adj_run_* through adjoint.Trajectory and the
torch.autograd.Function XHIAfterRun) against central finite differences of the
forward run itself.

Case: 64^3 log-normal density, 16 sources, uniform clumping C = 2, 20 steps of
1 Myr, a checkpoint every 5 steps; loss = mean (x_HI - x_HI_target)^2 where the
target is the same run with every source rate times 1.3.  Gradients with
respect to the 16 log source rates, log C and the per-cell log density.

Step sizes and tolerances: h = 1e-4 for the source rates, C and the density
directions with a large derivative (uniform, and the gradient itself scaled
to unit rms).  The forward is piecewise smooth: a density change moves a
few rays across the photon floor or a thermal sub-step count, which shifts
the loss by ~5e-8 (measured: FD at h = 1e-5 .. 1e-2 along a random per-cell
direction scatters by that much in the loss difference).  Along the uniform
and gradient directions that is below 1 % of the difference at h = 1e-4
(measured 0.6 % and 0.2 % in both forms); a block-random direction (8^3
blocks of +-1) has a ~10x smaller derivative and takes h = 1e-3 (measured
2.3 % there, 8 % at 1e-4).  The checkpoint interval does not change the
reverse sweep's arithmetic: the gradients at K = 5 and K = 3 agree bit for
bit."""
import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs

pytestmark = pytest.mark.gpu


def _case(precision, K=5):
    from flexrt_imaging_reion_amd.adjoint import Trajectory

    print("CASE adjoint run precision=%d adjoint64 (64^3, 16 src, 20 steps) K=%d" % (precision, K), flush=True)
    wl = inputs.make_workload("adjoint64")
    return wl, Trajectory(wl, precision=precision, K=K)


def _gradient(tr, L0, C0, r0, xt):
    import torch

    from flexrt_imaging_reion_amd.adjoint import XHIAfterRun

    tL = torch.tensor(L0, requires_grad=True)
    tC = torch.tensor(C0, requires_grad=True)
    tr_ = torch.tensor(r0, requires_grad=True)
    x = XHIAfterRun.apply(tL, tC, tr_, tr)
    loss = torch.mean((x - torch.from_numpy(xt)) ** 2)
    loss.backward()
    return float(loss.detach()), tL.grad.numpy().copy(), float(tC.grad), tr_.grad.numpy().copy()


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_run_against_finite_differences(precision):
    wl, tr = _case(precision)
    try:
        n3 = tr.n3
        n = int(wl["cfg"]["n"])
        assert wl["steps"] == 20 and n == 64 and len(wl["src_lum"]) == 16
        L0 = np.log(np.asarray(wl["src_lum"], dtype=np.float64))
        C0 = float(np.log(2.0))
        r0 = np.zeros(n3)
        xt, _ = tr.run(L0 + np.log(1.3), C0, r0)

        def loss(L, C, r):
            x, _ = tr.run(L, C, r)
            return float(np.mean((x - xt) ** 2))

        lval, gL, gC, gr = _gradient(tr, L0, C0, r0, xt)
        assert lval > 0 and np.all(np.isfinite(gL)) and np.isfinite(gC) and np.all(np.isfinite(gr))
        assert abs(lval - loss(L0, C0, r0)) <= 1e-12 * lval

        H = 1e-4
        gmax = np.max(np.abs(gL))
        for j in range(len(L0)):
            e = np.zeros_like(L0)
            e[j] = H
            fd = (loss(L0 + e, C0, r0) - loss(L0 - e, C0, r0)) / (2 * H)
            print("dlogL[%d] adjoint %.6e fd %.6e" % (j, gL[j], fd), flush=True)
            assert abs(gL[j] - fd) <= 0.02 * abs(fd) + 2e-3 * gmax, j

        fd = (loss(L0, C0 + H, r0) - loss(L0, C0 - H, r0)) / (2 * H)
        print("dlogC adjoint %.6e fd %.6e" % (gC, fd), flush=True)
        assert abs(gC - fd) <= 0.02 * abs(fd)

        rng = np.random.default_rng(1)
        blk = rng.choice([-1.0, 1.0], size=(n // 8,) * 3)
        blk = np.repeat(np.repeat(np.repeat(blk, 8, 0), 8, 1), 8, 2).reshape(-1)
        dirs = [("uniform", np.ones(n3), 1e-4, 0.02), ("gradient", gr / np.sqrt(np.mean(gr * gr)), 1e-4, 0.02),
                ("block8", blk, 1e-3, 0.06)]
        for name, v, h, tol in dirs:
            an = float(np.sum(gr * v))
            fd = (loss(L0, C0, r0 + h * v) - loss(L0, C0, r0 - h * v)) / (2 * h)
            print("dlogrho.%s adjoint %.6e fd %.6e" % (name, an, fd), flush=True)
            assert abs(an - fd) <= tol * abs(fd), name
    finally:
        tr.close()


def test_adjoint_run_checkpoint_interval_invariant():
    wl, tr5 = _case(0, K=5)
    try:
        L0 = np.log(np.asarray(wl["src_lum"], dtype=np.float64))
        C0 = float(np.log(2.0))
        r0 = np.zeros(tr5.n3)
        xt, _ = tr5.run(L0 + np.log(1.3), C0, r0)
        a = _gradient(tr5, L0, C0, r0, xt)
    finally:
        tr5.close()
    wl, tr3 = _case(0, K=3)
    try:
        b = _gradient(tr3, L0, C0, r0, xt)
    finally:
        tr3.close()
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and a[2] == b[2] and np.array_equal(a[3], b[3])


def test_adjoint_run_refuses_out_of_order():
    from flexrt_imaging_reion_amd.rt import FlexRT, FlexRTError

    wl = inputs.make_workload("adjoint64", n=16, n_src=2, steps=1)
    rt = FlexRT(dict(wl["cfg"], precision=0))
    try:
        inputs.setup(rt, wl)
        with pytest.raises(FlexRTError):
            rt.adj_run_begin()  # no tape
        rt.adj_enable(True)
        ck = rt.checkpoint()
        rt.step()
        rt.adj_run_begin()
        rt.adj_seed(np.zeros(16 ** 3), None)
        rt.adj_step_reverse()
        with pytest.raises(FlexRTError):
            rt.adj_step_reverse()  # one reverse per taped step
        with pytest.raises(FlexRTError):
            rt.adj_run_finish(np.float32(wl["cfg"]["dt"]))  # not returned to the initial state
        rt.restore(ck)
        g = rt.adj_run_finish(np.float32(wl["cfg"]["dt"]))
        assert not np.any(g["lum"]) and not np.any(g["log_rho"])  # zero seed
        ck.free()
    finally:
        rt.close()

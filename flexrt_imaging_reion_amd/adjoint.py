"""Gradients of one find_avg_unu (ray_tracing.cc:623-1216, C2Ray branch) as a
torch.autograd.Function over the library's FP64 form.

The forward re-runs update_rays + find_avg_unu from a saved state with the
given ray photons, clumping factors and per-cell log-density shifts, and
returns the n_HI field; the backward calls the library's hand-written adjoint
kernels (frt_adj_find_avg_unu), with the ray paths, the dump decisions and
the recombination coefficients held at their forward values.  A backward must
follow its own forward: the next forward records a new tape.

Source luminosities enter through the photons of the rays they cast (new rays
carry spectrum * L * ray_dt / 12 / phot_norm photons, ray_tracing.cc:302-304),
so d/dL follows from the photon gradient of those rays.
"""
import numpy as np
import torch

# fields that scale with the gas density at fixed ionized fractions
DENSITY_FIELDS = ("nH", "nHe", "nH1", "nHe1", "nH2", "nHe2", "ne", "n_tot", "avg_nh1", "avg_nhe1", "rho")


class FrozenStep:
    """A FlexRT context (FP64 form) and a saved state (every field, the ray
    list, the time) from which one find_avg_unu is re-run."""

    def __init__(self, rt, fields, rays, t):
        from ._abi import FIELDS

        self.rt = rt
        self.fields = {k: np.array(fields[k]) for k in FIELDS}
        self.rays = {k: np.array(v) for k, v in rays.items()}
        self.t = float(t)
        self.nfreq = rt.nfreq
        rt.adj_enable(True)

    def run(self, nphot, clump, log_rho):
        """n_HI after find_avg_unu for ray photons nphot [ray][nfreq] (stored
        units), clumping factors and log-density shifts (row-major cells)."""
        rt = self.rt
        scale = np.exp(np.asarray(log_rho, dtype=np.float64))
        for k, v in self.fields.items():
            if k == "clump":
                v = np.asarray(clump, dtype=np.float64)
            elif k in DENSITY_FIELDS:
                v = v.astype(np.float64) * scale
            rt.upload_field(k, np.asarray(v, dtype=np.float32))
        rays = dict(self.rays)
        rays["nphot"] = np.asarray(nphot, dtype=np.float32).reshape(-1)
        rt.upload_rays(rays)
        rt.set_time(self.t)
        rt.get_nphot_min_max()
        rt.update_rays()
        rt.find_avg_unu()
        return rt.download_field("nH1").astype(np.float64)


class FindAvgUnu(torch.autograd.Function):
    """n_HI = find_avg_unu(nphot, clump, log_rho) on a FrozenStep."""

    @staticmethod
    def forward(ctx, nphot, clump, log_rho, step):
        out = step.run(nphot.detach().cpu().numpy(), clump.detach().cpu().numpy(),
                       log_rho.detach().cpu().numpy())
        ctx.step = step
        ctx.scale = np.exp(log_rho.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, grad_nh1):
        step = ctx.step
        g = step.rt.adj_find_avg_unu(grad_nh1.detach().cpu().numpy())
        f, s = step.fields, ctx.scale
        # d/dlog(rho): every input density scales with rho
        g_log_rho = sum(g[k] * f[k].astype(np.float64) * s for k in ("nH", "nHe", "nH1", "nHe1"))
        return (torch.from_numpy(np.array(g["nphot"])), torch.from_numpy(g["clump"]),
                torch.from_numpy(g_log_rho), None)


def log_luminosity_gradient(grad_nphot, rays, src_cells):
    """d/dlog(L) of each source cell from the photon gradient of one step.

    The rays a source cast at the end of the previous step (new_rays,
    rt_funcs.cc:240-261) have not moved yet: r = 0 and their cell is the
    source cell, and their photons are proportional to its luminosity
    (ray_tracing.cc:302-304).  Rays cast earlier depend on it through earlier
    steps, which this one-step gradient holds fixed."""
    g = np.asarray(grad_nphot, dtype=np.float64)
    nph = np.asarray(rays["nphot"], dtype=np.float64).reshape(g.shape)
    new = np.asarray(rays["r"]) == 0
    cell = np.asarray(rays["cell"])
    per_ray = np.sum(g * nph, axis=1)
    return np.array([float(per_ray[new & (cell == c)].sum()) for c in np.asarray(src_cells)])



# Fields of the initial gas state (inputs.init_gas) that scale with the density.
INITIAL_DENSITY_FIELDS = DENSITY_FIELDS + ("nH1_prev",)


class Trajectory:
    """A run of `steps` RT steps from a workload's initial state, as a function
    of the source luminosities, a uniform clumping factor and the per-cell
    density, with its gradient from the library's adjoint across steps.

    The forward keeps a device checkpoint every K steps.  The reverse sweep
    walks the blocks backwards; for each step it returns to the block's
    checkpoint, recomputes the march and the steps up to it, runs the step
    with the tape on and reverses it (frt_adj_step_reverse): the thermal
    update, find_avg_unu's five passes (chemistry and attenuation march with
    their paths and dump decisions held), split and compaction, and the cast,
    which hands the adjoint of each new ray's photons to its source.  The
    split, dump, survival and thermal sub-step decisions are held at their
    forward values; a run whose steps merge rays is refused."""

    def __init__(self, wl, precision=0, K=5):
        from .rt import FlexRT

        self.wl = wl
        self.K = int(K)
        self.steps = int(wl["steps"])
        self.rt = FlexRT(dict(wl["cfg"], precision=precision))
        self.n3 = int(wl["cfg"]["n"]) ** 3
        self.ck = {}

    def close(self):
        for c in self.ck.values():
            c.free()
        self.ck = {}
        self.rt.close()

    def _setup(self, log_L, log_C, log_rho):
        from . import inputs

        wl = dict(self.wl)
        scale = np.exp(np.asarray(log_rho, dtype=np.float64))
        f = {}
        for k, v in wl["fields"].items():
            if k in INITIAL_DENSITY_FIELDS:
                v = (np.asarray(v, dtype=np.float64) * scale).astype(np.float32)
            f[k] = v
        f["clump"] = np.full(self.n3, np.float32(np.exp(float(log_C))), dtype=np.float32)
        wl["fields"] = f
        wl["src_lum"] = np.exp(np.asarray(log_L, dtype=np.float64))
        inputs.setup(self.rt, wl)

    def run(self, log_L, log_C, log_rho, record=False):
        """x_HI = nH1/nH after the steps (float64 of the stored floats) and nH."""
        rt = self.rt
        for c in self.ck.values():
            c.free()
        self.ck = {}
        self._setup(log_L, log_C, log_rho)
        rt.adj_enable(record)
        for s in range(self.steps):
            if record and s % self.K == 0:
                self.ck[s] = rt.checkpoint()
            rt.step()
        nH = rt.download_field("nH").astype(np.float64)
        return rt.download_field("nH1").astype(np.float64) / nH, nH

    def gradient(self, a_nH1, a_nH):
        """Adjoint of the recorded run for the seed (a_nH1, a_nH) of its final
        state: d/dL per source, d/dC per cell, d/dlog rho per cell."""
        rt = self.rt
        rt.adj_run_begin()
        rt.adj_seed(a_nH1, a_nH)
        for b in sorted(self.ck, reverse=True):
            for s in reversed(range(b, min(b + self.K, self.steps))):
                rt.restore(self.ck[b])
                for _ in range(b, s):
                    rt.step()
                rt.step()
                rt.adj_step_reverse()
        rt.restore(self.ck[0])
        return rt.adj_run_finish(np.float32(self.wl["cfg"]["dt"]))


class XHIAfterRun(torch.autograd.Function):
    """x_HI after a Trajectory's steps as a function of (log L per source,
    log C, log rho per cell)."""

    @staticmethod
    def forward(ctx, log_L, log_C, log_rho, traj):
        x, nH = traj.run(log_L.detach().cpu().numpy(), float(log_C.detach()), log_rho.detach().cpu().numpy(),
                         record=True)
        ctx.traj = traj
        ctx.x, ctx.nH = x, nH
        ctx.L = np.exp(log_L.detach().cpu().numpy().astype(np.float64))
        ctx.C = float(np.exp(float(log_C.detach())))
        return torch.from_numpy(x)

    @staticmethod
    def backward(ctx, grad_x):
        gx = grad_x.detach().cpu().numpy().astype(np.float64)
        g = ctx.traj.gradient(gx / ctx.nH, -gx * ctx.x / ctx.nH)
        return (torch.from_numpy(g["lum"] * ctx.L), torch.tensor(float(np.sum(g["clump"])) * ctx.C,
                                                                 dtype=torch.float64),
                torch.from_numpy(g["log_rho"]), None)

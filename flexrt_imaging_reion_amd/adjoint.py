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


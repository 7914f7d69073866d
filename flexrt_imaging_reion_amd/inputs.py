"""Synthetic inputs of the benchmark workloads and the driver-side setup.

The reference reads a density cube and a halo catalogue from disk
(io_funcs.cc:33-62,157-233) and initializes the gas in init_funcs.cc:10-143.
Here the same quantities are generated from a seed, identically for the GPU
library and for the CPU checker, and handed to either through the same
calls (``setup``), following main.cc:29-105.

Workloads (grid, box, redshift, density, sources, steps):
  uniform32      32^3, 5 cMpc/h, z=8, uniform gas, 1 source of 1e53 /s at the
                 box centre, 10 steps of 1 Myr, C=1
  lognormal64    64^3, 20 cMpc/h, z=7, lognormal (sigma_ln 1, P(k)~k^-2, seed 1),
                 64 sources on the 64 densest cells, Ndot in [1e52, 1e55],
                 20 steps of 1 Myr
  batch128       128^3, 40 cMpc/h, z=7, lognormal seed 2, 4096 sources drawn
                 with probability ~(1+delta)^2, Ndot in [1e51, 1e56], C=3,
                 20 steps of 0.5 Myr
  adjoint64      64^3, 20 cMpc/h, z=7, lognormal (sigma_ln 0.8, seed 3), 16
                 sources on the densest cells, Ndot in [1e53, 1e55], C=2,
                 20 steps of 1 Myr
  full256        256^3, 80 cMpc/h, z=6.5, lognormal (sigma_ln 1.2, seed 4),
                 T0 uniform in [50, 200] K, 32768 sources ~(1+delta)^2,
                 Ndot in [1e51, 1e56], C = 1 + 2 (1+delta)^0.5, 50 steps of 0.5 Myr
  subgrid32      32^3, 5 cMpc/h, z=8, lognormal (sigma_ln 0.8, seed 5), 4
                 sources on the densest cells, Ndot in [1e53, 1e54], C=2, 10
                 steps of 1 Myr, the sub-grid I-front model (FP64 form) with
                 placeholder RadHydro tables, one bin at 19 eV as the
                 reference's production setting
All but subgrid32 use 5 log-spaced bins over 13.6-54.4 eV with L_nu ~ nu^-1.5.
"""
import math

import numpy as np

# global_constants.h
Y_HE = 0.246
M_H = 1.672622e-24
H_EV = 4.135667e-15
KPC_TO_CM = 3.086e21
YR_TO_S = 3.154e7
CL = 2.998e10
RHO_C_CGS = 1.87890e-29  # critical density / h^2 (io_funcs.cc:41)

FH1_0 = np.float32(1.0 - 1.2e-3)   # user_inputs.h:98
FHE1_0 = np.float32(1.0 - 1.2e-3)  # user_inputs.h:99
FHE2_0 = np.float32(1.2e-3)        # user_inputs.h:100


def reference_defaults(n):
    """Python copy of frt_config_default (user_inputs.h / global_variables.cc values)."""
    return dict(
        n=n, nfreq=1, path_length=4, itercount=5, precision=0, hpx_lvl=0, l_max=12,
        ray_per_cell_max=22, rec_case_b=1, temp_ev=1, recomb_cool=1, free_free=1, compton=1,
        frac_limit=1, subgrid=0, c2ray_off=0, ray_capacity=0, seed=1,
        box_ckpc=40 * 1e3 / 0.68, a_rt=1.0 / (1.0 + 12.0038), zz=12.0038, dt=0.0,
        cl_factor=0.3, nl_const=0.5, safety_factor=2.0, min_dist=1e-4, temp_min=1e2,
        hh=0.68, omega_m=0.305147, omega_l=0.6948,
    )


# ---- spectrum -------------------------------------------------------------------------
def sigma_hi_verner(e_ev):
    """Verner et al. 1996 HI photo-ionization cross-section (cm^2) at energy e_ev."""
    e = np.asarray(e_ev, dtype=np.float64)
    x = e / 4.298e-1
    y = x
    F = ((x - 1.0) ** 2) * y ** (0.5 * 2.963 - 5.5) * (1.0 + np.sqrt(y / 32.88)) ** (-2.963)
    return np.where((e >= 13.6) & (e <= 5e4), 5.475e-14 * F, 0.0)


def powerlaw_bins(alpha=1.5, emin=13.6, emax=54.4, nbins=5):
    """Log-spaced bins over [emin, emax] eV for L_nu ~ nu^-alpha.

    Returns (nu_phot [Hz], weights): weights are the photon fractions per bin
    (integral of nu^-(alpha+1)), nu_phot is the frequency whose HI
    cross-section equals the photon-weighted bin average of the Verner fit.
    """
    edges = np.geomspace(emin, emax, nbins + 1)
    w = np.array([(lo ** (-alpha) - hi ** (-alpha)) / alpha for lo, hi in zip(edges[:-1], edges[1:])])
    w = w / w.sum()
    nu = np.empty(nbins)
    for b in range(nbins):
        lo, hi = edges[b], edges[b + 1]
        e = np.geomspace(lo, hi, 4001)
        pw = e ** (-alpha - 1.0)
        sig = sigma_hi_verner(e)
        avg = np.trapezoid(sig * pw, e) / np.trapezoid(pw, e)
        a, c = lo, hi  # sigma decreases with energy: bisect sigma(E) = avg
        for _ in range(200):
            m = 0.5 * (a + c)
            if sigma_hi_verner(m) > avg:
                a = m
            else:
                c = m
        nu[b] = 0.5 * (a + c) / H_EV
    return nu, w.astype(np.float32)


def spectrum_from_table(alpha_table, table, alpha_spec=1.5):
    """set_spectrum with multi_freq (gas_funcs.cc:44-70): linear interpolation of a
    spectrum table over the spectral index, then normalization (float)."""
    alpha_table = np.asarray(alpha_table, dtype=np.float64)
    table = np.asarray(table, dtype=np.float64)
    na = alpha_table.size
    m = int(math.floor((alpha_spec - alpha_table[0]) / (alpha_table[-1] - alpha_table[0]) * (na - 1)))
    m = min(m, na - 2)
    p = m + 1
    spec = (table[m] + (table[p] - table[m]) / (alpha_table[p] - alpha_table[m]) *
            (alpha_spec - alpha_table[m])).astype(np.float32)
    s = np.float32(0.0)
    for v in spec:
        s = np.float32(s + v)
    return (spec / s).astype(np.float32)


# ---- density and sources ----------------------------------------------------------------
def gaussian_field(n, seed, slope=-2.0):
    """Unit-variance Gaussian random field on the periodic n^3 grid, P(k) ~ k^slope."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n))
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.zeros_like(k)
    nz = k > 0
    amp[nz] = k[nz] ** (0.5 * slope)
    g = np.fft.irfftn(fk * amp, s=(n, n, n), axes=(0, 1, 2))
    g -= g.mean()
    g /= g.std()
    return g


def lognormal_overdensity(n, sigma_ln, seed, slope=-2.0):
    """1 + delta = exp(sigma g - sigma^2/2), renormalized to unit mean."""
    g = gaussian_field(n, seed, slope)
    rho = np.exp(sigma_ln * g - 0.5 * sigma_ln ** 2)
    return rho / rho.mean()


def mean_baryon_density(z, h, omega_b):
    """Proper mean baryon density in g/cm^3 at redshift z."""
    return RHO_C_CGS * omega_b * h * h * (1.0 + z) ** 3


def sources_densest(one_plus_delta, count, lo, hi, seed):
    """`count` sources on the densest cells, Ndot log-uniform in [10^lo, 10^hi]."""
    flat = one_plus_delta.reshape(-1)
    cells = np.argsort(-flat, kind="stable")[:count]
    rng = np.random.default_rng([seed, 2])
    lum = 10.0 ** rng.uniform(lo, hi, size=count)
    return cells.astype(np.int64), lum


def sources_weighted(one_plus_delta, count, lo, hi, seed, power=2.0):
    """`count` sources drawn with replacement with probability ~ (1+delta)^power."""
    flat = one_plus_delta.reshape(-1).astype(np.float64)
    p = flat ** power
    p /= p.sum()
    rng = np.random.default_rng([seed, 1])
    cells = rng.choice(flat.size, size=count, replace=True, p=p)
    rng2 = np.random.default_rng([seed, 2])
    lum = 10.0 ** rng2.uniform(lo, hi, size=count)
    return cells.astype(np.int64), lum


def grid_sources(cells, lum, n3):
    """read_source_catalog (io_funcs.cc:190-232): sources in one cell are summed,
    then listed in cell-index order."""
    grid = np.zeros(n3, dtype=np.float64)
    for c, l in zip(cells, lum):
        grid[int(c)] += l
    idx = np.nonzero(grid != 0.0)[0]
    return idx.astype(np.int32), grid[idx]


# ---- gas initialization (init_funcs.cc:10-143, fresh start) -----------------------------------
def init_gas(rho, temp0, clump):
    """Fields of init() for a fresh start.  fion is still 0 when init() runs
    (init_rt sets it later, main.cc:54,72), so the fion == 0 branch applies.
    f_He3 = 1 - f_He1 - f_He2 (init_funcs.cc:41).  mfp and mfp_912 are left 0
    here: they are outputs (gas file) that the first C2Ray step sets."""
    rho = np.asarray(rho, dtype=np.float32).reshape(-1)
    n3 = rho.size
    f64 = np.float64
    nH = ((1.0 - Y_HE) * rho.astype(f64) / M_H).astype(np.float32)
    nHe = (Y_HE * rho.astype(f64) / (4.0 * M_H)).astype(np.float32)
    nH2 = (nH.astype(f64) * (1.0 - f64(FH1_0))).astype(np.float32)
    nHe2 = (nHe.astype(f64) * (1.0 - f64(FHE1_0))).astype(np.float32)
    nH1 = (nH - nH2).astype(np.float32)
    nHe1 = (nHe - nHe2).astype(np.float32)
    ne = (nH2 + nHe2).astype(np.float32)
    n_tot = ((nH + nHe) + ne).astype(np.float32)
    f_H1 = np.full(n3, FH1_0, dtype=np.float32)
    f_H2 = np.full(n3, np.float32(1.0 - f64(FH1_0)), dtype=np.float32)
    f_He1 = np.full(n3, FHE1_0, dtype=np.float32)
    f_He2 = np.full(n3, FHE2_0, dtype=np.float32)
    temp = np.broadcast_to(np.asarray(temp0, dtype=np.float32), (n3,)).astype(np.float32).copy()
    clump = np.broadcast_to(np.asarray(clump, dtype=np.float32), (n3,)).astype(np.float32).copy()
    f_He3 = np.full(n3, np.float32(1.0 - f64(FHE1_0) - f64(FHE2_0)), dtype=np.float32)
    return dict(nH=nH, nHe=nHe, nH1=nH1, nHe1=nHe1, nH2=nH2, nHe2=nHe2, ne=ne, n_tot=n_tot,
                f_H1=f_H1, f_H2=f_H2, f_He1=f_He1, f_He2=f_He2, f_He3=f_He3, temp=temp, clump=clump,
                nH1_prev=nH1.copy(), avg_nh1=nH1.copy(), avg_nhe1=nHe1.copy(), rho=rho.copy(),
                # the backward-difference state of the c2ray_off mode (init_funcs.cc:77,92-95)
                nH2_prev=nH2.copy(), nHe1_prev=nHe1.copy(), nHe2_prev=nHe2.copy())


# ---- workloads ----------------------------------------------------------------------------------
WORKLOADS = {
    "uniform32": dict(n=32, box_cmpc_h=5.0, z=8.0, sigma_ln=None, seed=0, n_src=1, src="center",
                      lum=(53.0, 53.0), steps=10, dt_myr=1.0, clump=1.0, temp=(100.0, 100.0)),
    "lognormal64": dict(n=64, box_cmpc_h=20.0, z=7.0, sigma_ln=1.0, seed=1, n_src=64, src="densest",
                        lum=(52.0, 55.0), steps=20, dt_myr=1.0, clump=1.0, temp=(100.0, 100.0)),
    "batch128": dict(n=128, box_cmpc_h=40.0, z=7.0, sigma_ln=1.0, seed=2, n_src=4096, src="weighted",
                     lum=(51.0, 56.0), steps=20, dt_myr=0.5, clump=3.0, temp=(100.0, 100.0)),
    "adjoint64": dict(n=64, box_cmpc_h=20.0, z=7.0, sigma_ln=0.8, seed=3, n_src=16, src="densest",
                      lum=(53.0, 55.0), steps=20, dt_myr=1.0, clump=2.0, temp=(100.0, 100.0)),
    "full256": dict(n=256, box_cmpc_h=80.0, z=6.5, sigma_ln=1.2, seed=4, n_src=32768, src="weighted",
                    lum=(51.0, 56.0), steps=50, dt_myr=0.5, clump="subgrid", temp=(50.0, 200.0)),
    "subgrid32": dict(n=32, box_cmpc_h=5.0, z=8.0, sigma_ln=0.8, seed=5, n_src=4, src="densest",
                      lum=(53.0, 54.0), steps=10, dt_myr=1.0, clump=2.0, temp=(100.0, 100.0), subgrid=True),
}
COSMO = dict(h=0.7, omega_m=0.3, omega_b=0.048)


def path_length_for(cfg):
    """Segments a ray can need in one step: ceil(sqrt(3) c_red dt / dx) + 3."""
    dx = (np.float32(cfg["box_ckpc"]) * np.float32(cfg["a_rt"]) / np.float32(cfg["n"])) * KPC_TO_CM
    cells = cfg["cl_factor"] * CL * cfg["dt"] / float(dx)
    return int(math.ceil(math.sqrt(3.0) * cells)) + 3


def make_workload(name, n=None, n_src=None, steps=None, precision=0, seed_offset=0, c2ray_off=False):
    """Build the inputs of workload `name`; n / n_src / steps override its size.
    c2ray_off runs it in the reference's c2ray_iter = FALSE mode (FP64 form)."""
    w = dict(WORKLOADS[name])
    if n is not None:
        w["n"] = int(n)
    if n_src is not None:
        w["n_src"] = int(n_src)
    if steps is not None:
        w["steps"] = int(steps)
    nn = w["n"]
    n3 = nn ** 3
    seed = w["seed"] + seed_offset
    h = COSMO["h"]
    cfg = reference_defaults(nn)
    cfg.update(nfreq=5, precision=precision, box_ckpc=w["box_cmpc_h"] * 1e3 / h, a_rt=1.0 / (1.0 + w["z"]),
               zz=w["z"], dt=w["dt_myr"] * 1e6 * YR_TO_S, hh=h, omega_m=COSMO["omega_m"],
               omega_l=1.0 - COSMO["omega_m"], seed=1000 + seed)
    cfg["path_length"] = path_length_for(cfg)
    if c2ray_off:
        cfg["c2ray_off"] = 1
    # path entries are indexed in int32: capacity x path_length stays below 2^31
    # (at 256^3 the reference's 30 n^3 slots would not)
    limit = (2 ** 31 - 1) // cfg["path_length"] - 1
    if 30 * n3 > limit:
        cfg["ray_capacity"] = limit
    if w["sigma_ln"] is None:
        opd = np.ones((nn, nn, nn))
    else:
        opd = lognormal_overdensity(nn, w["sigma_ln"], seed)
    rho = (mean_baryon_density(w["z"], h, COSMO["omega_b"]) * opd).astype(np.float32)
    if w["src"] == "center":
        c = nn // 2
        cells = np.array([(c * nn + c) * nn + c], dtype=np.int64)
        lum = np.full(1, 10.0 ** w["lum"][0])
    elif w["src"] == "densest":
        cells, lum = sources_densest(opd, min(w["n_src"], n3), w["lum"][0], w["lum"][1], seed)
    else:
        cells, lum = sources_weighted(opd, w["n_src"], w["lum"][0], w["lum"][1], seed)
    src_cells, src_lum = grid_sources(cells, lum, n3)
    if w["temp"][0] == w["temp"][1]:
        temp0 = np.float32(w["temp"][0])
    else:
        rng = np.random.default_rng([seed, 3])
        temp0 = rng.uniform(w["temp"][0], w["temp"][1], size=n3).astype(np.float32)
    if w["clump"] == "subgrid":
        clump = (1.0 + 2.0 * np.sqrt(opd.reshape(-1))).astype(np.float32)
    else:
        clump = np.float32(w["clump"])
    fields = init_gas(rho, temp0, clump)
    nu, spec = powerlaw_bins()
    out = dict(name=name, cfg=cfg, fields=fields, src_cells=src_cells, src_lum=src_lum,
               nu_phot=nu, spectrum=spec, steps=w["steps"], params=w)
    if w.get("subgrid"):
        # the reference's sub-grid production setting is monochromatic,
        # Nfreq = 1 at 19 eV (user_inputs.h:79,95): with several bins the
        # I-front count's per-bin min() can empty one bin while others
        # survive, and the next front cell's tau update (ray_tracing.cc:911)
        # is then 0/0 for it
        from . import subgrid
        cfg["subgrid"] = 1
        cfg["nfreq"] = 1
        out["nu_phot"] = np.array([19.0 / H_EV])
        out["spectrum"] = np.array([1.0], dtype=np.float32)
        out["subgrid"] = subgrid.make_params(cfg, rho, temp_0=w["temp"][0])
    return out


def setup(rt, wl):
    """main.cc:29-105 for a fresh start on an RT backend (GPU or checker).

    Every cell field is uploaded (the ones init() does not set as zeros, as
    the reference's zero-initialized globals), so a context can be reset."""
    from ._abi import FIELDS, PER_BIN

    n3 = int(wl["cfg"]["n"]) ** 3
    nf = int(wl["cfg"]["nfreq"])
    for k in FIELDS:
        if k in wl["fields"]:
            rt.upload_field(k, wl["fields"][k])
        else:
            rt.upload_field(k, np.zeros(n3 * nf if k in PER_BIN else n3, dtype=np.float32))
    rt.set_spectrum(wl["nu_phot"], wl["spectrum"])
    rt.set_sources(wl["src_cells"], wl["src_lum"])
    if "subgrid" in wl:
        rt.set_subgrid(wl["subgrid"])
    rt.init_rt()
    rt.new_rays()
    return rt.get_nphot_min_max()

"""Inputs of the sub-grid I-front model (cfg["subgrid"] = 1).

The reference reads 27 RadHydro tables at start (init_radhydro_data and
import_subgrid_file, subgrid.cc:20-44,132-184): per (z_re, delta, Gamma)
node a text file of Nt rows (time since ionization [yr], frequency-averaged
MFP and MFP at 912 A [ckpc/h], total HI fraction, temperature [K]).  In the
production setting (equilibrium = TRUE, temp_ev = TRUE) only the time and
HI-fraction columns are read by the step; ``read_radhydro`` loads them from
the reference's files, ``placeholder_tables`` makes smooth synthetic ones of
the same shape (the reference's tables are not part of its repository).

``make_params`` returns the dict that ``FlexRT.set_subgrid`` (and the
checker's) take: the two [zre][delta][gamma][nt] arrays and the driver's
cosmic-time constants.
"""
import math
import os

import numpy as np

YR_TO_S = 3.154e7      # global_constants.h:43
ZRE_DOM = (12.0, 8.0, 6.0)                  # subgrid.cc:137-139
GAMMA_TAGS = ("G14", "G13", "G12")          # subgrid.cc:152
ZRE_TAGS = ("zre12", "zre8", "zre6")
DENSITY_TAGS = ("dm3", "d0", "dp3")


def cosmic_time(z, hh=0.68, omega_m=0.305147):
    """toTime (cosmo_funcs.cc:95-101): cosmic time at redshift z, s, in the
    matter + Lambda closed form the reference uses for MP-Gadget."""
    h0 = 3.24086e-18
    a = 1.0 / (1.0 + z)
    return 1.0 / (1.5 * h0 * hh * math.sqrt(1.0 - omega_m)) * math.asinh(
        math.sqrt((1.0 - omega_m) / omega_m) * math.sqrt(a) * a)


def read_radhydro(subgrid_dir, model="equilibrium"):
    """The time and HI-fraction columns of the reference's RadHydro files
    (file names of subgrid.cc:152-177), as import_subgrid_file converts them:
    float values, time multiplied by yr_to_s.  Returns (nt, time, fh1) with
    arrays of shape [3 zre][3 delta][3 gamma][nt]."""
    ext = {"equilibrium": "_noclump", "relaxed": "_relaxed", "full_model": ""}[model]
    tabs = {}
    nt = None
    for q, gt in enumerate(GAMMA_TAGS):
        for m, zt in enumerate(ZRE_TAGS):
            for p, dt in enumerate(DENSITY_TAGS):
                path = os.path.join(subgrid_dir, model, "%s_%s_%s%s.txt" % (gt, zt, dt, ext))
                rows = np.loadtxt(path, dtype=np.float64).reshape(-1, 5)
                nt = rows.shape[0] if nt is None else nt
                if rows.shape[0] != nt:
                    raise ValueError("%s: %d rows, expected %d" % (path, rows.shape[0], nt))
                t = (rows[:, 0].astype(np.float32).astype(np.float64) * YR_TO_S).astype(np.float32)
                tabs[(m, p, q)] = (t, rows[:, 3].astype(np.float32))
    time = np.zeros((3, 3, 3, nt), np.float32)
    fh1 = np.zeros((3, 3, 3, nt), np.float32)
    for (m, p, q), (t, f) in tabs.items():
        time[m, p, q] = t
        fh1[m, p, q] = f
    return nt, time, fh1


def placeholder_tables(nt=400, t_max_yr=1.0e9):
    """Synthetic tables of the reference's shape: a common uniform time grid,
    and a total HI fraction that falls with Gamma (~Gamma^-0.8), rises with
    density (~(1+delta)^0.9 over the nodes -sqrt3, 0, +sqrt3 of delta/sigma)
    and with later reionization, and relaxes with the time since ionization
    ((1 + t / 50 Myr)^-0.4); capped to (1e-7, 0.45)."""
    t_yr = np.linspace(0.0, t_max_yr, nt)
    time = np.empty((3, 3, 3, nt), np.float32)
    fh1 = np.empty((3, 3, 3, nt), np.float32)
    for m in range(3):
        for p, d in enumerate((-math.sqrt(3.0), 0.0, math.sqrt(3.0))):
            for q, g in enumerate((3e-14, 3e-13, 3e-12)):
                time[m, p, q] = (t_yr.astype(np.float32).astype(np.float64) * YR_TO_S).astype(np.float32)
                f = 2e-4 * (g / 3e-13) ** -0.8 * math.exp(0.9 * d / math.sqrt(3.0)) * (1.0 + 0.3 * m) \
                    * (1.0 + t_yr / 5e7) ** -0.4
                fh1[m, p, q] = np.clip(f, 1e-7, 0.45).astype(np.float32)
    return nt, time, fh1


def make_params(cfg, rho, zinit=None, tables=None, delta_nl=(-0.6, 0.0, 1.8), temp_0=100.0,
                fh1_0=1.0 - 1.2e-3):
    """frt_subgrid values for a run with configuration `cfg` on density `rho`:
    cosmic times from ``cosmic_time`` with the run's cosmology, avg_rho the
    volume average of rho (float), the given tables or placeholder ones."""
    nt, time, fh1 = tables if tables is not None else placeholder_tables()
    hh, om = float(cfg["hh"]), float(cfg["omega_m"])
    z0 = float(cfg["zz"]) if zinit is None else float(zinit)
    ct = lambda z: float(np.float32(cosmic_time(z, hh, om)))
    rho = np.asarray(rho, dtype=np.float32).reshape(-1)
    return dict(nt=int(nt), time_rh=time, fh1_rh=fh1, tre_dom=[ct(z) for z in ZRE_DOM], t_z12=ct(12.0),
                t_z8=ct(8.0), tinit=ct(z0), avg_rho=float(np.float32(rho.astype(np.float64).mean())),
                temp_0=float(np.float32(temp_0)), fh1_0=float(np.float32(fh1_0)),
                delta_nl=[float(np.float32(v)) for v in delta_nl])

"""The checker's write_otf averages (io_funcs.cc:500-632) pinned on the CPU
by numpy's float64 means over the same cells: volume averages, mass-weighted
averages, the "binary" and mean-density selections, extremes and the photon
count.  A swapped weight, a dropped selection or a wrong normalization moves
an average far beyond the 1e-6 allowed for float rounding."""
import math

import numpy as np

from flexrt_imaging_reion_amd import inputs
from oracle.pyoracle import Oracle

F32 = np.float32


def close(a, b, rel=1e-6):
    if isinstance(a, float) and math.isnan(a):
        return math.isnan(b)
    return abs(a - b) <= rel * abs(b)


def test_checker_otf_against_numpy():
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=3)
    o = Oracle(wl["cfg"])
    try:
        inputs.setup(o, wl)
        for _ in range(3):
            o.step()
        f = {k: o.download_field(k).astype(np.float64) for k in
             ("f_H1", "f_H2", "f_He1", "nH", "nHe", "nH1", "nH2", "temp", "gamma_H1_tot", "mfp", "heat_rate",
              "recrate_H2", "rho", "avg_nh1")}
        G = f["gamma_H1_tot"]
        o.upload_field("avg_gamma_prev", (G * 1.5).astype(F32))  # within a decade: binary = (G >= 3e-14)
        avg_rho = F32(f["rho"].mean())
        r = o.otf_averages(avg_rho, 100.0)
        assert close(r["fH1"][0], f["f_H1"].mean())
        assert close(r["fH1"][1], (f["f_H1"] * f["nH"]).sum() / f["nH"].sum())
        assert close(r["fHe1"][1], (f["f_He1"] * f["nHe"]).sum() / f["nHe"].sum())
        assert close(r["nH1"], f["avg_nh1"].mean())
        b = G.astype(F32) >= F32(3e-14)
        assert 0 < b.sum() < b.size, "the selection must be partial"
        assert close(r["nH"][2], f["nH"][b].mean())
        assert close(r["T"][2], f["temp"][b].mean())
        assert close(r["GH1"][3], (G * f["nH"]).sum() / f["nH"].sum())
        assert close(r["GH1"][4], (G * f["nH1"]).sum() / f["nH1"].sum())
        assert r["GH1"][5] == F32(G.min()) and r["GH1"][6] == F32(G.max())
        assert close(r["RH2"][4], (f["recrate_H2"] * f["nH2"]).sum() / f["nH2"].sum())
        assert close(r["MFP"][3], (f["mfp"] * f["nH"]).sum() / f["nH"].sum())
        assert r["MFP"][4] == F32(f["mfp"].min()) and r["MFP"][5] == F32(f["mfp"].max())
        assert close(r["Heat"][3], (f["heat_rate"] * f["nH"]).sum() / f["nH"].sum())
        rr = f["rho"].astype(F32) / avg_rho
        m = (rr > 0.9) & (rr < 1.1)
        assert 0 < m.sum() < m.size
        t0 = np.where(m, f["temp"] * f["f_H2"] + 100.0 * (1 - f["f_H2"]), 100.0)
        assert close(r["T"][4], t0[m].mean(), rel=1e-5)
        rays = o.rays()
        nph = rays["nphot"].astype(np.float64).sum() * 1e60
        assert close(r["npht"], nph)
        assert r["nsrc"] == len(wl["src_cells"])
        assert r["fion"] == [1.0, 1.0]
    finally:
        o.close()


def test_checker_otf_empty_binary():
    """A fresh start has avg_gamma_prev = 0 (init_funcs.cc:118): no cell is
    "binary", and its averages are 0/0 as in the reference."""
    wl = inputs.make_workload("lognormal64", n=16, n_src=8, steps=1)
    o = Oracle(wl["cfg"])
    try:
        inputs.setup(o, wl)
        o.step()
        r = o.otf_averages(F32(o.download_field("rho").mean()))
        assert math.isnan(r["nH"][2]) and math.isnan(r["GH1"][2])
    finally:
        o.close()

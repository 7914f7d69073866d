"""Synthetic inputs: spectrum binning, density field, sources, gas init.

tests/golden/spectrum_tables_5bins.txt is the reference's own spectrum table
(input_files/spectrum_tables_5bins.txt): rows of photon fractions in 5 bins
for spectral indices -0.5..3.0.
"""
import os

import numpy as np
import pytest

from flexrt_imaging_reion_amd import inputs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "spectrum_tables_5bins.txt")
H_EV = 4.135667e-15


def load_table():
    rows = [l.split() for l in open(GOLDEN) if l.strip()]
    energies = np.array([float(x) for x in rows[0]])
    alpha = np.array([float(r[0]) for r in rows[1:]])
    table = np.array([[float(x) for x in r[1:]] for r in rows[1:]])
    return energies, alpha, table


def test_set_spectrum_table_interpolation():
    # set_spectrum (gas_funcs.cc:44-70): at a tabulated index it returns that row,
    # normalized; alpha = 1.5 gives the equal-photon bins of the reference table
    e, alpha, table = load_table()
    s = inputs.spectrum_from_table(alpha, table, 1.5)
    assert np.allclose(s, 0.2, atol=1e-7)
    for k in (0, 3, 7, 13):
        s = inputs.spectrum_from_table(alpha, table, alpha[k])
        assert np.allclose(s, table[k] / table[k].sum(), rtol=1e-6)
    mid = inputs.spectrum_from_table(alpha, table, 0.5 * (alpha[5] + alpha[6]))
    lin = 0.5 * (table[5] + table[6])
    assert np.allclose(mid, lin / lin.sum(), rtol=1e-6)


def test_powerlaw_bins():
    nu, w = inputs.powerlaw_bins()
    assert w.dtype == np.float32 and abs(float(w.sum()) - 1.0) < 1e-6
    e = nu * H_EV
    edges = np.geomspace(13.6, 54.4, 6)
    assert np.all(e > edges[:-1]) and np.all(e < edges[1:])
    # weights fall as nu^-1.5 photon counts: each bin's share of int nu^-2.5
    ratio = (edges[:-1] ** -1.5 - edges[1:] ** -1.5)
    assert np.allclose(w, ratio / ratio.sum(), rtol=1e-6)
    # sigma at nu_b equals the photon-weighted bin average of the Verner fit
    for b in range(5):
        x = np.geomspace(edges[b], edges[b + 1], 20001)
        wgt = x ** -2.5
        avg = np.trapezoid(inputs.sigma_hi_verner(x) * wgt, x) / np.trapezoid(wgt, x)
        assert inputs.sigma_hi_verner(e[b]) == pytest.approx(avg, rel=1e-6)


def test_lognormal_field():
    opd = inputs.lognormal_overdensity(32, 1.0, 7)
    assert opd.shape == (32, 32, 32)
    assert abs(opd.mean() - 1.0) < 1e-12
    assert np.std(np.log(opd)) == pytest.approx(1.0, rel=1e-6)
    g = inputs.gaussian_field(32, 7)
    # P(k) ~ k^-2: power in the lowest shell exceeds the highest by orders of magnitude
    pk = np.abs(np.fft.rfftn(g)) ** 2
    k = np.sqrt(np.fft.fftfreq(32)[:, None, None] ** 2 + np.fft.fftfreq(32)[None, :, None] ** 2 +
                np.fft.rfftfreq(32)[None, None, :] ** 2)
    lo = pk[(k > 0.02) & (k < 0.05)].mean()
    hi = pk[(k > 0.4) & (k < 0.45)].mean()
    assert 20 < lo / hi < 2000


def test_sources_and_gridding():
    opd = inputs.lognormal_overdensity(16, 1.0, 3)
    cells, lum = inputs.sources_weighted(opd, 500, 51, 56, 3)
    assert cells.min() >= 0 and cells.max() < 16 ** 3
    assert np.all((lum >= 1e51) & (lum <= 1e56))
    gc, gl = inputs.grid_sources(cells, lum, 16 ** 3)
    assert np.all(np.diff(gc) > 0)
    assert gl.sum() == pytest.approx(lum.sum(), rel=1e-12)
    dc, dl = inputs.sources_densest(opd, 10, 52, 55, 3)
    assert set(dc) == set(np.argsort(-opd.reshape(-1))[:10])


def test_init_gas_fresh_start():
    rho = np.full(8, 1e-28, np.float32)
    f = inputs.init_gas(rho, np.float32(100.0), np.float32(3.0))
    assert np.all(f["nH1"] + f["nH2"] == f["nH"])
    assert np.allclose(f["nH2"] / f["nH"], 1.2e-3, rtol=1e-4)
    assert np.all(f["ne"] == f["nH2"] + f["nHe2"])
    assert f["nHe"][0] == pytest.approx(f["nH"][0] * 0.246 / (4 * 0.754), rel=1e-6)


def test_workloads_build():
    wl = inputs.make_workload("uniform32")
    assert wl["cfg"]["path_length"] == 10 and len(wl["src_cells"]) == 1
    wl = inputs.make_workload("batch128", n=16, n_src=64)
    assert wl["cfg"]["n"] == 16 and len(wl["src_cells"]) <= 64
    assert np.all(wl["fields"]["clump"] == 3.0)

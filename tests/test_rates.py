"""Pins of the rate fits on the RT path (rates.cc restated in the CPU checker).

Known answers from the literature and closed forms derived here from physical
constants, independent of the fit coefficients being tested.
"""
import math

import pytest

from oracle import pyoracle as P

H_EV = 4.135667e-15
K_B = 1.380649e-16      # CODATA, independent of the code's 1.38065e-16
M_E = 9.1093837e-28
C_LIGHT = 2.99792458e10
E_CHARGE = 4.80320471e-10
H_PLANCK = 6.62607015e-27
SIGMA_T = 6.6524587e-25
A_RAD = 7.5657e-15


def test_sigma_hi_threshold():
    # Verner et al. 1996: sigma(HI) at 13.6 eV = 6.30e-18 cm^2
    assert P.rate("sigmapi_H1", 13.6 / H_EV) == pytest.approx(6.30e-18, rel=5e-3)
    # zero below threshold, and the hydrogenic ~E^-3 fall-off between 20 and 40 eV
    assert P.rate("sigmapi_H1", 13.5 / H_EV) == 0.0
    s20, s40 = P.rate("sigmapi_H1", 20 / H_EV), P.rate("sigmapi_H1", 40 / H_EV)
    slope = math.log(s40 / s20) / math.log(2.0)
    assert -3.3 < slope < -2.6


def test_sigma_hei_threshold():
    # Verner et al. 1996: sigma(HeI) at 24.59 eV = 7.42e-18 cm^2
    assert P.rate("sigmapi_He1", 24.59 / H_EV) == pytest.approx(7.42e-18, rel=1e-2)
    assert P.rate("sigmapi_He1", 24.5 / H_EV) == 0.0


def test_recombination_coefficients_at_1e4K():
    # Osterbrock & Ferland: alpha_B(HII, 1e4 K) = 2.59e-13, alpha_A = 4.18e-13,
    # alpha_B(HeII, 1e4 K) = 2.73e-13 cm^3/s
    assert P.rate("alphaB_H2", 1e4) == pytest.approx(2.59e-13, rel=2e-2)
    assert P.rate("alphaA_H2", 1e4) == pytest.approx(4.18e-13, rel=2e-2)
    assert P.rate("alphaB_He2", 1e4) == pytest.approx(2.73e-13, rel=5e-2)
    # alpha_B ~ T^-0.7 around 1e4 K
    s = math.log(P.rate("alphaB_H2", 2e4) / P.rate("alphaB_H2", 5e3)) / math.log(4.0)
    assert -0.85 < s < -0.6


def test_recombination_cooling_case_b():
    # mean energy carried off per case-B recombination at 1e4 K is ~0.68 kT
    # (Osterbrock & Ferland table 3.2: beta_B / alpha_B = 0.684 kT)
    T = 1e4
    ratio = P.rate("rcrB_H2", T) / (K_B * T * P.rate("alphaB_H2", T))
    assert 0.62 < ratio < 0.76
    # HeII recombination cooling is kT alpha_B(HeII) by definition in the fit
    assert P.rate("rcrB_He2", T) == pytest.approx(1.38065e-16 * T * P.rate("alphaB_He2", T), rel=1e-14)


def test_free_free_from_constants():
    # thermal bremsstrahlung, Z = 1: (2^5 pi e^6 / 3 h m_e c^3) sqrt(2 pi k T / 3 m_e) g_ff
    pref = (2 ** 5 * math.pi * E_CHARGE ** 6) / (3 * H_PLANCK * M_E * C_LIGHT ** 3) * \
        math.sqrt(2 * math.pi * K_B / (3 * M_E))
    for T in (1e3, 1e4, 1e5):
        assert P.rate("free_free_rate", T) == pytest.approx(pref * math.sqrt(T) * 1.30, rel=1e-2)


def test_compton_from_constants():
    # Compton coupling to the CMB: 4 sigma_T a T_cmb^4 k_B / (m_e c) (1+z)^4
    coef = 4 * SIGMA_T * A_RAD * 2.726 ** 4 * K_B / (M_E * C_LIGHT)
    for z in (0.0, 6.5, 7.0):
        assert P.rate("compton_rate", z) == pytest.approx(coef * (1 + z) ** 4, rel=1e-2)


def test_dielectronic_heii():
    # Aldrovandi & Pequignot 1973: 1.9e-3 T^-1.5 exp(-4.7e5/T) (1 + 0.3 exp(-9.4e4/T))
    for T in (5e4, 1e5, 5e5):
        ap = 1.9e-3 * T ** -1.5 * math.exp(-4.7e5 / T) * (1 + 0.3 * math.exp(-9.4e4 / T))
        assert P.rate("Dalpha_He2", T) == pytest.approx(ap, rel=3e-2)


def test_hubble_rate():
    # H(a=1) = 100 h km/s/Mpc; matter + Lambda scaling with a
    h = 0.7
    assert P.hubble(1.0, h, 0.3, 0.7) == pytest.approx(h * 100 / 3.086e19, rel=1e-6)
    a = 1 / 8.0
    expect = h * 100 / 3.086e19 * math.sqrt(0.3 / a ** 3 + 0.7)
    assert P.hubble(a, h, 0.3, 0.7) == pytest.approx(expect, rel=1e-6)

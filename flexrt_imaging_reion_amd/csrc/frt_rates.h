// frt_rates.h -- atomic rates and cross-sections on the RT path, host and device.
//
// Closed-form fits evaluated exactly as rates.cc writes them (Verner et al.
// 1996 photo-ionization cross-sections; Hui & Gnedin 1997 recombination and
// cooling coefficients), with libm replaced by frt_detmath.h.
#pragma once

#include "frt_detmath.h"

// Physical constants, global_constants.h:8-45.
#define FRT_PI_REF 3.14159
#define FRT_H_PLANCK 6.626068e-27
#define FRT_H_EV 4.135667e-15
#define FRT_K_B 1.38065e-16
#define FRT_KPC_TO_CM 3.086e21
#define FRT_CL 2.998e10
#define FRT_Y_HE 0.246
#define FRT_EV_CGS 1.60218e-12
#define FRT_EV_TO_ERG 1.6022e-12
#define FRT_MPC_TO_KM 3.086e19
#define FRT_YR_TO_S 3.154e7

// Verner et al. 1996 fit parameters, rates.cc:5-39.
#define FRT_ETH_H1 13.60
#define FRT_ETH_HE1 24.59
#define FRT_ETH_HE2 54.42

// x = 2 E_th / (k_B T) with E_th in eV (rates.cc:41-48).
FRT_HD double frt_lambda(double T, double Eth) {
  double K2eV = FRT_K_B / FRT_EV_CGS;
  double eV2K = 1.0 / K2eV;
  return 2 * Eth * eV2K / T;
}

// Verner fit shape F(y) (rates.cc:50-66).
FRT_HD double frt_F_of_y(double nu, double Eth, double Emax, double E0, double y0, double y1,
                         double yw, double ya, double P) {
  double E = FRT_H_EV * nu;
  double x = E / E0 - y0;
  double y = dm_pow(dm_pow(x, 2) + dm_pow(y1, 2), 0.5);
  double F = (dm_pow(x - 1, 2) + dm_pow(yw, 2)) * dm_pow(y, 0.5 * P - 5.5) *
             dm_pow(1 + dm_pow(y / ya, 0.5), -P);
  if (FRT_H_EV * nu >= Eth && FRT_H_EV * nu <= Emax) return F;
  return 0;
}

// sigmapi_H1 / sigmapi_He1 in cm^2 (rates.cc:71-78).
FRT_HD double frt_sigmapi_H1(double nu) {
  return 5.475e-14 * frt_F_of_y(nu, FRT_ETH_H1, 5e+4, 4.298e-1, 0, 0, 0, 3.288e+1, 2.963);
}
FRT_HD double frt_sigmapi_He1(double nu) {
  return 9.492e-16 * frt_F_of_y(nu, FRT_ETH_HE1, 5e+4, 1.361e+1, 4.434e-1, 2.136, 2.039, 1.469, 3.188);
}
FRT_HD double frt_sigmapi_He2(double nu) {  // rates.cc:81-83, the c2ray_off mode's update_gamma
  return 1.369e-14 * frt_F_of_y(nu, FRT_ETH_HE2, 5e+4, 1.720e0, 0, 0, 0, 3.288e+1, 2.963);
}

// Recombination coefficients in cm^3/s (rates.cc:128-171).
FRT_HD double frt_alphaA_H2(double T) {
  double x = frt_lambda(T, FRT_ETH_H1);
  return 1.269e-13 * dm_pow(x, 1.503) / dm_pow((1.0 + dm_pow((x / 0.522), 0.470)), 1.923);
}
FRT_HD double frt_alphaB_H2(double T) {
  double x = frt_lambda(T, FRT_ETH_H1);
  return 2.753e-14 * dm_pow(x, 1.500) / dm_pow((1.0 + dm_pow((x / 2.740), 0.407)), 2.2420);
}
FRT_HD double frt_alphaA_He2(double T) {
  double x = frt_lambda(T, FRT_ETH_HE1);
  return 3.0e-14 * dm_pow(x, 0.654);
}
FRT_HD double frt_alphaB_He2(double T) {
  double x = frt_lambda(T, FRT_ETH_HE1);
  return 1.26e-14 * dm_pow(x, 0.750);
}
// Hui & Gnedin 1997 He III recombination, rates.cc:154-166 (the c2ray_off mode).
FRT_HD double frt_alphaA_He3(double T) {
  double x = frt_lambda(T, FRT_ETH_HE2);
  return 2. * 1.269e-13 * dm_pow(x, 1.503) / dm_pow((1.0 + dm_pow((x / 0.522), 0.470)), 1.923);
}
FRT_HD double frt_alphaB_He3(double T) {
  double x = frt_lambda(T, FRT_ETH_HE2);
  return 2 * 2.753e-14 * dm_pow(x, 1.500) / dm_pow((1.0 + dm_pow((x / 2.740), 0.407)), 2.2420);
}
FRT_HD double frt_Dalpha_He2(double T) {
  double x = frt_lambda(T, FRT_ETH_HE2);
  return 1.90e-3 * dm_pow(T, -1.5) * dm_exp(-0.75 * x / 2) * (1 + 0.3 * dm_exp(-0.15 * x / 2));
}

// Cooling rates in erg cm^3/s (rates.cc:176-231).
FRT_HD double frt_rcrA_H2(double T) {
  double x = frt_lambda(T, FRT_ETH_H1);
  return 1.778e-29 * T * dm_pow(x, 1.965) / (dm_pow(1. + dm_pow(x / 0.541, 0.502), 2.697));
}
FRT_HD double frt_rcrB_H2(double T) {
  double x = frt_lambda(T, FRT_ETH_H1);
  return 3.435e-30 * T * dm_pow(x, 1.970) / (dm_pow(1. + dm_pow(x / 2.250, 0.376), 3.720));
}
FRT_HD double frt_rcrA_He2(double T) { return FRT_K_B * T * frt_alphaA_He2(T); }
FRT_HD double frt_rcrB_He2(double T) { return FRT_K_B * T * frt_alphaB_He2(T); }
FRT_HD double frt_free_free_rate(double T) { return 1.42e-27 * dm_pow(T, 0.5) * 1.30; }
FRT_HD double frt_compton_rate(double z) { return 5.6072e-36 * dm_pow(1 + z, 4.); }

// Hubble rate H(a) in 1/s, float result as cosmo_funcs.cc:13-15.
FRT_HD float frt_hubble(float a, double hh, double om, double ol) {
  return (float)(hh * 100. / FRT_MPC_TO_KM * dm_pow(om / dm_pow(a, 3.) + ol, 0.5));
}

// The reference's float min/max (general_funcs.cc:65-82): arguments narrow to float.
FRT_HD float frt_mind(float a, float b) { return (a >= b) ? b : a; }
FRT_HD float frt_maxd(float a, float b) { return (b >= a) ? b : a; }
FRT_HD float frt_absd(float x) { return (x >= 0.f) ? x : (float)(-1. * x); }

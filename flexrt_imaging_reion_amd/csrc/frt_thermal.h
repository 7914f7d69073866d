// frt_thermal.h -- the temperature update of one cell, shared by the step
// kernels and the adjoint (which replays it to get the sub-step schedule).
#pragma once

#include "frt_internal.h"

// ---------------------------------------------------------------------------
// update_thermal for one cell (gas_funcs.cc:362-419 with set_Treion_default
// :104-129, update_heat_cool :134-242, get_thermal_tstep :353-358), subgrid off.
struct ThermalIn {
  float nH, nHe, nH1, nH2, nHe2, ne, ntot, dne, fH2, ah1, ahe1, nH1_prev;
  float ddelta;   // ddelta_dt (0 until a hydro-step reload sets it)
  float rho_rel;  // rho / avg_rho (float division)
};

__device__ __forceinline__ void thermal_density_terms(const CellPtrs& cf, const GridConst& g, int64_t k,
                                                      ThermalIn& in) {
  in.ddelta = cf.f[C_DDELTA_DT][k];
  in.rho_rel = in.ddelta != 0.f ? cf.f[C_RHO][k] / g.avg_rho : 1.f;
}

// Sub-steps of one cell's update, kept for the adjoint: the temperature
// entering each sub-step and its length, the temperature after the
// post-front mixing, and that before the final floor.
struct ThermalRec {
  static constexpr int MAX = 64;
  float t_in[MAX];
  float dt[MAX];
  float t_mix;
  float t_raw;
  float heat;  // photo-heating (the same in every sub-step)
};

template <bool REC = false>
__device__ __forceinline__ int thermal_cell(const GridConst& g, const ThermalIn& in,
                                            const float* gh1nu, const float* ghe1nu,
                                            int64_t n3, int64_t k, float& temp, float& heat,
                                            float& cool, float& if_speed, float& inc_flux,
                                            ThermalRec* rec = nullptr) {
  // set_Treion_default
  const float fion_new = (in.nH - in.nH1) / in.nH;
  const float fion_old = (in.nH - in.nH1_prev) / in.nH;
  const float deltax = g.dx * (fion_new - fion_old);
  const float vIF = deltax / g.dt;
  const float Cn[5] = {9.5432f, -1.6441e-2f, -9.8010e-3f, 4.1664e-3f, -2.2710e-4f};
  const double L = dm_log((double)vIF / 1e5);
  float lnT = 0.f;
  for (int t = 0; t < 5; ++t) lnT = (float)((double)lnT + (double)Cn[t] * dm_pow(L, (double)t));
  if ((fion_new > fion_old) && ((double)fion_new < 0.98))
    temp = (fion_old * temp + (fion_new - fion_old) * dm_expf(lnT)) / fion_new;
  if (REC) rec->t_mix = temp;
  const float vpos = frt_maxd(0.f, vIF);
  if_speed = vpos;
  inc_flux = (float)((1. + (double)g.nhe_to_nh) * (double)in.nH * (double)vpos);

  const double k23 = 2. / 3. / FRT_K_B / (double)in.ntot;
  const double Hd = (double)g.hubble;
  const float dne_over_n = in.dne / in.ntot;
  const float min_dt = (float)(0.1 * (double)g.dt);
  const bool photoheat = (double)in.fH2 > 0.98;
  float t_temp = 0.f;
  int nsub = 0;
  while (t_temp < g.dt) {
    // get_thermal_tstep uses the heating/cooling of the previous sub-step
    const float dTdt = (float)(k23 * (double)(heat - cool) - (2. * Hd + (double)dne_over_n) * (double)temp);
    const float tstep = frt_maxd(min_dt, frt_mind((float)(0.1 * (double)temp / (double)frt_absd(dTdt)), g.dt));
    const float dt_temp = frt_mind(tstep, g.dt - t_temp);
    // update_heat_cool
    heat = 0.f;
    cool = 0.f;
    if (photoheat) {
      for (int nu = 0; nu < g.nfreq; ++nu) {
        heat = (float)((double)heat + (double)(in.ah1 * gh1nu[nu * n3 + k]) * FRT_H_PLANCK *
                                          (FRT_H_EV * g.nu[nu] - FRT_ETH_H1) / FRT_H_EV);
        heat = (float)((double)heat + (double)(in.ahe1 * ghe1nu[nu * n3 + k]) * FRT_H_PLANCK *
                                          (FRT_H_EV * g.nu[nu] - FRT_ETH_HE1) / FRT_H_EV);
      }
    }
    const double Tm = (double)frt_maxd(g.temp_min, temp);
    if (g.recomb_cool) {
      if (g.rec_case_b) {
        cool = (float)((double)cool + frt_rcrB_H2(Tm) * (double)in.ne * (double)in.nH2);
        cool = (float)((double)cool + frt_rcrB_He2(Tm) * (double)in.ne * (double)in.nHe2);
      } else {
        cool = (float)((double)cool + frt_rcrA_H2(Tm) * (double)in.ne * (double)in.nH2);
        cool = (float)((double)cool + frt_rcrA_He2(Tm) * (double)in.ne * (double)in.nHe2);
      }
    }
    if (g.free_free) {
      const float nff = in.nH2 + in.nHe2;  // + 4 nHe3, nHe3 = 0 in C2Ray mode
      cool = (float)((double)cool + frt_free_free_rate(Tm) * (double)in.ne * (double)nff);
    }
    if (g.compton) cool = (float)((double)cool - g.compton_coef * (double)in.ne * (g.t_cmb - Tm));
    double den = 1. + 2. * Hd * (double)dt_temp + 1. / (double)in.ntot * (double)in.dne * (double)dt_temp;
    // - 2 ddelta_dt/3/(rho/avg_rho) dt_temp: skipped while ddelta_dt is 0, where
    // it subtracts zero (and where rho may not have been uploaded at all)
    if (in.ddelta != 0.f) den = den - 2. * (double)in.ddelta / 3. / (double)in.rho_rel * (double)dt_temp;
    if (REC) {
      rec->t_in[nsub] = temp;
      rec->dt[nsub] = dt_temp;
      rec->heat = heat;
    }
    temp = (float)(((double)temp + k23 * (double)(heat - cool) * (double)dt_temp) / den);
    t_temp += dt_temp;
    if (++nsub >= ThermalRec::MAX) break;
  }
  if (REC) rec->t_raw = temp;
  temp = frt_maxd(temp, g.temp_min);
  return nsub;
}


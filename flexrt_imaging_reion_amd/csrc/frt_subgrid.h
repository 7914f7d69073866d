// frt_subgrid.h -- the sub-grid I-front model on the device, included by
// frt_kernels.hip inside namespace frt (it uses that file's launch helpers).
//
// The reference's production setting (user_inputs.h:11-15): subgrid, ifront
// and equilibrium TRUE, relaxed FALSE, temp_ev TRUE, FP64 form.  Per C2Ray
// pass:
//   k_sg_reset     cell reset (ray_tracing.cc:630-697) and, on pass 0, the
//                  RadHydro interpolation of ionized cells (get_gamma_interp)
//                  or the neutral MFP guess;
//   k_sg_walk      the ray loop (:701-938) of every moved ray: attenuation by
//                  the MFP of the ionized part, the photon record of each
//                  segment, and a stop at each segment in a cell with
//                  fion < 1, whose neutral part absorbs photons first come
//                  first served (:867-919);
//   k_sg_resolve   one thread per such cell walks its segments in slot
//                  order (= ray list order) and hands out the room left in
//                  the neutral part to every segment whose ray has reached
//                  it, stopping at the first one that has not;
//   (k_sg_walk and k_sg_resolve alternate until no ray waits: the rounds
//   reproduce the serial list-order count exactly, whatever the schedule)
//   k_sg_chem      the cell loop's subgrid branch (:954-1064) from the photon
//                  records, and on the last pass update_fion, the ionization
//                  state, the rates and (fused) update_thermal's subgrid branch.

namespace sgd {
constexpr float sigma_eff_rh = 2.517139346618065e-18f;  // global_variables.cc:241
constexpr float max_subgrid_frac = 0.49f;               // global_variables.cc:245
constexpr float gamma_min = 1e-16f;                     // global_variables.cc:303
// global_variables.cc:302 sets mfp_min = 1e-2*dx[0] while dx[] is still zero
constexpr float mfp_min = 0.f;
__constant__ float gamma_dom[5] = {3e-15f, 3e-14f, 3e-13f, 3e-12f, 3e-11f};
}  // namespace sgd

// log10 of a float argument, as std::log10(float) in subgrid.cc: in double, rounded
__device__ __forceinline__ float sg_lg10f(float x) { return (float)dm_log10((double)x); }

// Row m of table (zre im, delta ip, gamma node iq = 1..3): file gamma k sits at
// index k + 1 of the reference's arrays (subgrid.cc:22).
__device__ __forceinline__ int64_t sg_row(const SgConst& c, int m, int im, int ip, int iq) {
  return (int64_t)((im * 3 + ip) * 3 + (iq - 1)) * c.nt + m;
}

// get_tradhydro_index (subgrid.cc:187-190)
__device__ __forceinline__ int sg_tindex(const SgConst& c, float tsince) {
  int idx = (int)floorf(tsince / c.t_norm * (float)(c.nt - 1));
  return max(0, min(idx, c.nt - 2));
}

// deltaNL_to_deltasig (subgrid.cc:193-204)
__device__ __forceinline__ float sg_deltasig(float deltaNL, float redshift) {
  const float dcmode = 5.10655f;
  const float sigma0 = (float)((double)dcmode / dm_pow(3., 0.5));
  const float d_of_z = (float)(1. / (double)(1 + redshift) / 0.78205);
  const double o = 1. + (double)deltaNL;
  const float delta = (float)(-1.35 * dm_pow(o, -2. / 3.) + 0.78785 * dm_pow(o, -0.58661) -
                              1.12431 * dm_pow(o, -0.5) + 1.68647);
  return delta / d_of_z / sigma0;
}

// Linear interpolation in time from row it, clamped to [0, 1] (subgrid.cc:345-357).
__device__ __forceinline__ float sg_lin_fh1(const SgPtrs& sp, const SgConst& c, int it, int im, int ip, int iq,
                                            float ts) {
  const int64_t r = sg_row(c, it, im, ip, iq);
  const float a = sp.fh1_rh[r], b = sp.fh1_rh[r + 1];
  const float ta = sp.time_rh[r], tb = sp.time_rh[r + 1];
  const float v = a + (b - a) / (tb - ta) * (ts - ta);
  return frt_mind(frt_maxd(v, 0.f), 1.f);
}

// get_interpolate_gamma_f, fractype TRUE (subgrid.cc:210-251), and the
// fixed-slope end nodes (:445-446).
__device__ void sg_interp_gamma(const SgConst& c, float tsince, float tre, float deltaNL, const int it[2],
                                const int id[2], const float f[2][2][3], float out[5]) {
  const float lt = sg_lg10f(tre + tsince);
  const float lt0 = sg_lg10f(c.tre_dom[it[0]] + tsince);
  const float lt1 = sg_lg10f(c.tre_dom[it[1]] + tsince);
  const double ld = dm_log10(1. + (double)deltaNL);
  const double ld0 = dm_log10(1. + (double)c.delta_nl[id[0]]);
  const double ld1 = dm_log10(1. + (double)c.delta_nl[id[1]]);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float ft[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const float a = sg_lg10f(f[0][p][q]), b = sg_lg10f(f[1][p][q]);
      ft[p] = a + (b - a) * (lt - lt0) / (lt1 - lt0);
    }
    const float v = (float)((double)ft[0] + (double)(ft[1] - ft[0]) * (ld - ld0) / (ld1 - ld0));
    out[q + 1] = frt_mind((float)dm_pow(10., (double)v), sgd::max_subgrid_frac);
  }
  out[0] = (float)(dm_pow((double)(sgd::gamma_dom[0] / sgd::gamma_dom[1]), 0.) * (double)out[1]);
  out[4] = (float)(dm_pow((double)(sgd::gamma_dom[4] / sgd::gamma_dom[3]), -1.) * (double)out[3]);
}

// get_gamma_interp (subgrid.cc:257-550), its fh1_gamma_interp nodes (the MFP
// and temperature nodes only feed interp_mfp and interp_temp, which this
// setting does not call); the crossing-time bins are evaluated in bin order
// and summed as they come.
__device__ void sg_gamma_interp_fh1(const SgPtrs& sp, const SgConst& c, float zz, int64_t n3, int64_t k,
                                    float tre, float deltaNL, float tsince) {
  constexpr int NT = FRT_SG_NTCROSS;
  tre = frt_maxd(c.t_z12, tre);
  int it[2], id[2];
  if (tre > c.t_z8) { it[1] = 2; it[0] = 1; } else { it[1] = 1; it[0] = 0; }
  if (sg_deltasig(deltaNL, zz) < 0.) { id[1] = 1; id[0] = 0; } else { id[1] = 2; id[0] = 1; }
  const int itf = sg_tindex(c, tsince);
  float f[2][2][3];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) f[m][p][q] = sg_lin_fh1(sp, c, itf, it[m], id[p], q + 1, tsince);
  float nodes[5];
  sg_interp_gamma(c, tsince, tre, deltaNL, it, id, f, nodes);
  if (sp.tcross_flag[k] != 0.f) {
    // crossing-time bins (:481-547): the nodes of every bin, then the
    // reference's average over the bins the I-front has crossed
    float nn[NT][5];
    float prev = 0.f;
    for (int tt = 0; tt < NT; ++tt) {
      const float tcn = sp.tcross_new[(int64_t)tt * n3 + k];
      const float mid = (tt == 0) ? (float)((double)tcn / 2.) : (float)((double)prev + (double)(tcn - prev) / 2.);
      prev = tcn;
      int itc[2];
      if (tre + tcn > c.t_z8) { itc[1] = 2; itc[0] = 1; } else { itc[1] = 1; itc[0] = 0; }
      const int itftc = sg_tindex(c, tsince - mid);
      float ftc[2][2][3];
      for (int m = 0; m < 2; ++m)
        for (int p = 0; p < 2; ++p)
          for (int q = 0; q < 3; ++q) ftc[m][p][q] = sg_lin_fh1(sp, c, itftc, itc[m], id[p], q + 1, tsince - mid);
      sg_interp_gamma(c, tsince - mid, tre + mid, deltaNL, itc, id, ftc, nn[tt]);
    }
    const float fmax = sp.fion_max[k];
    const float tc_frac = (float)(1. / ((float)NT));
    for (int q = 0; q < 5; ++q) {
      float acc = 0.f;
      if (fmax == 1.) {
        for (int tt = 0; tt < NT; ++tt) acc += tc_frac * nn[tt][q];
      } else {
        for (int tt = 0; tt < NT; ++tt) {
          if ((fmax >= tt * tc_frac) && ((double)fmax < (tt + 1.) * (double)tc_frac)) {
            for (int ttt = 0; ttt < tt; ++ttt) acc += tc_frac * nn[ttt][q];
            acc += (frt_maxd(0.01f, fmax) - tt * tc_frac) * nn[tt][q];
          }
        }
      }
      nodes[q] = frt_mind(acc / frt_maxd(0.01f, fmax), sgd::max_subgrid_frac);
    }
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) sp.fh1_nodes[(int64_t)q * n3 + k] = nodes[q];
}

// interp_nh1 (subgrid.cc:616-627): the HI fraction, log-log in Gamma.
__device__ __forceinline__ float sg_interp_xh1(const SgPtrs& sp, int64_t n3, int64_t k, float gammaH1, float fh1_0) {
  const float gamma = frt_maxd(gammaH1, sgd::gamma_min);
  int gi = (int)floor(((float)3 + 2 - 1.) * (double)(sg_lg10f(gamma) - sg_lg10f(sgd::gamma_dom[0])) /
                      (double)(sg_lg10f(sgd::gamma_dom[4]) - sg_lg10f(sgd::gamma_dom[0])));
  gi = max(0, min(gi, 3));
  const float f0 = sp.fh1_nodes[(int64_t)gi * n3 + k], f1 = sp.fh1_nodes[(int64_t)(gi + 1) * n3 + k];
  const float v = sg_lg10f(f0) + (sg_lg10f(f1) - sg_lg10f(f0)) /
                                     (sg_lg10f(sgd::gamma_dom[gi + 1]) - sg_lg10f(sgd::gamma_dom[gi])) *
                                     (sg_lg10f(gamma) - sg_lg10f(sgd::gamma_dom[gi]));
  return frt_maxd(frt_mind((float)dm_pow(10., (double)v), fh1_0), 0.f);
}

// Equilibrium MFPs of an ionized cell (ray_tracing.cc:981-986).
__device__ __forceinline__ void sg_mfp_eq(const GridConst& g, float ag, float clump, float temp, float nH, float& mfp,
                                          float& mfp912) {
  const double ab = frt_alphaB_H2((double)frt_maxd(g.temp_min, temp));
  const double n2 = dm_pow((double)nH, 2.);
  const float gg = frt_maxd(sgd::gamma_min, ag) / clump;
  mfp = (float)((double)(gg / sgd::sigma_eff_rh) / ab / n2 / 1.082);
  mfp912 = (float)((double)gg / g.sig912 / ab / n2 / 1.082);
  mfp = frt_maxd(sgd::mfp_min, mfp);
  mfp912 = frt_maxd(sgd::mfp_min, mfp912);
}

// Photoheat_HIHeI (rates.cc:240-254) for a power law of index alpha.
__device__ __forceinline__ double sg_photht_hi(double alpha, double T, double ne) {
  return FRT_ETH_H1 / FRT_H_EV * FRT_H_PLANCK *
         ((alpha + 2.8) / (1.8 + alpha) * (1. - dm_pow(4., -1.8 - alpha)) / (1. - dm_pow(4., -2.8 - alpha)) - 1.) *
         frt_alphaA_H2(T) * ne;
}
__device__ __forceinline__ double sg_photht_hei(double alpha, double T, double ne) {
  return FRT_ETH_HE1 / FRT_H_EV * FRT_H_PLANCK *
         ((alpha + 1.7) / (0.7 + alpha) * (1. - dm_pow(2.2, -0.7 - alpha)) / (1. - dm_pow(2.2, -1.7 - alpha)) - 1.) *
         frt_alphaA_He2(T) * ne;
}

// update_thermal for one ionized cell, subgrid branches of update_heat_cool
// (gas_funcs.cc:171-241) and the same sub-stepping as the C2Ray mode.
__device__ int sg_thermal_cell(const GridConst& g, float nH, float nHe, float ntot, float dne, float fion,
                               float gh1tot, float ddelta, float rho_rel, float& temp, float& heat, float& cool) {
  const double k23 = 2. / 3. / FRT_K_B / (double)ntot;
  const double Hd = (double)g.hubble;
  const float dne_over_n = dne / ntot;
  const float min_dt = (float)(0.1 * (double)g.dt);
  const float nn = nH + nHe;
  float t_temp = 0.f;
  int nsub = 0;
  while (t_temp < g.dt) {
    const float dTdt = (float)(k23 * (double)(heat - cool) - (2. * Hd + (double)dne_over_n) * (double)temp);
    const float tstep = frt_maxd(min_dt, frt_mind((float)(0.1 * (double)temp / (double)frt_absd(dTdt)), g.dt));
    const float dt_temp = frt_mind(tstep, g.dt - t_temp);
    const float Tf = frt_maxd(g.temp_min, temp);
    const double Tm = (double)Tf;
    heat = ((double)gh1tot > 3e-14)
               ? (float)(sg_photht_hi(1.5, Tm, (double)nn) * (double)nH + sg_photht_hei(1.5, Tm, (double)nn) * (double)nHe)
               : 0.f;
    cool = 0.f;
    if (g.recomb_cool) {
      if (g.rec_case_b) {
        cool = (float)((double)cool + frt_rcrB_H2(Tm) * (double)nn * (double)nH);
        cool = (float)((double)cool + frt_rcrB_He2(Tm) * (double)nn * (double)nHe);
      } else {
        cool = (float)((double)cool + frt_rcrA_H2(Tm) * (double)nn * (double)nH);
        cool = (float)((double)cool + frt_rcrA_He2(Tm) * (double)nn * (double)nHe);
      }
    }
    if (g.free_free) {
      const float nff = fion * nn;
      cool = (float)((double)cool + frt_free_free_rate(Tm) * dm_pow((double)nff, 2.));
    }
    if (g.compton) cool = (float)((double)cool - g.compton_coef * (double)nn * (g.t_cmb - Tm));
    double den = 1. + 2. * Hd * (double)dt_temp + 1. / (double)ntot * (double)dne * (double)dt_temp;
    if (ddelta != 0.f) den = den - 2. * (double)ddelta / 3. / (double)rho_rel * (double)dt_temp;
    temp = (float)(((double)temp + k23 * (double)(heat - cool) * (double)dt_temp) / den);
    t_temp += dt_temp;
    if (++nsub >= ThermalRec::MAX) break;
  }
  temp = frt_maxd(temp, g.temp_min);
  return nsub;
}

// update_thermal of one cell in the subgrid mode (gas_funcs.cc:375-411).
__device__ __forceinline__ int sg_thermal(const CellPtrs& cf, const GridConst& g, const SgPtrs& sp,
                                          const SgConst& c, int64_t k) {
  if (cf.f[C_TREION_FLAG][k] == 0.f) {
    cf.f[C_TEMP][k] = c.temp_0;  // neutral until first ionized (:380-382)
    return 0;
  }
  float temp = cf.f[C_TEMP][k], heat = cf.f[C_HEAT][k], cool = cf.f[C_COOL][k];
  const float ddelta = cf.f[C_DDELTA_DT][k];
  const float rho_rel = ddelta != 0.f ? cf.f[C_RHO][k] / g.avg_rho : 1.f;
  const int nsub = sg_thermal_cell(g, cf.f[C_NH][k], cf.f[C_NHE][k], cf.f[C_NTOT][k], cf.f[C_DNE_DT][k],
                                   sp.fion[k], cf.f[C_GAMMA_H1_TOT][k], ddelta, rho_rel, temp, heat, cool);
  cf.f[C_TEMP][k] = temp;
  cf.f[C_HEAT][k] = heat;
  cf.f[C_COOL][k] = cool;
  return nsub;
}

// ---- kernels ---------------------------------------------------------------------

// Cell reset of a pass (ray_tracing.cc:630-697): the I-front counts, the vIF
// outputs and the resolver's start; pass 0 adds the recombination
// coefficients, dne_dt, and the RadHydro nodes of ionized cells or the
// neutral MFP guess of the others.
__global__ void __launch_bounds__(BLK) k_sg_reset(CellPtrs cf, GridConst g, SgPtrs sp, SgConst c,
                                                  const int32_t* __restrict__ cstart,
                                                  const int32_t* __restrict__ cend, const StepDev* __restrict__ sd) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= g.n3) return;
  const int64_t n3 = g.n3;
  for (int q = 0; q < 3; ++q) sp.ngam3[q * n3 + k] = 0.;
  cf.f[C_IF_SPEED][k] = 0.f;
  cf.f[C_INC_FLUX][k] = 0.f;
  sp.delta_fion[k] = 0.f;
  sp.rpos[k] = sp.fion[k] < 1.f ? cstart[k] : cend[k];
  if (c.iter != 0) return;
  cf.f[C_AVG_NH1][k] = cf.f[C_NH1][k];
  cf.f[C_AVG_NHE1][k] = cf.f[C_NHE1][k];
  const float T = cf.f[C_TEMP][k];
  const double Tm = (double)frt_maxd(g.temp_min, T);
  float rH2, rHe2;
  if (g.rec_case_b) {
    rH2 = (float)frt_alphaB_H2(Tm);
    rHe2 = (float)frt_alphaB_He2(Tm);
  } else {
    rH2 = (float)frt_alphaA_H2(Tm);
    rHe2 = (float)frt_alphaA_He2(Tm);
  }
  if (((double)T >= 3e4) && ((double)T <= 1e6)) rHe2 += (float)frt_Dalpha_He2(Tm);
  cf.f[C_RECOMB_H2][k] = rH2;
  cf.f[C_RECOMB_HE2][k] = rHe2;
  cf.f[C_DNE_DT][k] = -(cf.f[C_NH][k] + cf.f[C_NHE][k]);
  if (cf.f[C_TREION_FLAG][k] != 0.f) {
    const float tre = cf.f[C_TREION][k];
    sg_gamma_interp_fh1(sp, c, g.zz, n3, k, tre + c.tinit, (float)((double)(cf.f[C_RHO][k] / c.avg_rho) - 1.),
                        sd->t - tre);
  } else {
    cf.f[C_MFP][k] = (float)(1. / (double)cf.f[C_NH1][k] / g.sigH[0]);
  }
}

// Walk state of every moved ray at the start of a pass.
__global__ void __launch_bounds__(BLK) k_sg_walk_init(SgWalk w, const int32_t* __restrict__ moved,
                                                      const StepDev* __restrict__ sd, int F, int64_t cap) {
  const int64_t nmoved = sd->nmoved;
  for (int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x; t < nmoved; t += (int64_t)gridDim.x * BLK) {
    const int64_t i = moved[t];
    for (int q = 0; q < 4 * F; ++q) w.tau[(int64_t)q * cap + i] = 0.;
    w.p[i] = 0;
    w.np[i] = 0;
    w.wait[i] = -1;
    w.flags[i] = 0;
  }
}

__global__ void __launch_bounds__(BLK) k_sg_slot_clear(uint8_t* __restrict__ slot, const StepDev* __restrict__ sd) {
  const int64_t n = sd->nseg;
  for (int64_t s = (int64_t)blockIdx.x * BLK + threadIdx.x; s < n; s += (int64_t)gridDim.x * BLK) slot[s] = 0;
}

__global__ void __launch_bounds__(BLK) k_sg_zero(int64_t* __restrict__ v) { *v = 0; }

// The ray loop of one pass (ray_tracing.cc:701-938), resumed where each ray
// stopped: at every segment the photons in are recorded (the cell pass
// turns them into Gamma), the opacity of the ionized part is added, and in a
// cell with fion < 1 the photons that reach its neutral part are posted to
// the segment's slot; the ray then waits for k_sg_resolve's answer.
__global__ void __launch_bounds__(BLK) k_sg_walk(RaySoA rb, const int32_t* __restrict__ moved,
                                                 const StepDev* __restrict__ sd, GridConst g, SgWalk w, SgPtrs sp,
                                                 int64_t scap, const float* __restrict__ mfp,
                                                 const int32_t* __restrict__ path_cell,
                                                 const float* __restrict__ path_dist,
                                                 const int32_t* __restrict__ seg_cnt,
                                                 const int32_t* __restrict__ path_slot, ContribPtrs cp,
                                                 uint8_t* __restrict__ alive, DevCounters* cnt,
                                                 float* __restrict__ psum, int last) {
  const int64_t nmoved = sd->nmoved;
  const int F = g.nfreq;
  const int64_t cap = g.cap;
  const double chi = (double)g.nhe_to_nh;
  unsigned long long waiting_rays = 0;
  for (int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x; t < nmoved; t += (int64_t)gridDim.x * BLK) {
    const int64_t i = moved[t];
    uint8_t fl = w.flags[i];
    if (fl & 2) continue;
    const int c = seg_cnt[i];
    double th[FRT_MAX_FREQ], the[FRT_MAX_FREQ], dth[FRT_MAX_FREQ], dthe[FRT_MAX_FREQ], nph[FRT_MAX_FREQ];
    {
      float n0[8];
      phot_load(rb, i, F, cap, n0);
      FRT_FOR_BINS(nu) {
        nph[nu] = (double)n0[nu] * FRT_PHOT_NORM;
        th[nu] = w.tau[(int64_t)(0 * F + nu) * cap + i];
        the[nu] = w.tau[(int64_t)(1 * F + nu) * cap + i];
        dth[nu] = w.tau[(int64_t)(2 * F + nu) * cap + i];
        dthe[nu] = w.tau[(int64_t)(3 * F + nu) * cap + i];
      }
    }
    int p = w.p[i], np = w.np[i], ws = w.wait[i];
    if (ws >= 0) {
      const uint8_t f = w.slot[ws];
      if (f == 1) {  // the cell has not reached this ray yet
        ++waiting_rays;
        continue;
      }
      if (f == 5) {  // all photons absorbed (:896-901)
        p = 99;
        FRT_FOR_BINS(nu) th[nu] = 999.;
      } else if (f == 2) {  // the rest spread over the spectrum as opacity (:903-909)
        FRT_FOR_BINS(nu) {
          const double out = w.ifr[(int64_t)nu * scap + ws];
          th[nu] += -dm_log(out / nph[nu] / dm_exp(-(th[nu] + the[nu])));
        }
      }
      ws = -1;
    }
    bool waiting = false;
    while (p < c) {
      const int64_t e = (int64_t)p * cap + i;
      const int slot = path_slot[e];
      const int cell = path_cell[e];
      const float dist = path_dist[e];
      const float fi = sp.fion[cell], mf = mfp[cell];
      double nin[FRT_MAX_FREQ];
      double tot = 0.;
      FRT_FOR_BINS(nu) {
        nin[nu] = nph[nu] * dm_exp(-th[nu] - the[nu]);
        tot += nin[nu] / FRT_PHOT_NORM;
      }
      const bool dump = tot < g.nphot_min;
      store_photon_record((double*)cp.rec + (int64_t)slot * cp.rec_len, nin, F, dist, dump, cp.rec_len);
      np = p + 1;
      if (dump) {
        p = 99;
        fl |= 1;
      } else {
        p += 1;
      }
      if (fi > 0.f) {  // MFP opacity of the ionized part (:850-857)
        FRT_FOR_BINS(nu) {
          dth[nu] = (double)(fi * dist / mf);
          dthe[nu] = chi * dth[nu];
        }
      }
      FRT_FOR_BINS(nu) {
        th[nu] += dth[nu];
        the[nu] += dthe[nu];
      }
      if (fi < 1.f) {  // photons into the neutral part (:867-919)
        FRT_FOR_BINS(nu) w.ifr[(int64_t)nu * scap + slot] = nin[nu] * dm_exp(-(dth[nu] + dthe[nu]));
        w.slot[slot] = 1;
        ws = slot;
        waiting = true;
        break;
      }
    }
    if (waiting) {
      FRT_FOR_BINS(nu) {
        w.tau[(int64_t)(0 * F + nu) * cap + i] = th[nu];
        w.tau[(int64_t)(1 * F + nu) * cap + i] = the[nu];
        w.tau[(int64_t)(2 * F + nu) * cap + i] = dth[nu];
        w.tau[(int64_t)(3 * F + nu) * cap + i] = dthe[nu];
      }
      w.p[i] = p;
      w.np[i] = np;
      w.wait[i] = ws;
      w.flags[i] = fl;
      ++waiting_rays;
      continue;
    }
    // the path is done: segments behind a dump or an absorption deposit nothing
    for (int q = np; q < c; ++q) {
      const int slot = path_slot[(int64_t)q * cap + i];
      store_behind_dump((double*)cp.rec + (int64_t)slot * cp.rec_len, F, cp.rec_len);
      w.slot[slot] = 4;
    }
    w.wait[i] = -1;
    w.flags[i] = fl | 2;
    if (last) {  // photon update and removal (:923-935)
      float n1[8];
      float sum = 0.f;
      FRT_FOR_BINS(nu) {
        const float v = (float)(nph[nu] * dm_exp(-th[nu] - the[nu]) / FRT_PHOT_NORM);
        n1[nu] = v;
        sum += v;
      }
      phot_store(rb, i, F, cap, n1);
      psum[i] = sum;
      if (fl & 1) atomicAdd(&cnt->dumped, 1ull);
      if ((double)sum > g.nphot_min) {
        ray_set_cell(rb, i, path_cell[(int64_t)(c - 1) * cap + i]);
      } else {
        alive[i] = 0;
        atomicAdd(&cnt->terminated, 1ull);
      }
    }
  }
  if (waiting_rays) atomicAdd((unsigned long long*)w.pending, waiting_rays);
}

// The neutral part of a cell with fion < 1 absorbs the photons of the
// segments posted to it, in slot order, up to the atoms it holds
// (ray_tracing.cc:867-919); stops at the first segment whose ray has not
// reached it (its share depends on every earlier one).
__global__ void __launch_bounds__(BLK) k_sg_resolve(CellPtrs cf, GridConst g, SgPtrs sp, SgWalk w, int64_t scap,
                                                    const int32_t* __restrict__ cend, ContribPtrs cp) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= g.n3) return;
  int s = sp.rpos[k];
  const int e = cend[k];
  if (s >= e) return;
  const int64_t n3 = g.n3;
  const int F = g.nfreq;
  const double chi = (double)g.nhe_to_nh;
  double ng0 = sp.ngam3[k], ng1 = sp.ngam3[n3 + k], ng2 = sp.ngam3[2 * n3 + k];
  const double free_ngam = g.vol * (double)cf.f[C_NH][k] * (1. + chi) * (1. - (double)sp.fion[k]);
  const double* R = (const double*)cp.rec;
  for (; s < e; ++s) {
    const uint8_t f = w.slot[s];
    if (f == 0) break;
    if (f != 1) continue;
    const double d = fabs(R[(int64_t)s * cp.rec_len + F]);
    const double sum_ngam = ng0;
    double a[FRT_MAX_FREQ];
    FRT_FOR_BINS(nu) {
      a[nu] = w.ifr[(int64_t)nu * scap + s];
      ng2 += a[nu];
    }
    if (sum_ngam < free_ngam) {
      double left = 0.;
      FRT_FOR_BINS(nu) {
        const double room = free_ngam - sum_ngam;
        const double ab = (a[nu] >= room) ? room : a[nu];  // mindd
        ng0 += ab;
        ng1 += ab * (double)sgd::sigma_eff_rh * d;
        const double out = a[nu] - ab;
        left += out;
        w.ifr[(int64_t)nu * scap + s] = out;
      }
      w.slot[s] = (left == 0.) ? 5 : 2;
    } else {
      w.slot[s] = 3;
    }
  }
  sp.rpos[k] = s;
  sp.ngam3[k] = ng0;
  sp.ngam3[n3 + k] = ng1;
  sp.ngam3[2 * n3 + k] = ng2;
}

// The cell loop of a pass, subgrid branch (ray_tracing.cc:954-1064), from the
// photon records in slot order; the last pass adds update_fion, the
// ionization state from the RadHydro nodes, the rates (:1139-1210) and, with
// THERMAL, update_thermal's subgrid branch.
template <bool THERMAL>
__global__ void __launch_bounds__(BLK) k_sg_chem(CellPtrs cf, GridConst g, SgPtrs sp, SgConst c, ContribPtrs cp,
                                                 const int32_t* __restrict__ cstart,
                                                 const int32_t* __restrict__ cend, DevCounters* cnt,
                                                 const StepDev* __restrict__ sd) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int nsub = 0;
  if (k < g.n3) {
    const int64_t n3 = g.n3;
    const int F = g.nfreq;
    const double frac_h1 = 1. - (FRT_Y_HE / 4.) / (1. - FRT_Y_HE + FRT_Y_HE / 4.);  // :766-767
    const double frac_he1 = (FRT_Y_HE / 4.) / (1. - FRT_Y_HE + FRT_Y_HE / 4.);
    const float fi = sp.fion[k], mf = cf.f[C_MFP][k];
    const float nH = cf.f[C_NH][k], nHe = cf.f[C_NHE][k], clump = cf.f[C_CLUMP][k];
    // Gamma and photon sums over the cell's segments in slot order (:786-834)
    double a = 0., b = 0.;
    double x[FRT_MAX_FREQ], y[FRT_MAX_FREQ], z[FRT_MAX_FREQ];
    FRT_FOR_BINS(nu) x[nu] = y[nu] = z[nu] = 0.;
    const double* R = (const double*)cp.rec;
    const int s0 = cstart[k], s1 = cend[k];
    for (int s = s0; s < s1; ++s) {
      const double* r = R + (int64_t)s * cp.rec_len;
      const double d = r[F];
      if (isinf(d)) continue;  // behind a dump
      const float dist = (float)fabs(d);
      double gs = 0., gse = 0.;
      FRT_FOR_BINS(nu) {
        const double nin = r[nu];
        z[nu] += nin / FRT_PHOT_NORM;
        if (fi > 0.f) {
          float dse;
          if (c.iter == 0)
            dse = (float)((double)(fi * dist) / frac_h1);
          else
            dse = (float)((double)mf * (1. - dm_exp((double)(-fi * dist / mf) / frac_h1)));
          const double gh = frac_h1 * nin * (double)sgd::sigma_eff_rh * (double)dse / (double)g.dt / g.vol / (double)fi;
          const double ghe = frac_he1 * nin * (double)sgd::sigma_eff_rh * (double)dse / (double)g.dt / g.vol / (double)fi;
          x[nu] += gh;
          y[nu] += ghe;
          gs += gh;
          gse += ghe;
        }
      }
      a += gs;
      b += gse;
    }
    const double nelec = (double)(nH - cf.f[C_AVG_NH1][k] + nHe - cf.f[C_AVG_NHE1][k]);
    float ag = (float)a;
    float treion_flag = cf.f[C_TREION_FLAG][k];
    bool first = false;
    const float t = sd->t;
    if (c.last && (treion_flag == 0.f) && ((sp.ngam3[n3 + k] > 0.) || (ag > 0.))) {  // :966-971
      const float tre = t - g.dt;
      cf.f[C_TREION][k] = tre;
      cf.f[C_TREION_FLAG][k] = treion_flag = 1.f;
      sg_gamma_interp_fh1(sp, c, g.zz, n3, k, tre + c.tinit, (float)((double)(cf.f[C_RHO][k] / c.avg_rho) - 1.),
                          t - tre);
      first = true;
    }
    float fion = fi;
    float temp = cf.f[C_TEMP][k];
    if ((treion_flag != 0.f) && (fion > 0.f)) sg_mfp_eq(g, ag, clump, temp, nH, cf.f[C_MFP][k], cf.f[C_MFP912][k]);
    if ((treion_flag != 0.f) && c.last && (fion < 1.f)) {  // I-front correction (:1004-1052)
      const double gamma_fion = sp.ngam3[n3 + k] / (double)g.dt / g.vol;
      // update_fion (subgrid.cc:689-742)
      const float vIF_min = (float)(5e1 * 1e5);
      const double Fx = sp.ngam3[2 * n3 + k] / dm_pow((double)g.dx, 2.) / (double)g.dt;
      const float vIF = (float)(Fx / (double)nH / (1. + (double)g.nhe_to_nh));
      const float fion_old = fion;
      const float fmax_old = sp.fion_max[k];
      fion = frt_mind(fion + vIF * g.dt / g.dx, 1.f);
      float fmax = fmax_old;
      if (fion > fmax) fmax = fion;
      if ((vIF > vIF_min) || (fion_old == 0.)) {  // set_Treion (subgrid.cc:652-666)
        const float v = frt_maxd(vIF, vIF_min);
        const float Cn[5] = {9.5432f, -1.6441e-2f, -9.8010e-3f, 4.1664e-3f, -2.2710e-4f};
        const double L = dm_log((double)v / 1e5);
        float lnT = 0.f;
        for (int q = 0; q < 5; ++q) lnT = (float)((double)lnT + (double)Cn[q] * dm_pow(L, (double)q));
        if (fion > 0.f) temp = (fion_old * temp + (fion - fion_old) * dm_expf(lnT)) / fion;
      }
      if ((fion_old < 0.5f) && (fion >= 0.5f)) sp.treion_out[k] = t - g.dt;
      const float tdiff = (fion >= 1.f) ? (float)(0.5 * (double)g.dt) : 0.f;
      cf.f[C_IF_SPEED][k] = vIF;
      cf.f[C_INC_FLUX][k] = (float)Fx;
      sp.delta_fion[k] = fion - fion_old;
      sp.tcross[k] += g.dt - tdiff;
      const float tc_frac = (float)(1. / ((float)FRT_SG_NTCROSS));
      for (int tt = 0; tt < FRT_SG_NTCROSS; ++tt) {
        float& tcn = sp.tcross_new[(int64_t)tt * n3 + k];
        float& chk = sp.tcross_check[(int64_t)tt * n3 + k];
        const double edge = (double)tc_frac * (tt + 1.);
        if (((double)fmax < edge) && (chk == 0.f)) {
          tcn += g.dt;
        } else if (((double)fmax >= edge) && ((double)fmax_old < edge) && (chk == 0.f)) {
          tcn = (float)((double)tcn + (double)g.dt * (edge - (double)fmax_old) / (double)(fmax - fmax_old));
          chk = 1.f;
        }
      }
      sp.fion_max[k] = fmax;
      sp.tcross_flag[k] = 1.f;
      // recombinations in a partly ionized cell with no ionizing flux (:1020-1027)
      if ((fion < 1.f) && (ag == 0.f) && !first)
        fion = frt_maxd(0.f, (float)((double)fion - (double)clump * frt_alphaB_H2((double)frt_maxd(g.temp_min, temp)) *
                                                        (1. + (double)g.nhe_to_nh) * (double)nH * (double)fion *
                                                        (double)g.dt));
      if (first) {  // :1034-1051
        cf.f[C_GAMMA_PREV][k] = 0.f;
        ag = (float)gamma_fion;
        sg_mfp_eq(g, ag, clump, temp, nH, cf.f[C_MFP][k], cf.f[C_MFP912][k]);
      }
    }
    if ((fion == 1.f) && (ag == 0.f) && c.last)  // fully ionized and empty (:1055-1059)
      fion = frt_maxd(0.f, (float)((double)fion - (double)clump * frt_alphaB_H2((double)frt_maxd(g.temp_min, temp)) *
                                                      (1. + (double)g.nhe_to_nh) * (double)nH * (double)fion *
                                                      (double)g.dt));
    sp.fion[k] = fion;
    cf.f[C_TEMP][k] = temp;
    if (c.last) {
      float nH2 = cf.f[C_NH2][k], nHe2 = cf.f[C_NHE2][k];
      if (treion_flag != 0.f) {  // :1067-1078
        const float nH1n = sg_interp_xh1(sp, n3, k, ag, c.fh1_0) * nH;
        const float nHe1n = nH1n * nHe / nH;
        cf.f[C_NH1][k] = nH1n;
        cf.f[C_NHE1][k] = nHe1n;
        cf.f[C_AVG_NHE1][k] = nHe1n;
        nH2 = nH * fion;
        nHe2 = nHe * fion;
        cf.f[C_NH2][k] = nH2;
        cf.f[C_NHE2][k] = nHe2;
        cf.f[C_GAMMA_PREV][k] = ag;
      }
      FRT_FOR_BINS(nu) {
        cf.gamma_h1_nu[nu * n3 + k] = (float)x[nu];
        cf.gamma_he1_nu[nu * n3 + k] = (float)y[nu];
        cf.ngam_nu[nu * n3 + k] = (float)z[nu];
      }
      cf.f[C_GAMMA_H1_TOT][k] = ag;
      cf.f[C_GAMMA_HE1_TOT][k] = (float)b;
      cf.f[C_RECRATE_H2][k] = (float)((double)(clump * cf.f[C_RECOMB_H2][k]) * nelec * (double)nH2);
      cf.f[C_RECRATE_HE2][k] = (float)((double)(clump * cf.f[C_RECOMB_HE2][k]) * nelec * (double)nHe2);
      const float nHe3 = 0.f;
      cf.f[C_NE][k] = (float)((double)(nH2 + nHe2) + 2. * (double)nHe3);
      float dne = cf.f[C_DNE_DT][k];
      dne += nH + nHe;
      cf.f[C_NTOT][k] = (float)(2. * (double)(nH + nHe));
      dne /= g.dt;
      cf.f[C_DNE_DT][k] = dne;
      cf.f[C_NH1_PREV][k] = cf.f[C_NH1][k];
      float fH2 = nH2 / nH, fHe2 = nHe2 / nHe, fHe3 = nHe3 / nHe;
      if (fion > 0.f) {
        fH2 *= fion;
        fHe2 *= fion;
        fHe3 *= fion;
      }
      cf.f[C_FH2][k] = fH2;
      cf.f[C_FHE2][k] = fHe2;
      cf.f[C_FHE3][k] = fHe3;
      cf.f[C_FH1][k] = (float)(1. - (double)fH2);
      cf.f[C_FHE1][k] = (float)(1. - (double)fHe2 - (double)fHe3);
      if (THERMAL) nsub = sg_thermal(cf, g, sp, c, k);
    }
  }
  if (THERMAL) block_add(&cnt->substeps, (unsigned long long)nsub);
}

__global__ void __launch_bounds__(BLK) k_sg_thermal(CellPtrs cf, GridConst g, SgPtrs sp, SgConst c,
                                                    DevCounters* cnt) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int nsub = 0;
  if (k < g.n3) nsub = sg_thermal(cf, g, sp, c, k);
  block_add(&cnt->substeps, (unsigned long long)nsub);
}

// ---- host side ------------------------------------------------------------------

// Per-cell state, ray walk state and the I-front counter of the subgrid mode
// (capacity-sized; the slot buffers follow the segment capacity).
int alloc_subgrid(frt_ctx* ctx) {
  const int64_t n3 = ctx->g.n3, cap = ctx->g.cap;
  const int F = ctx->g.nfreq;
  const size_t nfl = (size_t)(6 + 2 * FRT_SG_NTCROSS + 5) * n3;
  const size_t bytes = 4 * nfl + 8 * 3 * (size_t)n3 + 4 * (size_t)n3;
  hipError_t e = hipMalloc(&ctx->sg_block, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(ctx->sg_block, 0, bytes, ctx->stream);
  if (e != hipSuccess) return set_error(ctx, FRT_E_NOMEM, "subgrid state: " + std::string(hipGetErrorString(e)));
  float* f = (float*)ctx->sg_block;
  SgPtrs& sp = ctx->sg;
  sp.fion = f;
  sp.fion_max = f + n3;
  sp.tcross = f + 2 * n3;
  sp.tcross_flag = f + 3 * n3;
  sp.treion_out = f + 4 * n3;
  sp.delta_fion = f + 5 * n3;
  sp.tcross_new = f + 6 * n3;
  sp.tcross_check = sp.tcross_new + FRT_SG_NTCROSS * n3;
  sp.fh1_nodes = sp.tcross_check + FRT_SG_NTCROSS * n3;
  sp.ngam3 = (double*)(f + nfl);
  sp.rpos = (int32_t*)(sp.ngam3 + 3 * n3);
  SgWalk& w = ctx->sgw;
  e = hipMalloc((void**)&w.tau, 8 * (size_t)4 * F * cap);
  if (e == hipSuccess) e = hipMalloc((void**)&w.p, 4 * (size_t)cap);
  if (e == hipSuccess) e = hipMalloc((void**)&w.np, 4 * (size_t)cap);
  if (e == hipSuccess) e = hipMalloc((void**)&w.wait, 4 * (size_t)cap);
  if (e == hipSuccess) e = hipMalloc((void**)&w.flags, (size_t)cap);
  if (e == hipSuccess) e = hipMalloc((void**)&w.pending, 8);
  if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->sg_pending_h, 8);
  if (e != hipSuccess) return set_error(ctx, FRT_E_NOMEM, "subgrid walk: " + std::string(hipGetErrorString(e)));
  return FRT_OK;
}

void free_subgrid(frt_ctx* ctx) {
  hipFree(ctx->sg_block);
  hipFree(ctx->sg_tables);
  hipFree(ctx->sgw.tau);
  hipFree(ctx->sgw.p);
  hipFree(ctx->sgw.np);
  hipFree(ctx->sgw.wait);
  hipFree(ctx->sgw.flags);
  hipFree(ctx->sgw.pending);
  hipFree(ctx->sgw.ifr);
  hipFree(ctx->sgw.slot);
  if (ctx->sg_pending_h) hipHostFree(ctx->sg_pending_h);
  ctx->sg_block = nullptr;
  ctx->sg_tables = nullptr;
  ctx->sgw = SgWalk{};
  ctx->sg_pending_h = nullptr;
}

// Slot buffers for the current segment capacity.
static int ensure_sg_slots(frt_ctx* ctx) {
  if (ctx->sg_slot_cap >= ctx->seg_cap && ctx->sgw.slot) return FRT_OK;
  hipStreamSynchronize(ctx->stream);
  hipFree(ctx->sgw.ifr);
  hipFree(ctx->sgw.slot);
  ctx->sgw.ifr = nullptr;
  ctx->sgw.slot = nullptr;
  const int64_t sc = ctx->seg_cap;
  hipError_t e = hipMalloc((void**)&ctx->sgw.ifr, 8 * (size_t)ctx->g.nfreq * sc);
  if (e == hipSuccess) e = hipMalloc((void**)&ctx->sgw.slot, (size_t)sc);
  if (e != hipSuccess) {
    ctx->sg_slot_cap = 0;
    return set_error(ctx, FRT_E_NOMEM, "subgrid slots: " + std::string(hipGetErrorString(e)));
  }
  ctx->sg_slot_cap = sc;
  return FRT_OK;
}

// find_avg_unu in the subgrid mode: per pass the cell reset, the walk and
// resolve rounds until no ray waits (the host reads the waiting count every
// few rounds), then the cell pass; the step's thermal update is fused into
// the last one.
int launch_find_avg_unu_subgrid(frt_ctx* ctx, bool fuse_thermal) {
  if (!ctx->subgrid_set) return set_error(ctx, FRT_E_STATE, "subgrid: frt_set_subgrid has not been called");
  if (int rc0 = materialize_rays(ctx)) return rc0;
  if (int rc0 = ensure_sg_slots(ctx)) return rc0;
  const int I = ctx->cfg.itercount;
  GridConst g = ctx->g;
  RaySoA rb = cur_buf(ctx);
  StepDev* sd = ctx->sd;
  const unsigned gc = grid_for(g.n3);
  const unsigned rgrid = std::min(ray_grid(ctx, bound_rays(ctx)), MAX_GRID);
  const unsigned sgrid = std::min(ray_grid(ctx, bound_segs(ctx)), 4 * MAX_GRID);
  ctx->deposit_launches_last = 0;
  ctx->sg_rounds = 0;
  constexpr int CHECK = 4;            // rounds between reads of the waiting count
  constexpr int64_t MAX_ROUNDS = 1 << 20;
  for (int it = 0; it < I; ++it) {
    SgConst c = ctx->sgc;
    c.iter = it;
    c.last = it == I - 1;
    g.pass = it;
    FRT_LAUNCH(ctx, "sg_reset", gc, k_sg_reset, ctx->c, g, ctx->sg, c, ctx->cell_start, ctx->cell_end, sd);
    FRT_LAUNCH(ctx, "sg_walk", rgrid, k_sg_walk_init, ctx->sgw, ctx->moved_idx, sd, g.nfreq, g.cap);
    FRT_LAUNCH(ctx, "sg_walk", sgrid, k_sg_slot_clear, ctx->sgw.slot, sd);
    for (int64_t r = 0;; ++r) {
      if (r >= MAX_ROUNDS) return set_error(ctx, FRT_E_STATE, "subgrid: I-front rounds did not finish");
      FRT_LAUNCH(ctx, "sg_walk", 1, k_sg_zero, ctx->sgw.pending);
      FRT_LAUNCH(ctx, "sg_walk", rgrid, k_sg_walk, rb, ctx->moved_idx, sd, g, ctx->sgw, ctx->sg, ctx->sg_slot_cap,
                 ctx->c.f[C_MFP], ctx->path_cell, ctx->path_dist, ctx->seg_cnt, ctx->path_slot, ctx->contrib,
                 ctx->alive, ctx->dcnt, ctx->psum, c.last ? 1 : 0);
      FRT_LAUNCH(ctx, "sg_resolve", gc, k_sg_resolve, ctx->c, g, ctx->sg, ctx->sgw, ctx->sg_slot_cap, ctx->cell_end,
                 ctx->contrib);
      ctx->sg_rounds += 1;
      if (r % CHECK == CHECK - 1) {
        hipError_t e = hipMemcpyAsync(ctx->sg_pending_h, ctx->sgw.pending, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_check(ctx, e, "subgrid rounds");
        if (*ctx->sg_pending_h == 0) break;
      }
    }
    ctx->deposit_launches_last += 1;
    if (c.last && fuse_thermal)
      FRT_LAUNCH(ctx, "sg_chem", gc, k_sg_chem<true>, ctx->c, g, ctx->sg, c, ctx->contrib, ctx->cell_start,
                 ctx->cell_end, ctx->dcnt, sd);
    else
      FRT_LAUNCH(ctx, "sg_chem", gc, k_sg_chem<false>, ctx->c, g, ctx->sg, c, ctx->contrib, ctx->cell_start,
                 ctx->cell_end, ctx->dcnt, sd);
  }
  if (I > 0 && ctx->psum_state != FRT_PSUM_NONE) ctx->psum_state = FRT_PSUM_VALID;
  return FRT_OK;
}

int launch_thermal_subgrid(frt_ctx* ctx) {
  SgConst c = ctx->sgc;
  FRT_LAUNCH(ctx, "thermal", grid_for(ctx->g.n3), k_sg_thermal, ctx->c, ctx->g, ctx->sg, c, ctx->dcnt);
  return FRT_OK;
}

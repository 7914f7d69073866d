// Adjoint of one find_avg_unu (ray_tracing.cc:623-1216, C2Ray branch), with
// the ray paths, the dump decisions and the recombination coefficients held
// at their forward values.  The reverse sweep is FP64 in both forms; for the
// mixed form it differentiates the same expressions the FP32 forward
// evaluates, at the forward's recorded values.
//
// The forward call records a tape (frt_kernels.hip, launch_find_avg_unu): for
// every C2Ray pass k the opacities n_HI, n_HeI its deposit used and the cell
// sums Gamma_HI, Gamma_HeI its chemistry used, the starting f_H1, f_He1, and
// the rays' starting photons.  The reverse sweep runs the passes backwards:
//
//   k_adj_chem     per cell, the C2Ray update (ray_tracing.cc:1084-1127):
//                  from the adjoint of its outputs (avg n_HI / n_HeI of the
//                  pass, and n_HI of the last pass) to the adjoint of the
//                  Gamma sums, of the pass's opacities through n_e, of nH, nHe
//                  and of the clumping factor;
//   k_adj_deposit  per ray, the attenuation march and deposition
//                  (ray_tracing.cc:701-863): the forward transmission is
//                  replayed, then the segments are walked backwards; each
//                  segment's adjoint of its cell's opacities goes to the
//                  segment's cell-sorted slot, the photons' adjoint to the ray;
//   k_adj_cellsum  per cell, the slots added in slot order.
//
// All sums are in a fixed order, so the gradient is deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "frt_internal.h"
#include "frt_detmath.h"

namespace frt {

static constexpr int ABLK = 256;

template <bool LAST>
__global__ void __launch_bounds__(ABLK) k_adj_chem(CellPtrs cf, GridConst g, const double* __restrict__ tk,
                                                   const double* __restrict__ x0tape,
                                                   const double* __restrict__ a_out_h,
                                                   const double* __restrict__ a_out_he,
                                                   const double* __restrict__ a_nh1_out,
                                                   double* __restrict__ aGH, double* __restrict__ aGHe,
                                                   double* __restrict__ a_ah, double* __restrict__ a_ahe,
                                                   double* __restrict__ a_nH, double* __restrict__ a_nHe,
                                                   double* __restrict__ a_C) {
  const int64_t k = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  const int64_t n3 = g.n3;
  if (k >= n3) return;
  const float nHf = cf.f[C_NH][k], nHef = cf.f[C_NHE][k], Cf = cf.f[C_CLUMP][k];
  const float rH2 = cf.f[C_RECOMB_H2][k], rHe2 = cf.f[C_RECOMB_HE2][k];
  const double nH = nHf, nHe = nHef;
  const double ah = tk[k], ahe = tk[n3 + k], G = tk[2 * n3 + k], Ghe = tk[3 * n3 + k];
  const double x0 = 1. - x0tape[k], x0he = 1. - x0tape[n3 + k];
  const double dt = (double)g.dt;
  // n_e = nH - n_HI + nHe - n_HeI (ray_tracing.cc:955), C alpha as the forward's float product
  const double u = (double)(nHf - (float)ah + nHef - (float)ahe);
  const double cH = (double)(Cf * rH2), cHe = (double)(Cf * rHe2);
  double a_u = 0., a_G = 0., a_Ghe = 0., a_cH = 0., a_cHe = 0., a_nHd = 0., a_nHed = 0.;
  {
    // HI: avg = nH (1 - F), F = xeq + (x0 - xeq) h, h = (1 - e)/(dt pr), e = exp(-dt pr)
    const double pr = G + cH * u;
    const double xeq = G / pr;
    const double e = std::exp(-dt * pr);
    const double h = (1. - e) / (dt * pr);
    const double dh = (dt * pr * e - (1. - e)) / (dt * pr * pr);
    const double Fv = xeq + (x0 - xeq) * h;
    const double dF_dG = (1. - h) / pr;                               // at fixed pr
    const double dF_dpr = (1. - h) * (-G / (pr * pr)) + (x0 - xeq) * dh;  // at fixed G
    const double ao = a_out_h ? a_out_h[k] : 0.;
    double aF = -nH * ao;
    a_nHd += (1. - Fv) * ao;
    double a_pr = aF * dF_dpr;
    a_G += aF * dF_dG;
    if (LAST) {
      // n_HI at the end of the step: nH (1 - FL), FL = xeq + (x0 - xeq) e (:1118)
      const double FL = xeq + (x0 - xeq) * e;
      const double aL = a_nh1_out[k];
      const double aFL = -nH * aL;
      a_nHd += (1. - FL) * aL;
      a_G += aFL * (1. - e) / pr;
      a_pr += aFL * ((1. - e) * (-G / (pr * pr)) + (x0 - xeq) * (-dt * e));
    }
    a_G += a_pr;  // pr = G + cH u
    a_u += a_pr * cH;
    a_cH += a_pr * u;
  }
  {
    // HeI, the same update (:1106-1113)
    const double pr = Ghe + cHe * u;
    const double xeq = Ghe / pr;
    const double e = std::exp(-dt * pr);
    const double h = (1. - e) / (dt * pr);
    const double dh = (dt * pr * e - (1. - e)) / (dt * pr * pr);
    const double Fv = xeq + (x0he - xeq) * h;
    const double ao = a_out_he ? a_out_he[k] : 0.;
    const double aF = -nHe * ao;
    a_nHed += (1. - Fv) * ao;
    const double a_pr = aF * ((1. - h) * (-Ghe / (pr * pr)) + (x0he - xeq) * dh);
    a_Ghe += aF * (1. - h) / pr + a_pr;
    a_u += a_pr * cHe;
    a_cHe += a_pr * u;
  }
  aGH[k] = a_G;
  aGHe[k] = a_Ghe;
  a_ah[k] = -a_u;  // the deposit's part is added by k_adj_cellsum
  a_ahe[k] = -a_u;
  a_nH[k] += a_nHd + a_u;
  a_nHe[k] += a_nHed + a_u;
  a_C[k] += a_cH * (double)rH2 + a_cHe * (double)rHe2;
}

// One ray per thread: the segments the forward pass processed and whether
// it dumped the ray there come from the tape (the forward's own decision, in
// either precision form); the optical depth in front of each segment is
// summed forward, then the processed segments are walked backwards.  R[nu]
// carries the adjoint of the optical depth accumulated in front of the
// current segment.
__global__ void __launch_bounds__(ABLK) k_adj_deposit(GridConst g, int64_t nrays, const int32_t* __restrict__ seg_cnt,
                                                      const int32_t* __restrict__ path_cell,
                                                      const float* __restrict__ path_dist,
                                                      const int32_t* __restrict__ path_slot,
                                                      const double* __restrict__ tk, const float* __restrict__ nph0,
                                                      const uint8_t* __restrict__ tape_np,
                                                      const double* __restrict__ aGH, const double* __restrict__ aGHe,
                                                      double2* __restrict__ slot_out, double* __restrict__ a_nph) {
  const int64_t i = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  if (i >= nrays) return;
  const int c = seg_cnt[i];
  if (c <= 0) return;
  const int F = g.nfreq;
  const int64_t n3 = g.n3, cap = g.cap;
  const double* tah = tk;
  const double* tahe = tk + n3;
  double nph[FRT_MAX_FREQ], tauh[FRT_MAX_FREQ], tauhe[FRT_MAX_FREQ], R[FRT_MAX_FREQ], anph[FRT_MAX_FREQ];
  for (int nu = 0; nu < F; ++nu) {
    nph[nu] = (double)nph0[i * F + nu] * FRT_PHOT_NORM;
    tauh[nu] = tauhe[nu] = R[nu] = anph[nu] = 0.;
  }
  const int rec = tape_np[i];
  const int np = rec & 0x7f;
  const bool dumped = (rec & 0x80) != 0;
  for (int p = 0; p < np; ++p) {
    const int cell = path_cell[(int64_t)p * cap + i];
    const double dist = (double)path_dist[(int64_t)p * cap + i];
    const double ah = tah[cell], ahe = tahe[cell];
    for (int nu = 0; nu < F; ++nu) {
      tauh[nu] += (ah * g.sigH[nu]) * dist;
      tauhe[nu] += (ahe * g.sigHe[nu]) * dist;
    }
  }
  for (int p = np; p < c; ++p) slot_out[path_slot[(int64_t)p * cap + i]] = make_double2(0., 0.);
  const double dtv = (double)g.dt * g.vol;
  for (int p = np - 1; p >= 0; --p) {
    const int cell = path_cell[(int64_t)p * cap + i];
    const int slot = path_slot[(int64_t)p * cap + i];
    const double dist = (double)path_dist[(int64_t)p * cap + i];
    const double ah = tah[cell], ahe = tahe[cell];
    const bool dump = dumped && (p == np - 1);
    const double wH = aGH[cell], wHe = aGHe[cell];
    double a_ah = 0., a_ahe = 0.;
    for (int nu = 0; nu < F; ++nu) {
      const double dth = (ah * g.sigH[nu]) * dist;
      const double dthe = (ahe * g.sigHe[nu]) * dist;
      tauh[nu] -= dth;  // optical depth in front of this segment
      tauhe[nu] -= dthe;
      const double ex = std::exp(-tauh[nu] - tauhe[nu]);
      const double nin = nph[nu] * ex;
      // HI/HeI split of the absorbed photons (ray_tracing.cc:770-771):
      // fh = a/(a+b), a = (1-eh) ehe, b = (1-ehe) eh.  Its derivatives in the
      // form fh fhe/(1-eh), -fh fhe/(1-ehe), which stays finite for optically
      // thick segments where (a+b)^2 underflows; both saturated: the logistic
      // limit the forward uses (frac_limit).
      const double eh = std::exp(-dth), ehe = std::exp(-dthe);
      const double omh = -std::expm1(-dth), omhe = -std::expm1(-dthe);
      const double a = omh * ehe, b = omhe * eh, sab = a + b;
      double fh, fhe, dfh_dh, dfh_dhe;
      if (sab > 0.) {
        fh = a / sab;
        fhe = b / sab;
        if (omh > 0. && omhe > 0.) {
          dfh_dh = fh * fhe / omh;
          dfh_dhe = -fh * fhe / omhe;
        } else if (omhe > 0.) {  // dth = 0: fh = 0
          dfh_dh = ehe / omhe;
          dfh_dhe = 0.;
        } else if (omh > 0.) {  // dthe = 0: fh = 1
          dfh_dh = 0.;
          dfh_dhe = -eh / omh;
        } else {
          dfh_dh = dfh_dhe = 0.;
        }
      } else if (dth > 0. || dthe > 0.) {
        fh = 1. / (1. + std::exp(dthe - dth));
        fhe = 1. - fh;
        dfh_dh = fh * fhe;
        dfh_dhe = -fh * fhe;
      } else {
        fh = fhe = dfh_dh = dfh_dhe = 0.;
      }
      // absorbed fraction; a dumped segment deposits all its photons (:782-813)
      const double E = std::exp(-dth - dthe);
      const double att = dump ? 1. : 1. - E;
      const double K = nin * att / dtv;
      const double gh = fh * K / ah, ghe = fhe * K / ahe;
      const double aK = wH * fh / ah + wHe * fhe / ahe;
      const double afh = wH * K / ah - wHe * K / ahe;
      a_ah += -wH * gh / ah;
      a_ahe += -wHe * ghe / ahe;
      const double anin = aK * att / dtv;
      const double aatt = dump ? 0. : aK * nin / dtv;
      const double adth = aatt * E + afh * dfh_dh + R[nu];
      const double adthe = aatt * E + afh * dfh_dhe + R[nu];
      anph[nu] += anin * ex;
      R[nu] += -anin * nin;
      a_ah += adth * g.sigH[nu] * dist;
      a_ahe += adthe * g.sigHe[nu] * dist;
    }
    slot_out[slot] = make_double2(a_ah, a_ahe);
  }
  // adjoint per unit of the stored photons (units of phot_norm)
  for (int nu = 0; nu < F; ++nu) a_nph[i * F + nu] += anph[nu] * FRT_PHOT_NORM;
}

__global__ void __launch_bounds__(ABLK) k_adj_cellsum(int64_t n3, const int32_t* __restrict__ cstart,
                                                      const int32_t* __restrict__ cend,
                                                      const double2* __restrict__ slot_in,
                                                      double* __restrict__ a_ah, double* __restrict__ a_ahe) {
  const int64_t k = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  if (k >= n3) return;
  double x = 0., y = 0.;
  for (int s = cstart[k]; s < cend[k]; ++s) {
    const double2 v = slot_in[s];
    x += v.x;
    y += v.y;
  }
  a_ah[k] += x;
  a_ahe[k] += y;
}

}  // namespace frt

using namespace frt;

extern "C" {

int frt_adj_enable(frt_ctx* ctx, int32_t on) {
  if (!ctx) return FRT_E_ARG;
  if (on && ctx->g.P > 127) return set_error(ctx, FRT_E_ARG, "the adjoint tape holds at most 127 segments per ray");
  if (on && !ctx->adj_tape) {
    const size_t nd = (size_t)(4 * ctx->cfg.itercount + 2) * (size_t)ctx->g.n3;
    if (hipMalloc((void**)&ctx->adj_tape, nd * sizeof(double)) != hipSuccess) {
      ctx->adj_tape = nullptr;
      return set_error(ctx, FRT_E_NOMEM, "adjoint tape");
    }
    if (hipMalloc((void**)&ctx->adj_nph, (size_t)ctx->g.cap * ctx->g.nfreq * sizeof(float)) != hipSuccess) {
      hipFree(ctx->adj_tape);
      ctx->adj_tape = nullptr;
      ctx->adj_nph = nullptr;
      return set_error(ctx, FRT_E_NOMEM, "adjoint tape (ray photons)");
    }
    if (hipMalloc((void**)&ctx->adj_np, (size_t)ctx->g.cap * ctx->cfg.itercount) != hipSuccess) {
      hipFree(ctx->adj_tape);
      hipFree(ctx->adj_nph);
      ctx->adj_tape = nullptr;
      ctx->adj_nph = nullptr;
      ctx->adj_np = nullptr;
      return set_error(ctx, FRT_E_NOMEM, "adjoint tape (dump decisions)");
    }
  }
  ctx->adj_on = on != 0;
  ctx->adj_valid = false;
  return FRT_OK;
}

int frt_adj_ray_count(const frt_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return FRT_E_ARG;
  *n = ctx->adj_valid ? ctx->adj_nrays : 0;
  return FRT_OK;
}

int frt_adj_find_avg_unu(frt_ctx* ctx, const double* a_nh1, double* g_nphot, double* g_clump, double* g_nH,
                         double* g_nHe, double* g_nH1, double* g_nHe1) {
  if (!ctx || !a_nh1) return FRT_E_ARG;
  if (!ctx->adj_valid || !ctx->adj_tape)
    return set_error(ctx, FRT_E_ARG, "no adjoint tape: call frt_adj_enable(ctx, 1), then update_rays and find_avg_unu");
  const int64_t n3 = ctx->g.n3;
  const int I = ctx->adj_iters;
  const int F = ctx->g.nfreq;
  const int64_t R = ctx->adj_nrays;
  std::vector<int32_t> perm((size_t)n3);  // internal index of each row-major cell
  for (int64_t r = 0; r < n3; ++r) perm[(size_t)r] = frt_cell_from_rowmajor(ctx->g.lay, (int)r);
  // device buffers: 11 cell arrays, the ray photons' adjoint and one
  // {HI, HeI} opacity adjoint per segment slot (the forward's slot buffers
  // hold float pairs in the mixed form, so the adjoint keeps its own)
  double* buf = nullptr;
  const size_t ncell = (size_t)11 * n3, nray = (size_t)std::max<int64_t>(R, 1) * F;
  const size_t nslot = 2 * (size_t)std::max<int64_t>(ctx->nseg, 1);
  const size_t slot_off = (ncell + nray + 1) & ~(size_t)1;  // 16-byte aligned double2 slots
  if (hipMalloc((void**)&buf, (slot_off + nslot) * sizeof(double)) != hipSuccess)
    return set_error(ctx, FRT_E_NOMEM, "adjoint buffers");
  double2* slots = (double2*)(buf + slot_off);
  double *aGH = buf, *aGHe = buf + n3, *ah0 = buf + 2 * n3, *ahe0 = buf + 3 * n3, *ah1 = buf + 4 * n3,
         *ahe1 = buf + 5 * n3, *anh1 = buf + 6 * n3, *aNH = buf + 7 * n3, *aNHe = buf + 8 * n3, *aC = buf + 9 * n3,
         *scratch = buf + 10 * n3, *aNph = buf + ncell;
  (void)scratch;
  std::vector<double> h((size_t)n3);
  for (int64_t r = 0; r < n3; ++r) h[(size_t)perm[(size_t)r]] = a_nh1[r];
  hipError_t e = hipMemcpyAsync(anh1, h.data(), n3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(aNH, 0, 3 * n3 * sizeof(double), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(aNph, 0, nray * sizeof(double), ctx->stream);
  int rc = FRT_OK;
  const unsigned gc = (unsigned)((n3 + ABLK - 1) / ABLK);
  const unsigned gr = (unsigned)((std::max<int64_t>(R, 1) + ABLK - 1) / ABLK);
  const double* x0tape = ctx->adj_tape + (size_t)4 * I * n3;
  // (a_out) adjoint of the pass's avg outputs = the next pass's opacity adjoint
  double *a_cur = ah0, *ae_cur = ahe0, *a_next = ah1, *ae_next = ahe1;
  for (int k = I - 1; k >= 0 && e == hipSuccess; --k) {
    const double* tk = ctx->adj_tape + (size_t)4 * k * n3;
    const bool last = (k == I - 1);
    if (last)
      hipLaunchKernelGGL(k_adj_chem<true>, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, ctx->g, tk, x0tape,
                         (const double*)nullptr, (const double*)nullptr, (const double*)anh1, aGH, aGHe, a_cur,
                         ae_cur, aNH, aNHe, aC);
    else
      hipLaunchKernelGGL(k_adj_chem<false>, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, ctx->g, tk, x0tape,
                         (const double*)a_next, (const double*)ae_next, (const double*)nullptr, aGH, aGHe, a_cur,
                         ae_cur, aNH, aNHe, aC);
    if (R > 0)
      hipLaunchKernelGGL(k_adj_deposit, dim3(gr), dim3(ABLK), 0, ctx->stream, ctx->g, R, ctx->seg_cnt,
                         ctx->path_cell, ctx->path_dist, ctx->path_slot, tk, ctx->adj_nph,
                         (const uint8_t*)(ctx->adj_np + (size_t)k * ctx->g.cap), (const double*)aGH,
                         (const double*)aGHe, slots, aNph);
    hipLaunchKernelGGL(k_adj_cellsum, dim3(gc), dim3(ABLK), 0, ctx->stream, n3, ctx->cell_start, ctx->cell_end,
                       (const double2*)slots, a_cur, ae_cur);
    e = hipGetLastError();
    std::swap(a_cur, a_next);
    std::swap(ae_cur, ae_next);
  }
  // after the loop a_next holds the adjoint of pass 0's opacities (n_HI, n_HeI at the start)
  std::vector<double> dv((size_t)n3 * 5);
  const double* src[5] = {aC, aNH, aNHe, a_next, ae_next};
  for (int j = 0; j < 5 && e == hipSuccess; ++j)
    e = hipMemcpyAsync(dv.data() + (size_t)j * n3, src[j], n3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  std::vector<double> rv(nray);
  if (e == hipSuccess) e = hipMemcpyAsync(rv.data(), aNph, nray * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(buf);
  rc = hip_check(ctx, e, "adjoint of find_avg_unu");
  if (rc) return rc;
  double* dst[5] = {g_clump, g_nH, g_nHe, g_nH1, g_nHe1};
  for (int j = 0; j < 5; ++j)
    if (dst[j])
      for (int64_t r = 0; r < n3; ++r) dst[j][r] = dv[(size_t)j * n3 + (size_t)perm[(size_t)r]];
  if (g_nphot) std::memcpy(g_nphot, rv.data(), (size_t)R * F * sizeof(double));
  return FRT_OK;
}

}  // extern "C"

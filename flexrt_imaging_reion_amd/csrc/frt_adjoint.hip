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

// Adjoint of one C2Ray cell pass (ray_tracing.cc:1084-1127).  In: the
// adjoint of the pass's time-averaged outputs a_out_h/he (the next pass's
// opacities, or for the last pass what the temperature update read), and for
// the last pass of the final n_HI, n_HeI.  Out: the adjoint of the Gamma sums
// (aGH, aGHe), of the pass's opacities through n_e (a_ah, a_ahe; the deposit
// part is added by k_adj_cellsum), and accumulated over the passes: nH, nHe,
// the clumping factor, the starting f_H1, f_He1 (through x0 = 1 - f) and the
// recombination coefficients (a_rH, a_rHe); the optional ones may be null.
template <bool LAST>
__global__ void __launch_bounds__(ABLK) k_adj_chem(CellPtrs cf, GridConst g, const double* __restrict__ tk,
                                                   const double* __restrict__ x0tape,
                                                   const double* __restrict__ a_out_h,
                                                   const double* __restrict__ a_out_he,
                                                   const double* __restrict__ a_nh1_out,
                                                   const double* __restrict__ a_nhe1_out,
                                                   double* __restrict__ aGH, double* __restrict__ aGHe,
                                                   double* __restrict__ a_ah, double* __restrict__ a_ahe,
                                                   double* __restrict__ a_nH, double* __restrict__ a_nHe,
                                                   double* __restrict__ a_C, double* __restrict__ a_fh,
                                                   double* __restrict__ a_fhe, double* __restrict__ a_rH,
                                                   double* __restrict__ a_rHe) {
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
  double a_u = 0., a_G = 0., a_Ghe = 0., a_cH = 0., a_cHe = 0., a_nHd = 0., a_nHed = 0., a_x0 = 0., a_x0he = 0.;
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
    a_x0 += aF * h;
    if (LAST) {
      // n_HI at the end of the step: nH (1 - FL), FL = xeq + (x0 - xeq) e (:1118)
      const double FL = xeq + (x0 - xeq) * e;
      const double aL = a_nh1_out[k];
      const double aFL = -nH * aL;
      a_nHd += (1. - FL) * aL;
      a_G += aFL * (1. - e) / pr;
      a_pr += aFL * ((1. - e) * (-G / (pr * pr)) + (x0 - xeq) * (-dt * e));
      a_x0 += aFL * e;
    }
    a_G += a_pr;  // pr = G + cH u
    a_u += a_pr * cH;
    a_cH += a_pr * u;
  }
  {
    // HeI, the same update (:1106-1113, final :1119-1120)
    const double pr = Ghe + cHe * u;
    const double xeq = Ghe / pr;
    const double e = std::exp(-dt * pr);
    const double h = (1. - e) / (dt * pr);
    const double dh = (dt * pr * e - (1. - e)) / (dt * pr * pr);
    const double Fv = xeq + (x0he - xeq) * h;
    const double ao = a_out_he ? a_out_he[k] : 0.;
    const double aF = -nHe * ao;
    a_nHed += (1. - Fv) * ao;
    double a_pr = aF * ((1. - h) * (-Ghe / (pr * pr)) + (x0he - xeq) * dh);
    a_Ghe += aF * (1. - h) / pr;
    a_x0he += aF * h;
    if (LAST && a_nhe1_out) {
      const double FL = xeq + (x0he - xeq) * e;
      const double aL = a_nhe1_out[k];
      const double aFL = -nHe * aL;
      a_nHed += (1. - FL) * aL;
      a_Ghe += aFL * (1. - e) / pr;
      a_pr += aFL * ((1. - e) * (-Ghe / (pr * pr)) + (x0he - xeq) * (-dt * e));
      a_x0he += aFL * e;
    }
    a_Ghe += a_pr;
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
  if (a_fh) {  // x0 = 1 - f_H1 at the start of the step
    a_fh[k] -= a_x0;
    a_fhe[k] -= a_x0he;
  }
  if (a_rH) {
    a_rH[k] += a_cH * (double)Cf;
    a_rHe[k] += a_cHe * (double)Cf;
  }
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
                                                      const double* __restrict__ aGnu,
                                                      const double* __restrict__ a_out,
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
  if (a_out) {
    // the ray's photons after the pass, n e^(-tau) over the processed
    // segments (ray_tracing.cc:925): adjoint of the photons in, and of tau,
    // which every processed segment's depth adds to
    for (int nu = 0; nu < F; ++nu) {
      const double ex = std::exp(-tauh[nu] - tauhe[nu]);
      const double ao = a_out[i * F + nu];
      a_nph[i * F + nu] += ao * ex;
      R[nu] = -ao * (nph[nu] / FRT_PHOT_NORM) * ex;
    }
  }
  for (int p = np - 1; p >= 0; --p) {
    const int cell = path_cell[(int64_t)p * cap + i];
    const int slot = path_slot[(int64_t)p * cap + i];
    const double dist = (double)path_dist[(int64_t)p * cap + i];
    const double ah = tah[cell], ahe = tahe[cell];
    const bool dump = dumped && (p == np - 1);
    double a_ah = 0., a_ahe = 0.;
    for (int nu = 0; nu < F; ++nu) {
      // the deposit adds to the Gamma sums and, in the last pass, to the
      // per-bin rates the temperature update reads
      const double wH = aGH[cell] + (aGnu ? aGnu[(int64_t)nu * n3 + cell] : 0.);
      const double wHe = aGHe[cell] + (aGnu ? aGnu[(int64_t)(F + nu) * n3 + cell] : 0.);
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
  if (on && (ctx->subgrid || ctx->c2ray_off))
    return set_error(ctx, FRT_E_ARG, "the adjoint is built for the C2Ray mode (subgrid = 0, c2ray_off = 0)");
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
    if (hipMalloc((void**)&ctx->adj_tstate, 5 * (size_t)ctx->g.n3 * sizeof(float)) != hipSuccess) {
      hipFree(ctx->adj_tape);
      hipFree(ctx->adj_nph);
      ctx->adj_tape = nullptr;
      ctx->adj_nph = nullptr;
      ctx->adj_tstate = nullptr;
      return set_error(ctx, FRT_E_NOMEM, "adjoint tape (thermal inputs)");
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
  if (on && ctx->adj_raydt_cap < ctx->nsrc) {
    hipFree(ctx->adj_raydt);
    ctx->adj_raydt = nullptr;
    ctx->adj_raydt_cap = 0;
    if (hipMalloc((void**)&ctx->adj_raydt, (size_t)ctx->nsrc * sizeof(float)) != hipSuccess)
      return set_error(ctx, FRT_E_NOMEM, "adjoint tape (ray_dt)");
    ctx->adj_raydt_cap = ctx->nsrc;
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
    const double* nul = nullptr;
    double* nulw = nullptr;
    if (last)
      hipLaunchKernelGGL(k_adj_chem<true>, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, ctx->g, tk, x0tape,
                         nul, nul, (const double*)anh1, nul, aGH, aGHe, a_cur, ae_cur, aNH, aNHe, aC, nulw, nulw,
                         nulw, nulw);
    else
      hipLaunchKernelGGL(k_adj_chem<false>, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, ctx->g, tk, x0tape,
                         (const double*)a_next, (const double*)ae_next, nul, nul, aGH, aGHe, a_cur, ae_cur, aNH,
                         aNHe, aC, nulw, nulw, nulw, nulw);
    if (R > 0)
      hipLaunchKernelGGL(k_adj_deposit, dim3(gr), dim3(ABLK), 0, ctx->stream, ctx->g, R, ctx->seg_cnt,
                         ctx->path_cell, ctx->path_dist, ctx->path_slot, tk, ctx->adj_nph,
                         (const uint8_t*)(ctx->adj_np + (size_t)k * ctx->g.cap), (const double*)aGH,
                         (const double*)aGHe, (const double*)nullptr, (const double*)nullptr, slots, aNph);
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

// ===========================================================================
// Adjoint across steps.  The carried adjoint (internal cell order, device)
// refers to the state after the last reversed step: the cell fields the next
// step reads (n_HI, n_HeI as its first opacities, f_H1, f_He1 through x0, n_e
// through the dn_e/dt seed, T, nH1_prev), the parameters' accumulated
// adjoints (nH, nHe, clumping) and the ray photons in the list order after
// that step.  frt_adj_step_reverse turns it into the adjoint of the state
// before the step that was just run with the tape on.
#include "frt_thermal.h"

namespace frt {

enum AdjState { S_NH1 = 0, S_NHE1, S_FH1, S_FHE1, S_NE, S_T, S_NH1PREV, P_NH, P_NHE, P_C, ADJR_NCELL };
enum AdjWork {
  W_NH1N = 0, W_NH2, W_NHE2, W_NE, W_NTOT, W_DNE, W_AH1, W_AHE1, W_NH1PREV0, W_T0, W_NHT,  // thermal adjoints
  W_GH, W_GHE, W_A0, W_AE0, W_A1, W_AE1, W_NH1OUT, W_NHE1OUT, W_RH, W_RHE, W_FH, W_FHE, W_NE0, ADJR_NWORK
};

// Cast (init_rays, ray_tracing.cc:302-304): each new ray carries
// spectrum L ray_dt / 12 / phot_norm photons; the adjoint of L gathers its 12 rays.
__global__ void __launch_bounds__(ABLK) k_adjr_cast(int nsrc, int F, int64_t base, GridConst g,
                                                    const float* __restrict__ raydt,
                                                    const double* __restrict__ a_ray, double* __restrict__ a_lum) {
  const int s = blockIdx.x * ABLK + threadIdx.x;
  if (s >= nsrc) return;
  double acc = 0.;
  for (int r = 0; r < 12; ++r) {
    const int64_t q = base + 12 * (int64_t)s + r;
    for (int nu = 0; nu < F; ++nu)
      acc += a_ray[q * F + nu] * (double)g.spectrum[nu] * (double)raydt[s] / 12. / FRT_PHOT_NORM;
  }
  a_lum[s] += acc;
}

// Inverse of the compaction's order: output q took child ch of ray i
// (sorted_src[q] = i << 2 | ch), pre-sort entry out_off[i] + ch.
__global__ void __launch_bounds__(ABLK) k_adjr_inv(int64_t total, const uint32_t* __restrict__ sorted_src,
                                                   const int32_t* __restrict__ out_off, int32_t* __restrict__ inv) {
  const int64_t q = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  if (q < total) {
    const uint32_t packed = sorted_src[q];
    inv[out_off[packed >> 2] + (int)(packed & 3u)] = (int32_t)q;
  }
}

// split_rays and the compaction (rt_funcs.cc:211-237, ray_tracing.cc:482-619):
// a kept ray's photons are copied, a split ray's children carry a quarter each,
// a removed ray none; the adjoint of ray i gathers its outputs in child order.
__global__ void __launch_bounds__(ABLK) k_adjr_split(int64_t n, int F, const int32_t* __restrict__ n_out,
                                                     const int32_t* __restrict__ out_off,
                                                     const int32_t* __restrict__ inv, const double* __restrict__ a_post,
                                                     double* __restrict__ a_pre) {
  const int64_t i = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  if (i >= n) return;
  const int no = n_out[i];
  const double w = no == 4 ? 0.25 : 1.;
  for (int nu = 0; nu < F; ++nu) {
    double acc = 0.;
    for (int c = 0; c < no; ++c) acc += a_post[(int64_t)inv[out_off[i] + c] * F + nu];
    a_pre[i * F + nu] = w * acc;
  }
}

// Rays that did not move keep their photons through find_avg_unu.
__global__ void __launch_bounds__(ABLK) k_adjr_rays_in(int64_t n, int F, const int32_t* __restrict__ seg_cnt,
                                                       const double* __restrict__ a_out, double* __restrict__ a_in) {
  const int64_t i = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  if (i >= n) return;
  const bool moved = seg_cnt[i] > 0;
  for (int nu = 0; nu < F; ++nu) a_in[i * F + nu] = moved ? 0. : a_out[i * F + nu];
}

// d ln(rate)/d ln(T) of the recombination coefficients and cooling rates
// (rates.cc:128-231; Hui & Gnedin 1997 fits, x = 2 E_th/(k_B T) falls as 1/T).
__device__ __forceinline__ double dlog_hg(double x, double a, double x0, double b, double p) {
  // rate ~ x^a / (1 + (x/x0)^b)^p: d ln / d ln x, then d ln x / d ln T = -1
  const double q = dm_pow(x / x0, b);
  return -(a - p * b * q / (1. + q));
}

__device__ __forceinline__ double dalpha_dT(const GridConst& g, double T, bool helium) {
  // derivative of the stored coefficient at T (0 below the floor, where Tm is clamped)
  if (T < (double)g.temp_min) return 0.;
  if (!helium) {
    if (g.rec_case_b) return frt_alphaB_H2(T) / T * dlog_hg(frt_lambda(T, FRT_ETH_H1), 1.500, 2.740, 0.407, 2.2420);
    return frt_alphaA_H2(T) / T * dlog_hg(frt_lambda(T, FRT_ETH_H1), 1.503, 0.522, 0.470, 1.923);
  }
  double d = g.rec_case_b ? frt_alphaB_He2(T) / T * (-0.750) : frt_alphaA_He2(T) / T * (-0.654);
  if (T >= 3e4 && T <= 1e6) {
    const double x = frt_lambda(T, FRT_ETH_HE2);
    const double e2 = dm_exp(-0.15 * x / 2);
    d += frt_Dalpha_He2(T) / T * (-1.5 + 0.375 * x + 0.3 * 0.075 * x * e2 / (1. + 0.3 * e2));
  }
  return d;
}

// Cooling of the sub-steps (update_heat_cool, gas_funcs.cc:184-241) and its
// derivatives in T (through Tm = max(temp_min, T)), n_e, nH2, nHe2.
struct CoolDiff {
  double c, dT, dne, dnH2, dnHe2;
};

__device__ __forceinline__ CoolDiff cooling_diff(const GridConst& g, double T, double ne, double nH2, double nHe2) {
  const bool clamped = T < (double)g.temp_min;
  const double Tm = clamped ? (double)g.temp_min : T;
  CoolDiff r = {0., 0., 0., 0., 0.};
  if (g.recomb_cool) {
    const double rH = g.rec_case_b ? frt_rcrB_H2(Tm) : frt_rcrA_H2(Tm);
    const double rHe = g.rec_case_b ? frt_rcrB_He2(Tm) : frt_rcrA_He2(Tm);
    const double lH = g.rec_case_b ? 1. + dlog_hg(frt_lambda(Tm, FRT_ETH_H1), 1.970, 2.250, 0.376, 3.720)
                                   : 1. + dlog_hg(frt_lambda(Tm, FRT_ETH_H1), 1.965, 0.541, 0.502, 2.697);
    const double lHe = g.rec_case_b ? 1. - 0.750 : 1. - 0.654;
    r.c += rH * ne * nH2 + rHe * ne * nHe2;
    r.dT += (rH * lH * ne * nH2 + rHe * lHe * ne * nHe2) / Tm;
    r.dne += rH * nH2 + rHe * nHe2;
    r.dnH2 += rH * ne;
    r.dnHe2 += rHe * ne;
  }
  if (g.free_free) {
    const double ff = frt_free_free_rate(Tm);
    r.c += ff * ne * (nH2 + nHe2);
    r.dT += 0.5 * ff / Tm * ne * (nH2 + nHe2);
    r.dne += ff * (nH2 + nHe2);
    r.dnH2 += ff * ne;
    r.dnHe2 += ff * ne;
  }
  if (g.compton) {
    r.c -= g.compton_coef * ne * (g.t_cmb - Tm);
    r.dT += g.compton_coef * ne;
    r.dne -= g.compton_coef * (g.t_cmb - Tm);
  }
  if (clamped) r.dT = 0.;
  return r;
}

// Adjoint of one cell's temperature update (gas_funcs.cc:362-419 with
// set_Treion_default :104-129), replayed from the taped start values so that
// the sub-step schedule, the post-front mixing and the floor are the
// forward's own (held fixed).  a_T: adjoint of T after the step.
__global__ void __launch_bounds__(ABLK) k_adjr_thermal(CellPtrs cf, GridConst g, const float* __restrict__ ts,
                                                       const double* __restrict__ a_Tout, double* __restrict__ W,
                                                       double* __restrict__ a_gnu) {
  const int64_t k = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  const int64_t n3 = g.n3;
  if (k >= n3) return;
  const int F = g.nfreq;
  double* w = W;
  auto wr = [&](int f) -> double& { return w[(int64_t)f * n3 + k]; };
  for (int f = W_NH1N; f <= W_NHT; ++f) wr(f) = 0.;
  for (int nu = 0; nu < 2 * F; ++nu) a_gnu[(int64_t)nu * n3 + k] = 0.;
  if (!g.temp_ev) return;  // T set by a threshold: no derivative
  ThermalIn in;
  in.nH = cf.f[C_NH][k];
  in.nHe = cf.f[C_NHE][k];
  in.nH1 = cf.f[C_NH1][k];
  in.nH2 = cf.f[C_NH2][k];
  in.nHe2 = cf.f[C_NHE2][k];
  in.ne = cf.f[C_NE][k];
  in.ntot = cf.f[C_NTOT][k];
  in.dne = cf.f[C_DNE_DT][k];
  in.fH2 = cf.f[C_FH2][k];
  in.ah1 = cf.f[C_AVG_NH1][k];
  in.ahe1 = cf.f[C_AVG_NHE1][k];
  in.nH1_prev = ts[3 * n3 + k];
  thermal_density_terms(cf, g, k, in);
  const float T0 = ts[k];
  float temp = T0, heat = ts[n3 + k], cool = ts[2 * n3 + k], ifs, inc;
  ThermalRec rec;
  const int m = thermal_cell<true>(g, in, cf.gamma_h1_nu, cf.gamma_he1_nu, n3, k, temp, heat, cool, ifs, inc, &rec);
  const double ntot = in.ntot, dne = in.dne, ne = in.ne, nH2 = in.nH2, nHe2 = in.nHe2;
  const double k23 = 2. / 3. / FRT_K_B / ntot;
  const double Hd = (double)g.hubble;
  double a = (rec.t_raw < g.temp_min) ? 0. : a_Tout[k];  // the floor
  double a_heat = 0., a_k23 = 0., a_ne = 0., a_nH2 = 0., a_nHe2 = 0., a_ntot = 0., a_dne = 0.;
  for (int s = m - 1; s >= 0; --s) {
    const double Tp = rec.t_in[s], dts = rec.dt[s];
    double D = 1. + 2. * Hd * dts + 1. / ntot * dne * dts;
    if (in.ddelta != 0.f) D = D - 2. * (double)in.ddelta / 3. / (double)in.rho_rel * dts;
    const CoolDiff cd = cooling_diff(g, Tp, ne, nH2, nHe2);
    const double num = Tp + k23 * ((double)rec.heat - cd.c) * dts;
    const double Tk = num / D;
    const double a_num = a / D, a_D = -a * Tk / D;
    a_heat += a_num * k23 * dts;
    const double a_cool = -a_num * k23 * dts;
    a_k23 += a_num * ((double)rec.heat - cd.c) * dts;
    a_ne += a_cool * cd.dne;
    a_nH2 += a_cool * cd.dnH2;
    a_nHe2 += a_cool * cd.dnHe2;
    a_dne += a_D * dts / ntot;
    a_ntot += a_D * (-dne * dts / (ntot * ntot));
    a = a_num + a_cool * cd.dT;  // d num / d Tp = 1 - k23 dcool/dT dts
  }
  a_ntot += a_k23 * (-k23 / ntot);
  // photo-heating, the same in every sub-step (update_heat_cool :149-163)
  if ((double)in.fH2 > 0.98 && m > 0) {
    double aah = 0., aahe = 0.;
    for (int nu = 0; nu < F; ++nu) {
      const double EH = FRT_H_PLANCK * (FRT_H_EV * g.nu[nu] - FRT_ETH_H1) / FRT_H_EV;
      const double EHe = FRT_H_PLANCK * (FRT_H_EV * g.nu[nu] - FRT_ETH_HE1) / FRT_H_EV;
      const double gH = cf.gamma_h1_nu[nu * n3 + k], gHe = cf.gamma_he1_nu[nu * n3 + k];
      aah += a_heat * gH * EH;
      aahe += a_heat * gHe * EHe;
      a_gnu[(int64_t)nu * n3 + k] = a_heat * (double)in.ah1 * EH;
      a_gnu[(int64_t)(F + nu) * n3 + k] = a_heat * (double)in.ahe1 * EHe;
    }
    wr(W_AH1) = aah;
    wr(W_AHE1) = aahe;
  }
  // post-front mixing (set_Treion_default): T = (fo T0 + (fn - fo) T_re)/fn
  const double nH = in.nH;
  const double fn = (double)((in.nH - in.nH1) / in.nH), fo = (double)((in.nH - in.nH1_prev) / in.nH);
  double a_T0 = a, a_fn = 0., a_fo = 0.;
  if ((fn > fo) && (fn < 0.98)) {
    const double v = (double)g.dx * (fn - fo) / (double)g.dt;
    const double L = dm_log(v / 1e5);
    const double Cn[5] = {9.5432, -1.6441e-2, -9.8010e-3, 4.1664e-3, -2.2710e-4};
    double lnT = 0., dlnT = 0.;
    for (int t = 0; t < 5; ++t) {
      lnT += Cn[t] * dm_pow(L, (double)t);
      if (t > 0) dlnT += t * Cn[t] * dm_pow(L, (double)(t - 1));
    }
    const double TR = dm_exp(lnT);
    const double dTR_dv = TR * dlnT / v;
    const double dv = (double)g.dx / (double)g.dt;
    const double gnum = fo * (double)T0 + (fn - fo) * TR;
    a_T0 = a * fo / fn;
    a_fn = a * ((TR + (fn - fo) * dTR_dv * dv) / fn - gnum / (fn * fn));
    a_fo = a * ((double)T0 - TR - (fn - fo) * dTR_dv * dv) / fn;
  }
  wr(W_NH1N) = -a_fn / nH;
  wr(W_NH1PREV0) = -a_fo / nH;
  wr(W_NH2) = a_nH2;
  wr(W_NHE2) = a_nHe2;
  wr(W_NE) = a_ne;
  wr(W_NTOT) = a_ntot;
  wr(W_DNE) = a_dne;
  wr(W_T0) = a_T0;
  // nH enters fn and fo: d/d nH of (nH - x)/nH = x/nH^2
  wr(W_NHT) = a_fn * (double)in.nH1 / (nH * nH) + a_fo * (double)in.nH1_prev / (nH * nH);
}

// From the carried adjoint of the state after the step and the thermal
// adjoints to the adjoint of find_avg_unu's outputs (the last pass's final
// n_HI, n_HeI and averages), through f = n/nH, ne = nH2 + nHe2,
// n_tot = nH + nHe + ne, dne_dt = (ne - ne_start)/dt (ray_tracing.cc:1139-1209).
__global__ void __launch_bounds__(ABLK) k_adjr_combine(CellPtrs cf, GridConst g, double* __restrict__ S,
                                                       double* __restrict__ W) {
  const int64_t k = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  const int64_t n3 = g.n3;
  if (k >= n3) return;
  auto s = [&](int f) -> double& { return S[(int64_t)f * n3 + k]; };
  auto w = [&](int f) -> double& { return W[(int64_t)f * n3 + k]; };
  const double nH = cf.f[C_NH][k], nHe = cf.f[C_NHE][k];
  const double nH1 = cf.f[C_NH1][k], nHe1 = cf.f[C_NHE1][k];
  const double a_ne = s(S_NE) + w(W_NE) + w(W_NTOT) + w(W_DNE) / (double)g.dt;
  const double a_nH2 = w(W_NH2) + a_ne, a_nHe2 = w(W_NHE2) + a_ne;
  w(W_NH1OUT) = s(S_NH1) + s(S_FH1) / nH + s(S_NH1PREV) + w(W_NH1N) - a_nH2;
  w(W_NHE1OUT) = s(S_NHE1) + s(S_FHE1) / nHe - a_nHe2;
  s(P_NH) += a_nH2 + w(W_NTOT) - s(S_FH1) * nH1 / (nH * nH) + w(W_NHT);
  s(P_NHE) += a_nHe2 + w(W_NTOT) - s(S_FHE1) * nHe1 / (nHe * nHe);
  w(W_NE0) = -w(W_DNE) / (double)g.dt;  // dne_dt seed -ne_start
  w(W_RH) = w(W_RHE) = w(W_FH) = w(W_FHE) = 0.;
}

// The carried adjoint of the state before the step.
__global__ void __launch_bounds__(ABLK) k_adjr_finish(GridConst g, const float* __restrict__ ts,
                                                      double* __restrict__ S, const double* __restrict__ W,
                                                      const double* __restrict__ a_op_h,
                                                      const double* __restrict__ a_op_he) {
  const int64_t k = (int64_t)blockIdx.x * ABLK + threadIdx.x;
  const int64_t n3 = g.n3;
  if (k >= n3) return;
  auto s = [&](int f) -> double& { return S[(int64_t)f * n3 + k]; };
  auto w = [&](int f) -> double { return W[(int64_t)f * n3 + k]; };
  const double T0 = ts[k];
  s(S_NH1) = a_op_h[k];
  s(S_NHE1) = a_op_he[k];
  s(S_FH1) = w(W_FH);
  s(S_FHE1) = w(W_FHE);
  s(S_NE) = w(W_NE0);
  s(S_T) = w(W_T0) + w(W_RH) * dalpha_dT(g, T0, false) + w(W_RHE) * dalpha_dT(g, T0, true);
  s(S_NH1PREV) = w(W_NH1PREV0);
}

}  // namespace frt

extern "C" {

int frt_adj_run_begin(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  if (!ctx->adj_tape) return set_error(ctx, FRT_E_STATE, "call frt_adj_enable(ctx, 1) first");
  const size_t n3 = (size_t)ctx->g.n3, F = (size_t)ctx->g.nfreq, cap = (size_t)ctx->g.cap;
  if (!ctx->adjr_cell) {
    hipError_t e = hipMalloc((void**)&ctx->adjr_cell, ADJR_NCELL * n3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->adjr_ray, cap * F * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->adjr_ray2, cap * F * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->adjr_work, (ADJR_NWORK + 2 * F) * n3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->adjr_inv, cap * sizeof(int32_t));
    if (e != hipSuccess) return set_error(ctx, FRT_E_NOMEM, "adjoint run buffers");
  }
  hipFree(ctx->adjr_lum);
  ctx->adjr_lum = nullptr;
  if (hipMalloc((void**)&ctx->adjr_lum, std::max<size_t>(ctx->nsrc, 1) * sizeof(double)) != hipSuccess)
    return set_error(ctx, FRT_E_NOMEM, "adjoint run buffers");
  hipError_t e = hipMemsetAsync(ctx->adjr_cell, 0, ADJR_NCELL * n3 * sizeof(double), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(ctx->adjr_lum, 0, std::max<size_t>(ctx->nsrc, 1) * sizeof(double), ctx->stream);
  if (e == hipSuccess && ctx->nrays > 0)
    e = hipMemsetAsync(ctx->adjr_ray, 0, (size_t)ctx->nrays * F * sizeof(double), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->adjr_nray = ctx->nrays;
  ctx->adjr_step = ctx->step;
  ctx->adjr_on = e == hipSuccess;
  return hip_check(ctx, e, "adjoint run begin");
}

// Adds to the carried adjoint of the current state: a_nH1 and a_nH (row-major
// cells, either may be NULL), e.g. the gradient of a loss of x_HI = nH1/nH.
int frt_adj_seed(frt_ctx* ctx, const double* a_nH1, const double* a_nH) {
  if (!ctx) return FRT_E_ARG;
  if (!ctx->adjr_on) return set_error(ctx, FRT_E_STATE, "call frt_adj_run_begin first");
  const int64_t n3 = ctx->g.n3;
  std::vector<double> cur((size_t)n3), add((size_t)n3);
  for (int j = 0; j < 2; ++j) {
    const double* src = j == 0 ? a_nH1 : a_nH;
    if (!src) continue;
    double* dst = ctx->adjr_cell + (size_t)(j == 0 ? S_NH1 : P_NH) * n3;
    hipError_t e = stream_copy(ctx, cur.data(), dst, n3 * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_check(ctx, e, "adjoint seed");
    for (int64_t r = 0; r < n3; ++r) cur[(size_t)frt_cell_from_rowmajor(ctx->g.lay, (int)r)] += src[r];
    e = stream_copy(ctx, dst, cur.data(), n3 * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_check(ctx, e, "adjoint seed");
  }
  return FRT_OK;
}

int frt_adj_step_reverse(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  if (!ctx->adjr_on) return set_error(ctx, FRT_E_STATE, "call frt_adj_run_begin first");
  if (!ctx->adj_valid || ctx->adj_split_n < 0 || ctx->adj_split_gen != ctx->split_gen)
    return set_error(ctx, FRT_E_STATE, "frt_adj_step_reverse must follow a frt_step run with the tape on");
  if (ctx->step != ctx->adjr_step)
    return set_error(ctx, FRT_E_STATE, "the step just run does not end at the state the carried adjoint refers to");
  if (ctx->stats.rays_merged_away > 0)
    return set_error(ctx, FRT_E_STATE, "the adjoint across steps does not cover merging (rays were merged in this step)");
  const int64_t n3 = ctx->g.n3, F = ctx->g.nfreq;
  const int64_t n_pre = ctx->adj_split_n, total = ctx->adj_split_total;
  const int64_t base = ctx->adj_cast_base;
  if (ctx->adjr_nray != ctx->nrays || (base >= 0 && base != total) || n_pre != ctx->adj_nrays)
    return set_error(ctx, FRT_E_STATE, "the carried ray adjoint does not match this step's rays");
  const unsigned gc = (unsigned)((n3 + ABLK - 1) / ABLK);
  auto gr = [](int64_t n) { return (unsigned)((std::max<int64_t>(n, 1) + ABLK - 1) / ABLK); };
  double* S = ctx->adjr_cell;
  double* W = ctx->adjr_work;
  double* Gnu = W + (size_t)ADJR_NWORK * n3;
  auto wk = [&](int f) { return W + (size_t)f * n3; };
  GridConst g = ctx->g;
  // cast, split and compaction, in reverse
  if (base >= 0)
    hipLaunchKernelGGL(k_adjr_cast, dim3(gr(ctx->nsrc)), dim3(ABLK), 0, ctx->stream, ctx->nsrc, (int)F, base, g,
                       (const float*)ctx->adj_raydt, (const double*)ctx->adjr_ray, ctx->adjr_lum);
  if (total > 0)
    hipLaunchKernelGGL(k_adjr_inv, dim3(gr(total)), dim3(ABLK), 0, ctx->stream, total,
                       (const uint32_t*)ctx->split_sorted, (const int32_t*)ctx->out_off, ctx->adjr_inv);
  hipLaunchKernelGGL(k_adjr_split, dim3(gr(n_pre)), dim3(ABLK), 0, ctx->stream, n_pre, (int)F,
                     (const int32_t*)ctx->n_out, (const int32_t*)ctx->out_off, (const int32_t*)ctx->adjr_inv,
                     (const double*)ctx->adjr_ray, ctx->adjr_ray2);
  // temperature update, then the outputs of find_avg_unu
  hipLaunchKernelGGL(k_adjr_thermal, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, g, (const float*)ctx->adj_tstate,
                     (const double*)(S + (size_t)S_T * n3), W, Gnu);
  hipLaunchKernelGGL(k_adjr_combine, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, g, S, W);
  // find_avg_unu's passes, last first; rays: adjr_ray2 holds the outgoing
  // photons' adjoint, adjr_ray collects the incoming
  hipLaunchKernelGGL(k_adjr_rays_in, dim3(gr(n_pre)), dim3(ABLK), 0, ctx->stream, n_pre, (int)F,
                     (const int32_t*)ctx->seg_cnt, (const double*)ctx->adjr_ray2, ctx->adjr_ray);
  const int I = ctx->adj_iters;
  const double* x0tape = ctx->adj_tape + (size_t)4 * I * n3;
  double* slots = nullptr;
  if (hipMalloc((void**)&slots, 2 * (size_t)std::max<int64_t>(ctx->nseg, 1) * sizeof(double)) != hipSuccess)
    return set_error(ctx, FRT_E_NOMEM, "adjoint slot buffer");
  double *a_cur = wk(W_A0), *ae_cur = wk(W_AE0), *a_next = wk(W_A1), *ae_next = wk(W_AE1);
  const double* nul = nullptr;
  for (int k = I - 1; k >= 0; --k) {
    const double* tk = ctx->adj_tape + (size_t)4 * k * n3;
    const bool last = (k == I - 1);
    if (last)
      hipLaunchKernelGGL(k_adj_chem<true>, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, g, tk, x0tape,
                         (const double*)wk(W_AH1), (const double*)wk(W_AHE1), (const double*)wk(W_NH1OUT),
                         (const double*)wk(W_NHE1OUT), wk(W_GH), wk(W_GHE), a_cur, ae_cur, S + (size_t)P_NH * n3,
                         S + (size_t)P_NHE * n3, S + (size_t)P_C * n3, wk(W_FH), wk(W_FHE), wk(W_RH), wk(W_RHE));
    else
      hipLaunchKernelGGL(k_adj_chem<false>, dim3(gc), dim3(ABLK), 0, ctx->stream, ctx->c, g, tk, x0tape,
                         (const double*)a_next, (const double*)ae_next, nul, nul, wk(W_GH), wk(W_GHE), a_cur, ae_cur,
                         S + (size_t)P_NH * n3, S + (size_t)P_NHE * n3, S + (size_t)P_C * n3, wk(W_FH), wk(W_FHE),
                         wk(W_RH), wk(W_RHE));
    if (n_pre > 0)
      hipLaunchKernelGGL(k_adj_deposit, dim3(gr(n_pre)), dim3(ABLK), 0, ctx->stream, g, n_pre, ctx->seg_cnt,
                         ctx->path_cell, ctx->path_dist, ctx->path_slot, tk, ctx->adj_nph,
                         (const uint8_t*)(ctx->adj_np + (size_t)k * ctx->g.cap), (const double*)wk(W_GH),
                         (const double*)wk(W_GHE), last ? (const double*)Gnu : nul,
                         last ? (const double*)ctx->adjr_ray2 : nul, (double2*)slots, ctx->adjr_ray);
    hipLaunchKernelGGL(k_adj_cellsum, dim3(gc), dim3(ABLK), 0, ctx->stream, n3, ctx->cell_start, ctx->cell_end,
                       (const double2*)slots, a_cur, ae_cur);
    std::swap(a_cur, a_next);
    std::swap(ae_cur, ae_next);
  }
  // a_next holds the adjoint of pass 0's opacities: n_HI, n_HeI at the start
  hipLaunchKernelGGL(k_adjr_finish, dim3(gc), dim3(ABLK), 0, ctx->stream, g, (const float*)ctx->adj_tstate, S,
                     (const double*)W, (const double*)a_next, (const double*)ae_next);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(slots);
  ctx->adjr_nray = n_pre;
  ctx->adjr_step -= 1;
  ctx->adj_valid = false;  // one reverse per taped step
  return hip_check(ctx, e, "adjoint step reverse");
}

// Adjoints of the parameters after reversing every step back to the initial
// state: the rays then held are the initial cast (ray_dt = initial ray_dt,
// 12 rays per source in source order), so their adjoint goes to the source
// rates; g_lum = d/dL (per source, photons/s), g_clump and g_log_rho per
// row-major cell, the latter with every density of the initial state scaling
// with rho (nH, nHe, n_HI, n_HeI, n_e, nH1_prev; fractions and T fixed).
// Any output may be NULL.  init_ray_dt is the ray_dt the initial cast used.
int frt_adj_run_finish(frt_ctx* ctx, float init_ray_dt, double* g_lum, double* g_clump, double* g_log_rho) {
  if (!ctx) return FRT_E_ARG;
  if (!ctx->adjr_on) return set_error(ctx, FRT_E_STATE, "call frt_adj_run_begin first");
  const int64_t n3 = ctx->g.n3, F = ctx->g.nfreq;
  if (ctx->step != ctx->adjr_step || ctx->adjr_nray != ctx->nsrc * 12 || ctx->nrays != ctx->adjr_nray)
    return set_error(ctx, FRT_E_STATE,
                     "reverse every step, then return the context to its initial state (its checkpoint): "
                     "that state holds the initial cast only");
  std::vector<float> rdt((size_t)std::max<int64_t>(ctx->nsrc, 1), init_ray_dt);
  float* d_rdt = nullptr;
  if (hipMalloc((void**)&d_rdt, rdt.size() * sizeof(float)) != hipSuccess)
    return set_error(ctx, FRT_E_NOMEM, "adjoint finish");
  hipError_t e = stream_copy(ctx, d_rdt, rdt.data(), rdt.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && ctx->nsrc > 0)
    hipLaunchKernelGGL(k_adjr_cast, dim3((unsigned)((ctx->nsrc + ABLK - 1) / ABLK)), dim3(ABLK), 0, ctx->stream,
                       ctx->nsrc, (int)F, (int64_t)0, ctx->g, (const float*)d_rdt, (const double*)ctx->adjr_ray,
                       ctx->adjr_lum);
  std::vector<double> S((size_t)ADJR_NCELL * n3);
  std::vector<float> fld[6];
  const int ff[6] = {C_NH, C_NHE, C_NH1, C_NHE1, C_NE, C_NH1_PREV};
  if (e == hipSuccess) e = stream_copy(ctx, S.data(), ctx->adjr_cell, S.size() * sizeof(double), hipMemcpyDeviceToHost);
  for (int j = 0; j < 6 && e == hipSuccess; ++j) {
    fld[j].resize((size_t)n3);
    e = stream_copy(ctx, fld[j].data(), ctx->c.f[ff[j]], n3 * sizeof(float), hipMemcpyDeviceToHost);
  }
  std::vector<double> L((size_t)std::max<int64_t>(ctx->nsrc, 1));
  if (e == hipSuccess) e = stream_copy(ctx, L.data(), ctx->adjr_lum, L.size() * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(d_rdt);
  if (e != hipSuccess) return hip_check(ctx, e, "adjoint finish");
  if (g_lum) std::memcpy(g_lum, L.data(), (size_t)ctx->nsrc * sizeof(double));
  for (int64_t r = 0; r < n3; ++r) {
    const size_t c = (size_t)frt_cell_from_rowmajor(ctx->g.lay, (int)r);
    auto s = [&](int f) { return S[(size_t)f * n3 + c]; };
    if (g_clump) g_clump[r] = s(P_C);
    if (g_log_rho)
      g_log_rho[r] = (s(P_NH)) * fld[0][c] + s(P_NHE) * fld[1][c] + s(S_NH1) * fld[2][c] + s(S_NHE1) * fld[3][c] +
                     s(S_NE) * fld[4][c] + s(S_NH1PREV) * fld[5][c];
  }
  return FRT_OK;
}

}  // extern "C"

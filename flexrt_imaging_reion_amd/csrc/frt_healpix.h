// frt_healpix.h -- HEALPix NESTED pixelisation, host and device.
//
// Written from the definitions of Gorski et al. 2005 (ApJ 622, 759):
// 12 base faces, Nside = 2^order, ring index i and in-ring longitude from the
// face coordinates (ix, iy), and the inverse through the equatorial lines
// jp/jm (|z| <= 2/3) or the polar-cap sigma coordinate (|z| > 2/3).
//
// The reference calls HEALPix C++ 3.40 (T_Healpix_Base<int>(order, NEST):
// pix2vec at ray_tracing.cc:290,444,461,520,533 and vec2pix at 357,459).
// That library is not vendored, so this file stands in for it on both the GPU
// path and the CPU oracle.  All trigonometry goes through frt_detmath.h so the
// two sides produce identical bits.
#pragma once

#include "frt_detmath.h"

namespace frt_hpx {

// Face layout: ring offset (in units of Nside) and longitude offset (in units
// of pi/4 per Nside) of each base face's lowest corner.
FRT_HD int face_ring(int f) { return (f < 4) ? 2 : ((f < 8) ? 3 : 4); }
FRT_HD int face_lon(int f) {
  int m = f & 3;
  return (f >= 4 && f < 8) ? 2 * m : 2 * m + 1;
}

// Bit interleaving helpers for the nested index (ix in even bits, iy in odd).
FRT_HD uint32_t spread_bits(uint32_t v) {
  v &= 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
FRT_HD uint32_t compact_bits(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

FRT_HD int xyf_to_nest(int order, int ix, int iy, int face) {
  return (face << (2 * order)) + (int)(spread_bits((uint32_t)ix) | (spread_bits((uint32_t)iy) << 1));
}

}  // namespace frt_hpx

// Centre of NESTED pixel `pix` at `order` as a unit vector (double).
FRT_HD void hpx_pix2vec_nest(int order, int pix, double* vx, double* vy, double* vz) {
  using namespace frt_hpx;
  const int nside = 1 << order;
  const int face = pix >> (2 * order);
  const uint32_t ipf = (uint32_t)pix & (uint32_t)(nside * nside - 1);
  const int ix = (int)compact_bits(ipf);
  const int iy = (int)compact_bits(ipf >> 1);
  // ring index counted from the north pole, 1 .. 4 nside - 1
  const int ring = face_ring(face) * nside - ix - iy - 1;
  int nr;          // number of pixels per quarter ring
  double z, sth;   // cos(theta), sin(theta)
  bool pole;       // use the accurate sin(theta) form near the poles
  if (ring < nside) {
    nr = ring;
    double q = (double)(nr * nr) / (3.0 * (double)nside * (double)nside);
    z = 1.0 - q;
    pole = (z > 0.99);
    sth = pole ? sqrt(q * (2.0 - q)) : 0.0;
  } else if (ring > 3 * nside) {
    nr = 4 * nside - ring;
    double q = (double)(nr * nr) / (3.0 * (double)nside * (double)nside);
    z = q - 1.0;
    pole = (z < -0.99);
    sth = pole ? sqrt(q * (2.0 - q)) : 0.0;
  } else {
    nr = nside;
    z = (double)(2 * (2 * nside - ring)) / (3.0 * (double)nside);
    pole = false;
    sth = 0.0;
  }
  // longitude in units of pi/(4 nr)
  int t = face_lon(face) * nr + ix - iy;
  if (t < 0) t += 8 * nr;
  const double phi = (0.7853981633974483 * (double)t) / (double)nr;
  if (!pole) sth = sqrt((1.0 - z) * (1.0 + z));
  double s, c;
  dm_sincos(phi, &s, &c);
  *vx = sth * c;
  *vy = sth * s;
  *vz = z;
}

// NESTED pixel at `order` containing direction (x, y, z) (need not be unit).
FRT_HD int hpx_vec2pix_nest(int order, double x, double y, double z) {
  using namespace frt_hpx;
  const int nside = 1 << order;
  const double rxy2 = x * x + y * y;
  const double inv_len = 1.0 / sqrt(rxy2 + z * z);
  const double zn = z * inv_len;
  const double za = fabs(zn);
  double tt = dm_atan2(y, x) * 0.6366197723675814;  // longitude / (pi/2)
  if (tt < 0.0) tt += 4.0;
  if (tt >= 4.0) tt -= 4.0;
  if (za <= 0.6666666666666666) {
    // equatorial belt: the two families of face edges through the point
    const double a = (double)nside * (0.5 + tt);
    const double b = (double)nside * (zn * 0.75);
    const int jp = (int)(a - b);
    const int jm = (int)(a + b);
    const int fp = jp >> order;
    const int fm = jm >> order;
    int face;
    if (fp == fm)
      face = fp | 4;
    else if (fp < fm)
      face = fp;
    else
      face = fm + 8;
    const int ix = jm & (nside - 1);
    const int iy = nside - (jp & (nside - 1)) - 1;
    return xyf_to_nest(order, ix, iy, face);
  }
  // polar caps: sigma = nside * sqrt(3 (1 - |z|)), split by longitude fraction
  int ntt = (int)tt;
  if (ntt > 3) ntt = 3;
  const double tp = tt - (double)ntt;
  double sig;
  if (za < 0.99)
    sig = (double)nside * sqrt(3.0 * (1.0 - za));
  else
    sig = (double)nside * (sqrt(rxy2) * inv_len) / sqrt((1.0 + za) / 3.0);
  int jp = (int)(tp * sig);
  int jm = (int)((1.0 - tp) * sig);
  if (jp > nside - 1) jp = nside - 1;
  if (jm > nside - 1) jm = nside - 1;
  if (zn >= 0.0) return xyf_to_nest(order, nside - jm - 1, nside - jp - 1, ntt);
  return xyf_to_nest(order, jp, jm, ntt + 8);
}

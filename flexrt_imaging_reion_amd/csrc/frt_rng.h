// frt_rng.h -- counter-based rotation-index stream, host and device.
//
// The reference draws the random rotation index of each ray cast and each
// merge pass with rand() % Ntheta (general_funcs.cc:170-173, used at
// ray_tracing.cc:233 and :327) after a time-based srand (general_funcs.cc:164),
// so no two runs agree.  Here the index is a splitmix64 hash of
// (seed, stream, counter): stream = cast pass or merge pass, counter = source
// rank.  It is identical on the CPU and the GPU and independent of thread
// order.
#pragma once

#include "frt_detmath.h"

#define FRT_STREAM_MERGE (1ull << 62)

FRT_HD uint32_t frt_itheta_hash(uint64_t seed, uint64_t stream, uint64_t counter) {
  uint64_t z = seed ^ (stream * 0x9E3779B97F4A7C15ull) ^ (counter * 0xD1B54A32D192ED03ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)(z % 42875ull);
}

// Opacity dither (a measurement of rounding sensitivity, off unless a seed is
// set): after each C2Ray pass the stored per-cell opacity is moved by one
// float ulp, up or down by one hashed bit of (seed, step, pass, row-major
// cell).  One ulp is the rounding of the reference's float storage of
// avg_nh1 / avg_nhe1 (global_variables.cc), so the spread this causes is the
// part of any difference between two precision forms that no form with float
// storage can remove.
#define FRT_STREAM_DITHER (1ull << 61)

FRT_HD bool frt_dither_up(uint64_t seed, uint64_t step, int pass, uint64_t cell, int which) {
  uint64_t z = seed ^ ((FRT_STREAM_DITHER + step * 64u + (uint64_t)pass * 2u + (uint64_t)which) * 0x9E3779B97F4A7C15ull) ^
               (cell * 0xD1B54A32D192ED03ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (z >> 17) & 1u;
}

// v moved one ulp up or down (positive finite v; 0 and non-finite v are kept).
FRT_HD float frt_ulp_step(float v, bool up) {
  if (!(v > 0.f) || !(v < 3.4e38f)) return v;
  uint32_t b;
  __builtin_memcpy(&b, &v, 4);
  b = up ? b + 1u : b - 1u;
  float r;
  __builtin_memcpy(&r, &b, 4);
  return r;
}

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

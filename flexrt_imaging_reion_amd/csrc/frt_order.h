// frt_order.h -- canonical order of the ray list.
//
// After every compaction the rays are stably sorted by the 4x4x4 tile of
// their cell (tiles numbered row-major like cells); within a tile they keep
// their previous relative order.  Rays that stay inside their tile from one
// step to the next therefore keep their place, so the compaction gather reads
// mostly consecutive input.  Shared by the GPU library and the CPU checker.
#pragma once

#ifdef __HIPCC__
#define FRT_ORDER_HD __host__ __device__
#else
#define FRT_ORDER_HD
#endif

#define FRT_TILE_SHIFT 2

FRT_ORDER_HD inline int frt_tiles_per_side(int n) { return (n + (1 << FRT_TILE_SHIFT) - 1) >> FRT_TILE_SHIFT; }

FRT_ORDER_HD inline unsigned frt_tile_key(int cell, int n) {
  const int i = cell / (n * n);
  const int rem = cell - i * n * n;
  const int j = rem / n;
  const int k = rem - j * n;
  const int t = frt_tiles_per_side(n);
  return (unsigned)((((i >> FRT_TILE_SHIFT) * t) + (j >> FRT_TILE_SHIFT)) * t + (k >> FRT_TILE_SHIFT));
}

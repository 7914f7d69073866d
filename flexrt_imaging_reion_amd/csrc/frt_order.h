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

// Internal cell numbering of the GPU library.  With n a multiple of 4 the
// grid is stored tile by tile: the 64 cells of a 4x4x4 tile are contiguous
// (local index (i&3)*16 + (j&3)*4 + (k&3)) and tiles follow in the row-major
// tile order above, so a tile is one 256-byte run of every float field and
// the ray-list tile key is simply index >> 6.  Other n use the row-major
// numbering i*n*n + j*n + k of the reference.  The C ABI always exchanges
// fields and cell indices in the reference's row-major numbering.
struct FrtCellLayout {
  int n;
  int tiled;  // 1: 4x4x4 tiles, 0: row-major
  int t;      // tiles per side
  int tshift; // log2(t) when t is a power of two, else -1
};

FRT_ORDER_HD inline FrtCellLayout frt_cell_layout(int n) {
  FrtCellLayout L;
  L.n = n;
  L.tiled = (n % 4 == 0) ? 1 : 0;
  L.t = n >> 2;
  L.tshift = -1;
  for (int s = 0; s < 30; ++s)
    if ((1 << s) == L.t) L.tshift = s;
  return L;
}

FRT_ORDER_HD inline int frt_cell_index(const FrtCellLayout& L, int i, int j, int k) {
  if (!L.tiled) return (i * L.n + j) * L.n + k;
  return ((((i >> 2) * L.t + (j >> 2)) * L.t + (k >> 2)) << 6) | ((i & 3) << 4) | ((j & 3) << 2) | (k & 3);
}

FRT_ORDER_HD inline void frt_cell_ijk(const FrtCellLayout& L, int c, int& i, int& j, int& k) {
  if (!L.tiled) {
    i = c / (L.n * L.n);
    const int rem = c - i * L.n * L.n;
    j = rem / L.n;
    k = rem - j * L.n;
    return;
  }
  const int tile = c >> 6, loc = c & 63;
  int ti, tj, tk;
  if (L.tshift >= 0) {  // the same values by shifts
    const int m = L.t - 1;
    tk = tile & m;
    tj = (tile >> L.tshift) & m;
    ti = tile >> (2 * L.tshift);
  } else {
    tk = tile % L.t;
    const int tij = tile / L.t;
    tj = tij % L.t;
    ti = tij / L.t;
  }
  i = (ti << 2) | (loc >> 4);
  j = (tj << 2) | ((loc >> 2) & 3);
  k = (tk << 2) | (loc & 3);
}

// Tile key of a cell given in the internal numbering.
FRT_ORDER_HD inline unsigned frt_tile_key_internal(const FrtCellLayout& L, int c) {
  if (L.tiled) return (unsigned)(c >> 6);
  int i, j, k;
  frt_cell_ijk(L, c, i, j, k);
  return frt_tile_key((i * L.n + j) * L.n + k, L.n);
}

FRT_ORDER_HD inline int frt_cell_to_rowmajor(const FrtCellLayout& L, int c) {
  int i, j, k;
  frt_cell_ijk(L, c, i, j, k);
  return (i * L.n + j) * L.n + k;
}

FRT_ORDER_HD inline int frt_cell_from_rowmajor(const FrtCellLayout& L, int r) {
  const int i = r / (L.n * L.n);
  const int rem = r - i * L.n * L.n;
  const int j = rem / L.n;
  const int k = rem - j * L.n;
  return frt_cell_index(L, i, j, k);
}

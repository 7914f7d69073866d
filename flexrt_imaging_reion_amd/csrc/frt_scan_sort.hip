// frt_scan_sort.hip -- exclusive scan and stable LSD radix sort whose item
// counts live in device memory.
//
// The step's scans (segments per ray, moved rays, merge candidates, split
// children) and sorts (segments by cell, rays by tile, merge candidates by
// bin) run on counts that an earlier kernel of the same step produced.
// Reading those counts back to size the launches would stop the stream at
// every scan, and a captured step graph could not run at all; here every
// kernel reads its count from device memory and the grid is a fixed number
// of blocks sized from a host-side upper bound.
//
// Scan: reduce-then-scan.  Block b of G takes a contiguous range of whole
// tiles; pass 1 sums each range, pass 2 scans the G sums in one block, pass 3
// rescans each range from its offset.  Sort: per digit of up to 8 bits, pass
// 1 builds each block's digit histogram, the digit-major (digit, block)
// table is scanned, and pass 3 scatters each block's range tile by tile with
// the ranks of a stable in-tile ranking (wavefront ballots per digit bit,
// then a prefix over the block's wavefronts), so the order of equal keys is
// their input order and the result equals any stable sort's.
//
// Cell sort of the step's segments: the radix passes order the segments by
// group of 64 cells (key >> 6; a 4x4x4 tile in the tiled cell numbering),
// then one block per group places the group's segments by cell in LDS-ranked
// rounds, in their order within the group, and writes each segment's final
// slot and the cells' slot ranges.  The result equals the stable sort by
// the whole cell key; the last pass's scattered writes stay within a
// group's rays instead of spanning the whole list.
#include <algorithm>

#include "frt_internal.h"

namespace frt {

static constexpr int SBLK = 256;            // threads per block
static constexpr int SCAN_ITEMS = 8;        // values per thread per scan tile
static constexpr int SCAN_TILE = SBLK * SCAN_ITEMS;
static constexpr int SORT_ROUNDS = 8;       // SBLK-key rounds per sort tile
static constexpr int SORT_TILE = SBLK * SORT_ROUNDS;
static constexpr int RADIX_MAX_BITS = 8;

static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ int64_t dev_count(const int64_t* dcount, int64_t add) {
  return (dcount ? *dcount : 0) + add;
}

// Block barrier for LDS traffic only: global loads issued before it stay in
// flight (a full __syncthreads also waits for every outstanding global
// access).  The kernels here share only LDS data between threads.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Range [lo, hi) of block b when `count` items are dealt to G blocks in
// whole tiles.
__device__ __forceinline__ void block_range(int64_t count, int G, int tile, int b, int64_t& lo, int64_t& hi) {
  const int64_t tiles = (count + tile - 1) / tile;
  const int64_t per = (tiles + G - 1) / G;
  lo = min((int64_t)b * per * tile, count);
  hi = min(lo + per * tile, count);
}

// Inclusive wavefront scan of v (64 lanes).
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive block scan of v; *total receives the block sum.  All threads call.
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[SBLK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  lds_barrier();
  int64_t before = 0, all = 0;
#pragma unroll
  for (int q = 0; q < SBLK / 64; ++q) {
    const int64_t s = wsum[q];
    before += q < w ? s : 0;
    all += s;
  }
  lds_barrier();
  *total = all;
  return before + inc - v;
}

// Scanned value of item i.  With MOVED (the segment scan) the value is the
// count plus 2^32 for each of the first *dcount items with a positive count,
// so one scan gives both the count offsets (low 32 bits; their total stays
// below 2^31) and the rank of every such item among them (high 32 bits).
// Scan modes: plain offsets; the segment scan with the moved-ray list
// (above); the compaction of flagged items (0/1 flags) into a list with
// their keys, writing no offsets.
enum ScanMode { SCAN_PLAIN = 0, SCAN_MOVED = 1, SCAN_COMPACT = 2 };

template <int MODE>
__device__ __forceinline__ int64_t scan_value(int32_t v, int64_t i, int64_t nflag) {
  return MODE == SCAN_MOVED ? (int64_t)v + ((i < nflag && v > 0) ? (int64_t(1) << 32) : 0) : (int64_t)v;
}

template <int MODE>
__global__ void __launch_bounds__(SBLK) k_scan_reduce(const int32_t* __restrict__ in, const int64_t* dcount,
                                                      int64_t add, int64_t* __restrict__ part) {
  const int64_t count = dev_count(dcount, add);
  const int64_t nflag = dev_count(dcount, 0);
  int64_t lo, hi;
  block_range(count, gridDim.x, SCAN_TILE, blockIdx.x, lo, hi);
  int64_t s = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += SBLK) s += scan_value<MODE>(in[i], i, nflag);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ int64_t ws[SBLK / 64];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int q = 0; q < SBLK / 64; ++q) t += ws[q];
    part[blockIdx.x] = t;
  }
}

// One block: exclusive scan of the G block sums in place; the grand total to
// *total64 (if set) and to part[G]; with total_hi set, the packed total of
// the segment scan is split: low 32 bits to *total64, high to *total_hi.
__global__ void __launch_bounds__(SBLK) k_scan_parts(int64_t* __restrict__ part, int G, int64_t* total64,
                                                     int64_t* total_hi) {
  int64_t carry = 0;
  for (int base = 0; base < G; base += SBLK) {
    const int i = base + threadIdx.x;
    const int64_t v = i < G ? part[i] : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan(v, &tot);
    if (i < G) part[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    part[G] = carry;
    if (total_hi) {
      *total64 = carry & 0xffffffffll;
      *total_hi = carry >> 32;
    } else if (total64) {
      *total64 = carry;
    }
  }
}

// SCAN_MOVED: out receives the low half of the scan and every flagged item
// i is written to list[rank] (the items with a positive count).
// SCAN_COMPACT: each item i with flag 1 goes to list[rank] = i with
// kdst[rank] = ksrc[i]; out is not written.
template <int MODE>
__global__ void __launch_bounds__(SBLK) k_scan_down(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                    const int64_t* dcount, int64_t add,
                                                    const int64_t* __restrict__ part, int32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ ksrc, uint32_t* __restrict__ kdst) {
  const int64_t count = dev_count(dcount, add);
  const int64_t nflag = dev_count(dcount, 0);
  int64_t lo, hi;
  block_range(count, gridDim.x, SCAN_TILE, blockIdx.x, lo, hi);
  int64_t carry = part[blockIdx.x];
  for (int64_t base = lo; base < hi; base += SCAN_TILE) {
    // thread t holds items base + t*ITEMS .. + ITEMS-1 (consecutive)
    const int64_t first = base + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      v[k] = first + k < hi ? scan_value<MODE>(in[first + k], first + k, nflag) : 0;
      s += v[k];
    }
    int64_t tot;
    int64_t run = carry + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      if (first + k < hi) {
        if (MODE == SCAN_COMPACT) {
          if (v[k]) {
            list[run] = (int32_t)(first + k);
            kdst[run] = ksrc[first + k];
          }
        } else {
          out[first + k] = (int32_t)(MODE == SCAN_MOVED ? (run & 0xffffffffll) : run);
          if (MODE == SCAN_MOVED && (v[k] >> 32)) list[run >> 32] = (int32_t)(first + k);
        }
      }
      run += v[k];
    }
    carry += tot;
  }
}

// ---- radix sort ----------------------------------------------------------------

// Lanes of this wavefront holding the same d (of nbits bits) as this lane,
// among the lanes where valid is set.
__device__ __forceinline__ uint64_t match_digit(int d, int nbits, bool valid) {
  uint64_t m = __ballot(valid);
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (d >> b) & 1;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

// Digit histogram of one pass over the block ranges of that pass's input,
// digit-major [d][block].
__global__ void __launch_bounds__(SBLK) k_rs_hist(const uint32_t* __restrict__ key, const int64_t* dcount,
                                                  int64_t add, int shift, int nbits, int32_t* __restrict__ hist) {
  __shared__ int32_t h[1 << RADIX_MAX_BITS];
  const int R = 1 << nbits;
  for (int d = threadIdx.x; d < R; d += SBLK) h[d] = 0;
  __syncthreads();
  const int64_t count = dev_count(dcount, add);
  int64_t lo, hi;
  block_range(count, gridDim.x, SORT_TILE, blockIdx.x, lo, hi);
  const uint32_t mask = (uint32_t)R - 1u;
  const int lane = threadIdx.x & 63;
  // SORT_ROUNDS loads in flight per thread; then one atomic per run of equal
  // digits across neighbouring lanes (the high-digit passes see nearly
  // sorted keys, where a lane-per-atomic histogram would serialize the
  // wavefront on one LDS address)
  for (int64_t tile = lo; tile < hi; tile += SORT_TILE) {  // block-uniform trip count
    int dj[SORT_ROUNDS];
#pragma unroll
    for (int j = 0; j < SORT_ROUNDS; ++j) {
      const int64_t i = tile + (int64_t)j * SBLK + threadIdx.x;
      dj[j] = i < hi ? (int)((key[i] >> shift) & mask) : -1;
    }
#pragma unroll
    for (int j = 0; j < SORT_ROUNDS; ++j) {
      const int d = dj[j];
      const bool valid = d >= 0;  // the valid lanes are a prefix of the wavefront
      const int dprev = __shfl_up(d, 1, 64);
      const bool head = valid && (lane == 0 || dprev != d);
      const uint64_t heads = __ballot(head);
      const uint64_t vmask = __ballot(valid);
      if (head) {
        const uint64_t after = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const int next = after ? __ffsll((unsigned long long)after) - 1 : __popcll(vmask);
        atomicAdd(&h[d], next - lane);
      }
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += SBLK) hist[(int64_t)d * gridDim.x + blockIdx.x] = h[d];
}

// One digit pass.  Block b walks its range in tiles of SORT_TILE keys; each
// wavefront ranks its own quarter of a tile (rounds of 64 consecutive keys:
// ballots give the rank among the lanes, the wavefront's private running
// digit counts in LDS the rank within its quarter -- no block barrier
// between rounds).  The per-digit counts of the four quarters then give each
// quarter's offset within the digit's run, the tile is reordered by digit in
// LDS and written out in that order, so consecutive threads write
// consecutive addresses of each run.  Running digit offsets carry the
// block's position across tiles, and the scanned digit-major table orders
// the blocks: equal digits keep their input order.
// (held to 5 wavefronts per SIMD: 104 -> 94 VGPRs, no spills, 6 % faster)
__global__ void __launch_bounds__(SBLK, 5) k_rs_scatter(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                     uint32_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                     const int64_t* dcount, int64_t add, int shift, int nbits,
                                                     const int32_t* __restrict__ hoff) {
  constexpr int W = SBLK / 64;
  constexpr int RMAX = 1 << RADIX_MAX_BITS;
  constexpr int QUARTER = SORT_TILE / W;  // keys per wavefront per tile
  constexpr int ROUNDS = QUARTER / 64;
  static_assert(RMAX <= SBLK, "one thread per digit");
  __shared__ int32_t goff[RMAX];    // this block's next output position of each digit
  __shared__ int32_t tcnt[RMAX];    // keys of each digit in the tile
  __shared__ int32_t tstart[RMAX];  // start of each digit's run in the reordered tile
  __shared__ int32_t wc[W][RMAX];   // per wavefront: running digit counts, then its offset in each run
  __shared__ uint32_t sk[SORT_TILE];
  __shared__ int32_t sv[SORT_TILE];
  const int R = 1 << nbits;
  const uint32_t mask = (uint32_t)R - 1u;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (t < R) goff[t] = hoff[(int64_t)t * gridDim.x + blockIdx.x];
  const int64_t count = dev_count(dcount, add);
  int64_t lo, hi;
  block_range(count, gridDim.x, SORT_TILE, blockIdx.x, lo, hi);
  // the tile's keys and values are loaded one tile ahead
  uint32_t kn[ROUNDS];
  int32_t vn[ROUNDS];
#pragma unroll
  for (int j = 0; j < ROUNDS; ++j) {
    const int64_t i = lo + (int64_t)w * QUARTER + j * 64 + lane;
    kn[j] = i < hi ? kin[i] : 0u;
    vn[j] = i < hi ? vin[i] : 0;
  }
  for (int64_t tile = lo; tile < hi; tile += SORT_TILE) {
    volatile int32_t* mine = wc[w];
    for (int d = lane; d < R; d += 64) mine[d] = 0;
    __builtin_amdgcn_wave_barrier();
    const int64_t qbase = tile + (int64_t)w * QUARTER;
    uint32_t k[ROUNDS];
    int32_t v[ROUNDS];
    int lr[ROUNDS], dd[ROUNDS];
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
      k[j] = kn[j];
      v[j] = vn[j];
      const int64_t i = qbase + SORT_TILE + j * 64 + lane;
      kn[j] = i < hi ? kin[i] : 0u;
      vn[j] = i < hi ? vin[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
      const bool valid = qbase + j * 64 + lane < hi;
      const int d = (int)((k[j] >> shift) & mask);
      const uint64_t m = match_digit(d, nbits, valid);
      const int rank = __popcll(m & lt);
      int base = 0;
      if (valid) base = mine[d];
      __builtin_amdgcn_wave_barrier();
      if (valid && rank == 0) mine[d] = base + __popcll(m);
      __builtin_amdgcn_wave_barrier();
      lr[j] = base + rank;
      dd[j] = valid ? d : -1;
    }
    lds_barrier();
    // each quarter's offset within the digit's run, and the tile's digit counts
    int total = 0;
    if (t < R) {
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const int c = wc[q][t];
        wc[q][t] = total;
        total += c;
      }
      tcnt[t] = total;
    }
    int64_t tot_tile;
    const int64_t ex = block_excl_scan(t < R ? total : 0, &tot_tile);
    if (t < R) tstart[t] = (int32_t)ex;
    lds_barrier();
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j)
      if (dd[j] >= 0) {
        const int p = tstart[dd[j]] + wc[w][dd[j]] + lr[j];
        sk[p] = k[j];
        sv[p] = v[j];
      }
    lds_barrier();
    const int ntile = (int)min((int64_t)SORT_TILE, hi - tile);
    for (int p = t; p < ntile; p += SBLK) {
      const uint32_t kk = sk[p];
      const int d = (int)((kk >> shift) & mask);
      const int pos = goff[d] + (p - tstart[d]);
      kout[pos] = kk;
      vout[pos] = sv[p];
    }
    lds_barrier();
    if (t < R) goff[t] += tcnt[t];
  }
}

// ---- cell sort: groups of 64 cells ----------------------------------------------

static constexpr int GROUP_SHIFT = 6;  // cells per group = 64 = one wavefront

__global__ void __launch_bounds__(SBLK) k_zero_i32(int32_t* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * SBLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SBLK) a[i] = 0;
}

// [start, end) of every non-empty group's run in the group-sorted keys
// (empty groups keep the zeros they were cleared to).  Each thread takes
// four consecutive keys; the neighbours across threads come from the
// adjacent lanes, and only a wavefront's edge lanes load one more key.
__global__ void __launch_bounds__(SBLK) k_group_ranges(const uint32_t* __restrict__ key, const int64_t* dcount,
                                                       int32_t* __restrict__ gstart, int32_t* __restrict__ gend) {
  const int64_t n = dev_count(dcount, 0);
  const int lane = threadIdx.x & 63;
  const int64_t nq = (n + 3) >> 2;  // groups of four keys
  // block-uniform trip count: the shuffles see every lane
  for (int64_t qb = (int64_t)blockIdx.x * SBLK; qb < nq; qb += (int64_t)gridDim.x * SBLK) {
    const int64_t q = qb + threadIdx.x;
    const int64_t s0 = q << 2;
    uint32_t g[4];
    if (s0 + 3 < n) {
      const uint4 k4 = *reinterpret_cast<const uint4*>(key + s0);  // 16-byte aligned: s0 is a multiple of 4
      g[0] = k4.x >> GROUP_SHIFT;
      g[1] = k4.y >> GROUP_SHIFT;
      g[2] = k4.z >> GROUP_SHIFT;
      g[3] = k4.w >> GROUP_SHIFT;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) g[j] = s0 + j < n ? key[s0 + j] >> GROUP_SHIFT : 0xffffffffu;
    }
    // the group of the key before s0 and after s0 + 3 (none: a value no group has)
    uint32_t before = __shfl_up(g[3], 1, 64);
    uint32_t after = __shfl_down(g[0], 1, 64);
    if (lane == 0) before = s0 > 0 && s0 - 1 < n ? key[s0 - 1] >> GROUP_SHIFT : 0xffffffffu;
    if (lane == 63) after = s0 + 4 < n ? key[s0 + 4] >> GROUP_SHIFT : 0xffffffffu;
    if (s0 < n) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t s = s0 + j;
        if (s >= n) break;
        const uint32_t prev = j == 0 ? before : g[j - 1];
        const uint32_t next = j == 3 ? after : g[j + 1];
        if (s == 0 || prev != g[j]) gstart[g[j]] = (int32_t)s;
        if (s == n - 1 || next != g[j]) gend[g[j]] = (int32_t)(s + 1);
      }
    }
  }
}

// One block per group of 64 cells: the cells' segment counts (LDS atomics),
// their slot ranges [cstart, cend) from an exclusive scan, then every
// segment's slot: its cell's start, plus the segments of that cell earlier
// in the group run -- counted over the previous rounds (running per-cell
// position), the earlier wavefronts of this round and the earlier lanes of
// its own wavefront (ballot match on the 6 cell bits).  inv[value] = slot.
// Blocks are remapped so each XCD places a contiguous eighth of the groups:
// neighbouring groups hold neighbouring rays, whose slot entries share
// cache lines.
__global__ void __launch_bounds__(SBLK) k_group_local(const uint32_t* __restrict__ key,
                                                      const int32_t* __restrict__ val,
                                                      const int32_t* __restrict__ gstart,
                                                      const int32_t* __restrict__ gend, int32_t* __restrict__ inv,
                                                      int32_t* __restrict__ cstart, int32_t* __restrict__ cend,
                                                      int64_t ncells) {
  constexpr int W = SBLK / 64;
  __shared__ int32_t next[64];   // next slot of each cell
  __shared__ int32_t wc[W][64];  // this round: segments of each cell per wavefront
  const unsigned G = gridDim.x, b = blockIdx.x, G8 = G & ~7u;
  const int64_t g = b >= G8 ? b : (int64_t)(b & 7u) * (G8 >> 3) + (b >> 3);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int s0 = gstart[g], s1 = gend[g];
  if (t < 64) next[t] = 0;
  __syncthreads();
  // the cell bits of the first KEEP rounds stay in registers for the placement
  constexpr int KEEP = 8;
  int kd[KEEP];
#pragma unroll
  for (int r = 0; r < KEEP; ++r) {
    const int s = s0 + r * SBLK + t;
    kd[r] = s < s1 ? (int)(key[s] & 63u) : 0;
  }
#pragma unroll
  for (int r = 0; r < KEEP; ++r)
    if (s0 + r * SBLK + t < s1) atomicAdd(&next[kd[r]], 1);
  for (int s = s0 + KEEP * SBLK + t; s < s1; s += SBLK) atomicAdd(&next[key[s] & 63u], 1);
  __syncthreads();
  if (w == 0) {
    const int c = next[lane];
    int inc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int u = __shfl_up(inc, off, 64);
      if (lane >= off) inc += u;
    }
    const int first = s0 + inc - c;
    const int64_t cell = (g << GROUP_SHIFT) + lane;
    if (cell < ncells) {  // empty cells: (0, 0)
      cstart[cell] = c ? first : 0;
      cend[cell] = c ? first + c : 0;
    }
    next[lane] = first;
  }
  __syncthreads();
  int r = 0;
  for (int base = s0; base < s1; base += SBLK, ++r) {  // block-uniform trip count
    const int s = base + t;
    const bool valid = s < s1;
    int d = 0;
#pragma unroll
    for (int j = 0; j < KEEP; ++j)
      if (r == j) d = kd[j];
    if (r >= KEEP) d = valid ? (int)(key[s] & 63u) : 0;
    const int32_t v = valid ? val[s] : 0;
    const uint64_t m = match_digit(d, GROUP_SHIFT, valid);
    wc[w][lane] = 0;
    __builtin_amdgcn_wave_barrier();
    if (valid && __popcll(m & lt) == 0) wc[w][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int pos = next[d] + __popcll(m & lt);
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      inv[v] = pos;
    }
    __syncthreads();
    if (t < 64) {
      int tot = 0;
#pragma unroll
      for (int q = 0; q < W; ++q) tot += wc[q][t];
      next[t] += tot;
    }
    __syncthreads();
  }
}

// ---- host side ---------------------------------------------------------------

static int scan_grid(int64_t bound) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(std::max<int64_t>(bound, 1), SCAN_TILE), FRT_SCAN_MAX_GRID));
}

static int sort_grid(int64_t bound) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(std::max<int64_t>(bound, 1), SORT_TILE), FRT_SORT_MAX_GRID));
}

#define SS_LAUNCH(ctx, name, grid, kern, ...)                                            \
  do {                                                                                   \
    ProfRec _rec;                                                                        \
    prof_begin(ctx, name, &_rec);                                                        \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(SBLK), 0, (ctx)->stream, __VA_ARGS__);     \
    prof_end(ctx, &_rec);                                                                \
    hipError_t _e = hipGetLastError();                                                   \
    if (_e != hipSuccess) return hip_check(ctx, _e, name);                               \
  } while (0)

int dev_scan(frt_ctx* ctx, const int32_t* in, int32_t* out, const int64_t* dcount, int64_t add, int64_t bound,
             int64_t* dtotal, const char* name) {
  const int G = scan_grid(bound);
  int64_t* part = ctx->ss_part;
  SS_LAUNCH(ctx, name, G, k_scan_reduce<SCAN_PLAIN>, in, dcount, add, part);
  SS_LAUNCH(ctx, name, 1, k_scan_parts, part, G, dtotal, (int64_t*)nullptr);
  SS_LAUNCH(ctx, name, G, k_scan_down<SCAN_PLAIN>, in, out, dcount, add, (const int64_t*)part, (int32_t*)nullptr,
            (const uint32_t*)nullptr, (uint32_t*)nullptr);
  return FRT_OK;
}

int dev_compact(frt_ctx* ctx, const int32_t* flag, const uint32_t* ksrc, int32_t* list, uint32_t* kdst,
                const int64_t* dcount, int64_t add, int64_t bound, int64_t* dtotal, const char* name) {
  const int G = scan_grid(bound);
  int64_t* part = ctx->ss_part;
  SS_LAUNCH(ctx, name, G, k_scan_reduce<SCAN_COMPACT>, flag, dcount, add, part);
  SS_LAUNCH(ctx, name, 1, k_scan_parts, part, G, dtotal, (int64_t*)nullptr);
  SS_LAUNCH(ctx, name, G, k_scan_down<SCAN_COMPACT>, flag, (int32_t*)nullptr, dcount, add, (const int64_t*)part,
            list, ksrc, kdst);
  return FRT_OK;
}

int dev_scan_moved(frt_ctx* ctx, const int32_t* in, int32_t* out, int32_t* moved, const int64_t* dcount,
                   int64_t add, int64_t bound, int64_t* dtotal, int64_t* dmoved, const char* name) {
  const int G = scan_grid(bound);
  int64_t* part = ctx->ss_part;
  SS_LAUNCH(ctx, name, G, k_scan_reduce<SCAN_MOVED>, in, dcount, add, part);
  SS_LAUNCH(ctx, name, 1, k_scan_parts, part, G, dtotal, dmoved);
  SS_LAUNCH(ctx, name, G, k_scan_down<SCAN_MOVED>, in, out, dcount, add, (const int64_t*)part, moved,
            (const uint32_t*)nullptr, (uint32_t*)nullptr);
  return FRT_OK;
}

int dev_sort_pairs(frt_ctx* ctx, uint32_t* key[2], int32_t* val[2], const int64_t* dcount, int64_t add, int64_t bound,
                   int bits, int* which, const char* name, int lo) {
  bits = std::max(bits - lo, 1);
  struct {
    int passes;
    int shift[4];
    int nbits[4];
  } plan{};
  plan.passes = (bits + RADIX_MAX_BITS - 1) / RADIX_MAX_BITS;
  if (plan.passes > 4) return set_error(ctx, FRT_E_ARG, "sort keys wider than 32 bits");
  for (int p = 0, sh = 0; p < plan.passes; ++p) {
    plan.nbits[p] = (bits - sh + (plan.passes - p) - 1) / (plan.passes - p);  // spread the bits evenly
    plan.shift[p] = lo + sh;
    sh += plan.nbits[p];
  }
  const int G = sort_grid(bound);
  int src = 0;
  for (int p = 0; p < plan.passes; ++p) {
    const int64_t R = (int64_t)1 << plan.nbits[p];
    // per-block counts of this pass's input order (the previous pass moved the keys)
    SS_LAUNCH(ctx, name, G, k_rs_hist, (const uint32_t*)key[src], dcount, add, plan.shift[p], plan.nbits[p],
              ctx->ss_hist);
    int rc = dev_scan(ctx, ctx->ss_hist, ctx->ss_hoff, nullptr, R * G, R * G, nullptr, name);
    if (rc) return rc;
    SS_LAUNCH(ctx, name, G, k_rs_scatter, (const uint32_t*)key[src], (const int32_t*)val[src], key[1 - src],
              val[1 - src], dcount, add, plan.shift[p], plan.nbits[p], (const int32_t*)ctx->ss_hoff);
    src = 1 - src;
  }
  *which = src;
  return FRT_OK;
}

int dev_sort_cells(frt_ctx* ctx, uint32_t* key[2], int32_t* val[2], const int64_t* dcount, int64_t bound,
                   int key_bits, int64_t ncells, int32_t* inv, int32_t* cstart, int32_t* cend, const char* name) {
  int which = 0;
  if (key_bits > GROUP_SHIFT) {
    int rc = dev_sort_pairs(ctx, key, val, dcount, 0, bound, key_bits, &which, name, GROUP_SHIFT);
    if (rc) return rc;
  }
  const int64_t G = (ncells + 63) >> GROUP_SHIFT;
  int32_t* gstart = ctx->ss_grp;
  int32_t* gend = ctx->ss_grp + G;
  SS_LAUNCH(ctx, name, (unsigned)std::min<int64_t>(ceil_div(2 * G, SBLK), 4096), k_zero_i32, ctx->ss_grp, 2 * G);
  SS_LAUNCH(ctx, name, (unsigned)std::min<int64_t>(ceil_div(std::max<int64_t>(bound, 1), SBLK), 4 * 4096),
            k_group_ranges, (const uint32_t*)key[which], dcount, gstart, gend);
  SS_LAUNCH(ctx, name, (unsigned)G, k_group_local, (const uint32_t*)key[which], (const int32_t*)val[which],
            (const int32_t*)gstart, (const int32_t*)gend, inv, cstart, cend, ncells);
  return FRT_OK;
}

int alloc_scan_sort(frt_ctx* ctx) {
  if (ctx->ss_part) return FRT_OK;
  const size_t nh = ((size_t)1 << RADIX_MAX_BITS) * FRT_SORT_MAX_GRID;
  hipError_t e = hipMalloc((void**)&ctx->ss_part, (FRT_SCAN_MAX_GRID + 1) * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&ctx->ss_hist, nh * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&ctx->ss_hoff, nh * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&ctx->ss_grp, 2 * (size_t)((ctx->g.n3 + 63) >> GROUP_SHIFT) * sizeof(int32_t));
  if (e != hipSuccess) return set_error(ctx, FRT_E_NOMEM, "scan/sort workspace");
  return FRT_OK;
}

void free_scan_sort(frt_ctx* ctx) {
  hipFree(ctx->ss_part);
  hipFree(ctx->ss_hist);
  hipFree(ctx->ss_hoff);
  hipFree(ctx->ss_grp);
  ctx->ss_grp = nullptr;
  ctx->ss_part = nullptr;
  ctx->ss_hist = nullptr;
  ctx->ss_hoff = nullptr;
}

}  // namespace frt

using namespace frt;

extern "C" {

// Runs the library's device-count exclusive scan on counts[0..n) and its
// stable radix sort on (keys[i], i), i < n, over the low `bits` key bits,
// with the count n held in device memory as in a step; results to host
// arrays (any may be NULL).
int frt_selftest_scan_sort(frt_ctx* ctx, int64_t n, int32_t bits, const uint32_t* keys, uint32_t* keys_out,
                           int32_t* perm_out, const int32_t* counts, int32_t* scan_out, int64_t* total) {
  if (!ctx || n < 0 || n >= (1ll << 31) || bits < 1 || bits > 32) return FRT_E_ARG;
  const size_t m = (size_t)std::max<int64_t>(n, 1);
  uint32_t* k[2] = {nullptr, nullptr};
  int32_t* v[2] = {nullptr, nullptr};
  int32_t *ci = nullptr, *co = nullptr;
  int64_t* dn = nullptr;
  hipError_t e = hipMalloc((void**)&k[0], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&k[1], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&v[0], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&v[1], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&ci, m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&co, m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dn, 2 * sizeof(int64_t));
  int rc = FRT_OK;
  if (e != hipSuccess) rc = set_error(ctx, FRT_E_NOMEM, "selftest buffers");
  std::vector<int32_t> iota(m);
  for (size_t i = 0; i < m; ++i) iota[i] = (int32_t)i;
  int64_t hn[2] = {n, 0};
  if (rc == FRT_OK) {
    e = stream_copy(ctx, dn, hn, sizeof(hn), hipMemcpyHostToDevice);
    if (e == hipSuccess && keys) e = stream_copy(ctx, k[0], keys, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = stream_copy(ctx, v[0], iota.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && counts) e = stream_copy(ctx, ci, counts, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_check(ctx, e, "selftest upload");
  }
  if (rc == FRT_OK && counts) rc = dev_scan(ctx, ci, co, dn, 0, n, dn + 1, "selftest_scan");
  int which = 0;
  if (rc == FRT_OK && keys) rc = dev_sort_pairs(ctx, k, v, dn, 0, n, bits, &which, "selftest_sort");
  if (rc == FRT_OK) {
    e = hipSuccess;
    if (keys && keys_out) e = stream_copy(ctx, keys_out, k[which], (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && keys && perm_out) e = stream_copy(ctx, perm_out, v[which], (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && counts && scan_out) e = stream_copy(ctx, scan_out, co, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && total) e = stream_copy(ctx, total, dn + 1, sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_check(ctx, e, "selftest download");
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(k[0]); hipFree(k[1]); hipFree(v[0]); hipFree(v[1]); hipFree(ci); hipFree(co); hipFree(dn);
  return rc;
}

// Runs the step's cell sort on (cells[i], i), i < n, cells[i] < the
// context's cell count: slot_out[i] = sorted position of item i, and every
// cell's [cstart_out, cend_out) (arrays of the cell count; any may be NULL).
int frt_selftest_sort_cells(frt_ctx* ctx, int64_t n, const uint32_t* cells, int32_t* slot_out, int32_t* cstart_out,
                            int32_t* cend_out) {
  if (!ctx || n < 0 || n >= (1ll << 31) || (n > 0 && !cells)) return FRT_E_ARG;
  const int64_t n3 = ctx->g.n3;
  for (int64_t i = 0; i < n; ++i)
    if ((int64_t)cells[i] >= n3) return set_error(ctx, FRT_E_ARG, "selftest_sort_cells: cell out of range");
  const size_t m = (size_t)std::max<int64_t>(n, 1);
  uint32_t* k[2] = {nullptr, nullptr};
  int32_t* v[2] = {nullptr, nullptr};
  int32_t *slot = nullptr, *cs = nullptr, *ce = nullptr;
  int64_t* dn = nullptr;
  hipError_t e = hipMalloc((void**)&k[0], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&k[1], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&v[0], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&v[1], m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&slot, m * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&cs, (size_t)n3 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&ce, (size_t)n3 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dn, sizeof(int64_t));
  int rc = FRT_OK;
  if (e != hipSuccess) rc = set_error(ctx, FRT_E_NOMEM, "selftest buffers");
  if (rc == FRT_OK) rc = alloc_scan_sort(ctx);
  std::vector<int32_t> iota(m);
  for (size_t i = 0; i < m; ++i) iota[i] = (int32_t)i;
  if (rc == FRT_OK) {
    e = stream_copy(ctx, dn, &n, sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = stream_copy(ctx, k[0], cells, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = stream_copy(ctx, v[0], iota.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_check(ctx, e, "selftest upload");
  }
  if (rc == FRT_OK)
    rc = dev_sort_cells(ctx, k, v, dn, n, ctx->key_bits, n3, slot, cs, ce, "selftest_sort_cells");
  if (rc == FRT_OK) {
    e = hipSuccess;
    if (slot_out && n) e = stream_copy(ctx, slot_out, slot, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && cstart_out) e = stream_copy(ctx, cstart_out, cs, (size_t)n3 * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && cend_out) e = stream_copy(ctx, cend_out, ce, (size_t)n3 * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_check(ctx, e, "selftest download");
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(k[0]); hipFree(k[1]); hipFree(v[0]); hipFree(v[1]); hipFree(slot); hipFree(cs); hipFree(ce); hipFree(dn);
  return rc;
}

}  // extern "C"

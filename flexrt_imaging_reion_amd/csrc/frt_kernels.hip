// frt_kernels.hip -- gfx950 kernels of the per-timestep RT core.
//
// Data layout in HBM
//   rays    : four 16-byte vectors per ray, each in its own array (see
//             RaySoA); after every compaction the list is stably sorted by
//             the 4x4x4 tile of each ray's cell (frt_order.h), split
//             children in their parent's place; new casts append in source
//             order;
//   paths   : path_cell/path_dist [path_length][capacity], written once per
//             step by the march and re-read by every C2Ray pass;
//   segments: stable-sorted by cell once per step; every pass writes each
//             segment's deposit into its cell-sorted slot and one thread per
//             cell sums its run in ray order (no float atomics, so the sums
//             do not depend on scheduling);
//   cells   : one float array per reference field, [nfreq][n^3] for per-bin
//             fields, cells numbered tile by tile (frt_order.h) so that the
//             cells a ray list tile touches are contiguous.
// Every floating-point expression keeps the reference's float/double
// promotions (noted per line); the library is built with -ffp-contract=off.
#include <hipcub/hipcub.hpp>

#include "frt_internal.h"
#include "frt_fexp.h"
#include "frt_order.h"
#include "frt_rng.h"

namespace frt {

static constexpr int BLK = 256;

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

// Periodic index one step from a valid one (a in [-1, b]): the reference's
// mod (general_funcs.cc:189-197) for that range, without an integer division.
__device__ __forceinline__ int wrap_mod(int a, int b) {
  return a < 0 ? a + b : (a >= b ? a - b : a);
}

// Grid-stride launches of the light per-ray kernels are capped at this many
// blocks.  The march, deposit and compaction kernels (long, latency-bound
// threads) launch one thread per ray instead: the dispatcher then refills
// the CUs with fresh wavefronts to the end, where a capped grid leaves a
// partial last round of long-running wavefronts.
static constexpr unsigned MAX_GRID = 4096;
static inline unsigned sgrid(int64_t n) {
  const int64_t g = (n + BLK - 1) / BLK;
  return (unsigned)(g < (int64_t)MAX_GRID ? g : (int64_t)MAX_GRID);
}

// Block index remapped so that each XCD runs a contiguous eighth of the
// grid.  Workgroups are dealt to the 8 XCDs round-robin in block order, so
// without the remap neighbouring blocks, which handle neighbouring rays and
// write neighbouring slots, sit in different L2s, and each L2 writes back
// partial lines of the same region.  A permutation of [0, gridDim.x), so a
// grid-stride loop over it still visits every index once.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned G = gridDim.x, b = blockIdx.x;
  const unsigned G8 = G & ~7u;
  if (b >= G8) return b;
  return (int64_t)(b & 7u) * (G8 >> 3) + (b >> 3);
}

// Add v summed over the block to counter *ctr (a field of DevCounters copy 0)
// with one atomic on the block's copy.  Every thread of the block must reach
// the call.
__device__ __forceinline__ void block_add(unsigned long long* ctr, unsigned long long v) {
  __shared__ unsigned long long part[BLK / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < BLK / 64; ++k) t += part[k];
    constexpr unsigned stride = sizeof(DevCounters) / sizeof(unsigned long long);
    static_assert(sizeof(DevCounters) % sizeof(unsigned long long) == 0, "whole-word counter copies");
    if (t) atomicAdd(ctr + (blockIdx.x % FRT_CTR_COPIES) * stride, t);
  }
  __syncthreads();
}

// Work distribution over n items.  Without a queue head: grid-stride over
// the XCD-remapped blocks.  With one (reset to 0 before the launch): every
// wavefront pulls `chunk` consecutive items at a time from *head until they
// run out, so a persistent grid of long-running wavefronts ends within one
// chunk of each other instead of on a fixed share.  The order of the calls
// changes no result: each item writes only its own outputs.
template <typename Body>
__device__ __forceinline__ void for_items(int64_t n, int64_t* head, int chunk, Body&& body) {
  if (!head) {
    for (int64_t i = xcd_block() * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) body(i);
    return;
  }
  const int lane = threadIdx.x & 63;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long*)head, (unsigned long long)chunk);
    base = __shfl(base, 0, 64);
    if ((int64_t)base >= n) break;
    const int64_t end = min((int64_t)base + chunk, n);
    for (int64_t i = (int64_t)base + lane; i < end; i += 64) body(i);
  }
}

// Ray order of the deposit kernels.  A wavefront runs as long as its longest
// path.  The FP64 deposit is bound by FP64 issue, so each block takes CHUNK
// consecutive rays at a time and reorders them by decreasing segment count
// (LDS counting sort, one bucket per count): the lanes of a wavefront then
// walk paths of (nearly) the same length, and rays without segments are
// skipped.  The mixed deposit keeps the list order (for_rays), where
// neighbouring lanes read neighbouring path entries; spreading a wavefront's
// rays over a chunk costs it more in memory requests than it saves in idle
// lanes.  The order changes no result: each ray writes only its own slots,
// and the counters are sums.
static constexpr int DEP_BUCKETS = 32;  // counts >= 31 share the first bucket
static inline unsigned chunk_grid(int64_t n, int chunk) {
  const int64_t g = (n + chunk - 1) / chunk;
  return (unsigned)(g < (int64_t)MAX_GRID ? g : (int64_t)MAX_GRID);
}

// Calls body(i, seg_cnt[i]) for every moved ray i = moved[t], t < nmoved,
// in list order (grid-stride).  The moved rays (seg_cnt > 0) are listed once
// per step, so no lane waits on a ray that did not move.
template <typename Body>
__device__ __forceinline__ void for_rays(const int32_t* __restrict__ seg_cnt, const int32_t* __restrict__ moved,
                                         int64_t nmoved, int64_t* head, int chunk, Body&& body) {
  for_items(nmoved, head, chunk, [&](int64_t t) {
    const int64_t i = moved[t];
    body(i, seg_cnt[i]);
  });
}

// Calls body(i, seg_cnt[i]) for every moved ray i = moved[t], t < nmoved, in
// decreasing segment count within each chunk of the moved list.  Every
// thread of the block must reach the call.
template <int CHUNK, typename Body>
__device__ __forceinline__ void for_rays_by_count(const int32_t* __restrict__ seg_cnt,
                                                  const int32_t* __restrict__ moved, int64_t nrays,
                                                  int64_t* head, Body&& body) {
  static_assert(CHUNK % BLK == 0 && CHUNK <= 32768, "chunk of whole blocks, int16 local indices");
  constexpr int PER = CHUNK / BLK;
  __shared__ int s_cnt[DEP_BUCKETS];
  __shared__ int s_off[DEP_BUCKETS];
  __shared__ int16_t s_idx[CHUNK];
  __shared__ int64_t s_base;
  // chunks dealt round-robin to the blocks, or pulled from the queue head
  auto next = [&](int64_t prev) -> int64_t {
    if (!head) return prev < 0 ? (int64_t)blockIdx.x * CHUNK : prev + (int64_t)gridDim.x * CHUNK;
    if (threadIdx.x == 0) s_base = (int64_t)atomicAdd((unsigned long long*)head, (unsigned long long)CHUNK);
    __syncthreads();
    const int64_t b = s_base;
    __syncthreads();
    return b;
  };
  for (int64_t base = next(-1); base < nrays; base = next(base)) {
    if (threadIdx.x < DEP_BUCKETS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    int bkt[PER], rank[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int64_t t = base + k * BLK + threadIdx.x;
      const int c = t < nrays ? seg_cnt[moved[t]] : 0;
      bkt[k] = c <= 0 ? DEP_BUCKETS - 1 : (c >= DEP_BUCKETS - 1 ? 0 : DEP_BUCKETS - 1 - c);
      rank[k] = atomicAdd(&s_cnt[bkt[k]], 1);
    }
    __syncthreads();
    if (threadIdx.x < DEP_BUCKETS) {  // exclusive scan of the bucket sizes
      const int own = s_cnt[threadIdx.x];
      int v = own;
#pragma unroll
      for (int off = 1; off < DEP_BUCKETS; off <<= 1) {
        const int t = __shfl_up(v, off, DEP_BUCKETS);
        if ((int)threadIdx.x >= off) v += t;
      }
      s_off[threadIdx.x] = v - own;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) s_idx[s_off[bkt[k]] + rank[k]] = (int16_t)(k * BLK + threadIdx.x);
    __syncthreads();
    const int nact = s_off[DEP_BUCKETS - 1];  // rays with segments come first
#pragma unroll 1
    for (int k = 0; k < PER; ++k) {
      const int pos = k * BLK + threadIdx.x;
      if (pos < nact) {
        const int64_t i = moved[base + s_idx[pos]];
        body(i, seg_cnt[i]);
      }
    }
    __syncthreads();  // the LDS arrays are reused by the next chunk
  }
}

// v' = M v in float with the reference's left-to-right sums (healpix_funcs.cc:112-133).
__device__ __forceinline__ void rotate3(const float* m, float x, float y, float z, float* ox,
                                        float* oy, float* oz) {
  *ox = m[0] * x + m[1] * y + m[2] * z;
  *oy = m[3] * x + m[4] * y + m[5] * z;
  *oz = m[6] * x + m[7] * y + m[8] * z;
}

// Loop over the frequency bins, unrolled to the compile-time maximum so that
// per-bin register arrays are indexed statically (no scratch).
#define FRT_FOR_BINS(nu) \
  _Pragma("unroll") for (int nu = 0; nu < FRT_MAX_FREQ; ++nu) if (nu < F)

static_assert(FRT_MAX_FREQ <= 8, "two 4-bin photon groups at most");

// One ray's geometry and identity in registers (pos, dir, meta vectors).
struct Ray {
  float x, y, z, time, xhat, yhat, zhat, r, dr;
  int cell, pixel, order, itheta;
};

__device__ __forceinline__ void ray_load(const RaySoA& b, int64_t i, Ray& r) {
  const float4 p = b.pos[i];
  const float4 d = b.dir[i];
  const int4 m = b.meta[i];
  r.x = p.x;
  r.y = p.y;
  r.z = p.z;
  r.time = p.w;
  r.xhat = d.x;
  r.yhat = d.y;
  r.zhat = d.z;
  r.r = d.w;
  r.dr = __int_as_float(m.x);
  r.cell = m.y;
  r.pixel = m.z;
  r.order = m.w & 0xffff;
  r.itheta = (int)((uint32_t)m.w >> 16);
}

__device__ __forceinline__ void ray_store(const RaySoA& b, int64_t i, const Ray& r) {
  b.pos[i] = make_float4(r.x, r.y, r.z, r.time);
  b.dir[i] = make_float4(r.xhat, r.yhat, r.zhat, r.r);
  b.meta[i] = make_int4(__float_as_int(r.dr), r.cell, r.pixel, (r.order & 0xffff) | (r.itheta << 16));
}

__device__ __forceinline__ int ray_cell(const RaySoA& b, int64_t i) { return b.meta[i].y; }
__device__ __forceinline__ void ray_set_cell(const RaySoA& b, int64_t i, int cell) {
  reinterpret_cast<int*>(b.meta + i)[1] = cell;
}

// Photon counts of ray i in n[0..7] (bins >= F are zero).  Per-bin loads
// (a 16-byte vector load into the predicated per-bin array is a code form
// avoided here).
__device__ __forceinline__ void phot_load(const RaySoA& b, int64_t i, int F, int64_t cap, float* n) {
  const float* p0 = reinterpret_cast<const float*>(b.phot + i);
  const float* hi = reinterpret_cast<const float*>(b.phot + cap) + i;
  FRT_FOR_BINS(nu) n[nu] = (nu < 4) ? p0[nu] : hi[(int64_t)(nu - 4) * cap];
  for (int nu = F; nu < 8; ++nu) n[nu] = 0.f;
}

__device__ __forceinline__ void phot_store(const RaySoA& b, int64_t i, int F, int64_t cap, const float* n) {
  b.phot[i] = make_float4(n[0], F > 1 ? n[1] : 0.f, F > 2 ? n[2] : 0.f, F > 3 ? n[3] : 0.f);
  float* hi = reinterpret_cast<float*>(b.phot + cap) + i;
  for (int nu = 4; nu < F; ++nu) hi[(int64_t)(nu - 4) * cap] = n[nu];
}

// sum_nphot (io_funcs.cc:10-16): float sum over bins in bin order.
__device__ __forceinline__ float phot_sum(const float* n, int F) {
  float s = 0.f;
  FRT_FOR_BINS(nu) s += n[nu];
  return s;
}

__device__ __forceinline__ float sum_nphot(const RaySoA& b, int64_t i, int F, int64_t cap) {
  float n[8];
  phot_load(b, i, F, cap, n);
  return phot_sum(n, F);
}

// Corner correction after a crossing (ray_tracing.cc:1400-1417).
__device__ __forceinline__ float corner(float v, float dx, float min_dist) {
  if (v <= 0.f) v = min_dist * dx;                                      // float*float
  if ((double)v >= (double)dx * (1. - (double)min_dist / 2.))          // double compare
    v = (float)((double)dx * (1. - (double)min_dist));
  return v;
}

// Crossing time to the cell face along one axis (ray_tracing.cc:1274-1302).
__device__ __forceinline__ float face_time(float hat, float pos, float dx, float clgt) {
  if (hat > 0.f) return (dx - pos) / hat / clgt;                         // float
  if (hat < 0.f) return (float)(((-1. * (double)pos) / (double)hat) / (double)clgt);  // -1.* is double
  return INFINITY;                                                       // 9e99 stored in float
}

// Geometry of child ch (0..3) of split parent i (split_ray, ray_tracing.cc:482-619):
// pixel 4p+ch at order l+1, direction rotated like the parent, position moved by
// (v_child - v_parent) r and walked into the containing periodic cell.
struct ChildGeo {
  int cell;
  int pix;
  float vx, vy, vz;
  float x, y, z;
};

__device__ __forceinline__ ChildGeo child_geometry(const Ray& par, int ch, const GridConst& g,
                                                   const float* __restrict__ rot_fwd) {
  const int l = par.order;
  const int pn = par.pixel;
  const float r0 = par.r;
  const float* m = rot_fwd + 9 * (int64_t)par.itheta;
  const double vox = (double)par.xhat, voy = (double)par.yhat, voz = (double)par.zhat;
  const int n = g.n;
  int ic, jc, kc;
  frt_cell_ijk(g.lay, par.cell, ic, jc, kc);
  const float rsf = r0 * g.sf;
  ChildGeo c;
  c.pix = 4 * pn + ch;
  double px, py, pz;
  hpx_pix2vec_nest(l + 1, c.pix, &px, &py, &pz);
  rotate3(m, (float)px, (float)py, (float)pz, &c.vx, &c.vy, &c.vz);
  float xn = (float)((double)par.x + ((double)c.vx - vox) * (double)rsf);
  float yn = (float)((double)par.y + ((double)c.vy - voy) * (double)rsf);
  float zn = (float)((double)par.z + ((double)c.vz - voz) * (double)rsf);
  int icn = ic, jcn = jc, kcn = kc;
  // periodic cell walk (ray_tracing.cc:549-572); the guard only bounds non-finite input
  int w = 0;
  while (xn > g.dx && w < 4096) { xn -= g.dx; icn = wrap_mod(icn + 1, n); ++w; }
  while ((double)xn < 0. && w < 4096) { xn += g.dx; icn = wrap_mod(icn - 1, n); ++w; }
  while (yn > g.dx && w < 4096) { yn -= g.dx; jcn = wrap_mod(jcn + 1, n); ++w; }
  while ((double)yn < 0. && w < 4096) { yn += g.dx; jcn = wrap_mod(jcn - 1, n); ++w; }
  while (zn > g.dx && w < 4096) { zn -= g.dx; kcn = wrap_mod(kcn + 1, n); ++w; }
  while ((double)zn < 0. && w < 4096) { zn += g.dx; kcn = wrap_mod(kcn - 1, n); ++w; }
  c.cell = frt_cell_index(g.lay, icn, jcn, kcn);
  c.x = xn;
  c.y = yn;
  c.z = zn;
  return c;
}

// The ray at position q of the compacted list after a split: ray i copied,
// or child ch of split ray i (split_ray, ray_tracing.cc:482-619), with its
// photons in n (a quarter of the parent's per bin); sorted_src[q] = i << 2 | ch
// (the values the split's tile sort carried).
__device__ __forceinline__ void output_ray(const RaySoA& src, int64_t q, const GridConst& g,
                                           const int32_t* __restrict__ n_out,
                                           const uint32_t* __restrict__ sorted_src,
                                           const float* __restrict__ rot_fwd, Ray& out, float* n) {
  const uint32_t packed = sorted_src[q];
  const int64_t i = (int64_t)(packed >> 2);
  const int ch = (int)(packed & 3u);
  ray_load(src, i, out);
  phot_load(src, i, g.nfreq, g.cap, n);
  if (n_out[i] == 1) return;
  const Ray par = out;
  const int l = par.order;
  const ChildGeo c = child_geometry(par, ch, g, rot_fwd);
  out.cell = c.cell;
  out.pixel = c.pix;
  out.order = l + 1;
  out.xhat = c.vx;
  out.yhat = c.vy;
  out.zhat = c.vz;
  out.dr = g.rmax_tab[l + 1] - par.r;
  out.x = c.x;
  out.y = c.y;
  out.z = c.z;
  const int F = g.nfreq;
  FRT_FOR_BINS(nu) n[nu] = (float)((double)n[nu] / 4.);
}

// ---------------------------------------------------------------------------
// advance_ray (ray_tracing.cc:1222-1512), C2Ray mode, fion = 1.
// With GATHER, the list is still the previous split's permutation of the
// other buffer (src): the first sd->ncompact rays are fetched through it
// (output_ray, the compaction's gather) and written whole to rb after the
// march, so the compaction and the march share one pass over the rays; the
// rays after them (new_rays's) are in rb already.
template <bool GATHER>
__global__ void __launch_bounds__(BLK) k_march(RaySoA rb, StepDev* sd, GridConst g,
                                               int32_t* __restrict__ path_cell,
                                               float* __restrict__ path_dist,
                                               int32_t* __restrict__ seg_cnt,
                                               uint8_t* __restrict__ alive,
                                               float* __restrict__ psum,
                                               DevCounters* cnt, int chunk, RaySoA src,
                                               const int32_t* __restrict__ n_out,
                                               const uint32_t* __restrict__ sorted_src,
                                               const float* __restrict__ rot_fwd,
                                               const float* __restrict__ fion) {
  unsigned long long nmoved = 0;
  bool overflow = false;
  const int64_t nrays = sd->nrays;
  const int64_t ngather = GATHER ? sd->ncompact : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) seg_cnt[nrays] = 0;  // scanned with the rays
  for_items(nrays, chunk > 0 ? &sd->queue[FRT_Q_MARCH] : nullptr, chunk, [&](int64_t i) {
    bool moved = false;
    alive[i] = 1;
    const int n = g.n;
    Ray ry;
    float nph[8];
    const bool whole = GATHER && i < ngather;  // write every field of the ray
    if (whole)
      output_ray(src, i, g, n_out, sorted_src, rot_fwd, ry, nph);
    else
      ray_load(rb, i, ry);
    int ci, cj, ck;
    frt_cell_ijk(g.lay, ry.cell, ci, cj, ck);
    const float xhat = ry.xhat, yhat = ry.yhat, zhat = ry.zhat;
    float x0 = ry.x, y0 = ry.y, z0 = ry.z;
    const float deltat = ry.time + g.dt;  // ray[x].time += dt
    float dist = g.clight * deltat;
    const float hi = g.nc * g.dx;                             // nc*dx[i] (float)
    const double lo = (1. - (double)g.nc) * (double)g.dx;     // (1. - nc)*dx[i] (double)
    const float ax = x0 + dist * xhat, ay = y0 + dist * yhat, az = z0 + dist * zhat;
    moved = (ax > hi) || (ay > hi) || (az > hi) || ((double)ax < lo) || ((double)ay < lo) ||
            ((double)az < lo);
    if (!moved) {
      // photon sum of a ray the deposit passes do not visit
      if (whole) {
        ry.time = deltat;
        ray_store(rb, i, ry);
        phot_store(rb, i, g.nfreq, g.cap, nph);
        psum[i] = phot_sum(nph, g.nfreq);
      } else {
        rb.pos[i] = make_float4(ry.x, ry.y, ry.z, deltat);
        psum[i] = sum_nphot(rb, i, g.nfreq, g.cap);
      }
      seg_cnt[i] = 0;
    } else {
      float r0 = ry.r;
      float dr = ry.dr;
      r0 += dist / g.sf;  // comoving radius
      dr -= dist / g.sf;
      int inew = ci, jnew = cj, knew = ck;
      int p = 0;
      float tlast = 0.f, tmin = 0.f;
      int guard = 0;
      while (tlast < deltat) {
        const float clgt = g.clight;
        const float tx = face_time(xhat, x0, g.dx, clgt);
        const float ty = face_time(yhat, y0, g.dx, clgt);
        const float tz = face_time(zhat, z0, g.dx, clgt);
        int tag;
        if ((tx <= ty) && (tx <= tz)) {
          tmin = tx;
          tag = 0;
        } else if ((ty <= tx) && (ty <= tz)) {
          tmin = ty;
          tag = 1;
        } else {
          tmin = tz;
          tag = 2;
        }
        if ((tlast + tmin) < deltat) {
          dist = tmin * clgt;
          if (p < g.P) {
            path_cell[(int64_t)p * g.cap + i] = frt_cell_index(g.lay, inew, jnew, knew);
            path_dist[(int64_t)p * g.cap + i] = dist;
          } else {
            overflow = true;
          }
          p += 1;
          x0 += dist * xhat;
          y0 += dist * yhat;
          z0 += dist * zhat;
          const float snap_lo = g.min_dist * g.dx;
          const float snap_hi = (float)((double)g.dx * (1. - (double)g.min_dist));
          if (tag == 0) {
            if (xhat > 0.f) {
              inew = wrap_mod(inew + 1, n);
              x0 = snap_lo;
            } else {
              inew = wrap_mod(inew - 1, n);
              x0 = snap_hi;
            }
          } else if (tag == 1) {
            if (yhat > 0.f) {
              jnew = wrap_mod(jnew + 1, n);
              y0 = snap_lo;
            } else {
              jnew = wrap_mod(jnew - 1, n);
              y0 = snap_hi;
            }
          } else {
            if (zhat > 0.f) {
              knew = wrap_mod(knew + 1, n);
              z0 = snap_lo;
            } else {
              knew = wrap_mod(knew - 1, n);
              z0 = snap_hi;
            }
          }
          x0 = corner(x0, g.dx, g.min_dist);
          y0 = corner(y0, g.dx, g.min_dist);
          z0 = corner(z0, g.dx, g.min_dist);
        }
        tlast += tmin;
        if (++guard > g.P + 64) {  // cannot happen for finite input; keeps the loop bounded
          overflow = true;
          break;
        }
      }
      // last segment: remaining time in the final cell (:1434-1437), or with
      // the I-front model in a partly ionized cell its whole crossing (:1430-1433)
      if (fion && fion[frt_cell_index(g.lay, inew, jnew, knew)] < 1.f)
        dist = tmin * g.clight;
      else
        dist = (deltat - tlast + tmin) * g.clight;
      dist = frt_maxd(dist, g.min_dist * g.dx);
      if (p < g.P) {
        path_cell[(int64_t)p * g.cap + i] = frt_cell_index(g.lay, inew, jnew, knew);
        path_dist[(int64_t)p * g.cap + i] = dist;
      } else {
        overflow = true;
      }
      seg_cnt[i] = (p + 1 < g.P) ? p + 1 : g.P;
      x0 += dist * xhat;
      y0 += dist * yhat;
      z0 += dist * zhat;
      x0 = corner(x0, g.dx, g.min_dist);
      y0 = corner(y0, g.dx, g.min_dist);
      z0 = corner(z0, g.dx, g.min_dist);
      ry.r = r0;
      ry.dr = dr;
      ry.x = x0;
      ry.y = y0;
      ry.z = z0;
      ry.time = 0.f;
      ray_store(rb, i, ry);
      if (whole) phot_store(rb, i, g.nfreq, g.cap, nph);
    }
    nmoved += moved ? 1 : 0;
  });
  block_add(&cnt->moved, nmoved);
  if (overflow) atomicOr(&cnt->path_overflow, 1u);
}

// Segment keys for the per-step cell sort.
__global__ void __launch_bounds__(BLK) k_seg_keys(const StepDev* __restrict__ sd, int64_t cap,
                                                  const int32_t* __restrict__ seg_cnt,
                                                  const int32_t* __restrict__ seg_off,
                                                  const int32_t* __restrict__ path_cell,
                                                  uint32_t* __restrict__ key,
                                                  int32_t* __restrict__ val) {
  const int64_t nrays = sd->nrays;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < nrays; i += (int64_t)gridDim.x * BLK) {
    const int c = seg_cnt[i];
    const int base = seg_off[i];
    for (int p = 0; p < c; ++p) {
      key[base + p] = (uint32_t)path_cell[(int64_t)p * cap + i];
      val[base + p] = (int32_t)((int64_t)p * cap + i);  // path-array index (< cap * path_length < 2^31)
    }
  }
}

// {n_HI, n_HeI} packed for the first deposit pass of a step.
__global__ void __launch_bounds__(BLK) k_pack_avg(const float* __restrict__ a, const float* __restrict__ b,
                                                  float2* __restrict__ out, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k < n) out[k] = make_float2(a[k], b[k]);
}

// Adjoint tape: a pass's opacities (or the step's starting fractions) as
// doubles, and the rays' photons at the start of the step.
__global__ void __launch_bounds__(BLK) k_tape_pair(const float* __restrict__ a, const float* __restrict__ b,
                                                   double* __restrict__ out, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k < n) {
    out[k] = (double)a[k];
    out[n + k] = (double)b[k];
  }
}

// Fill of the first *dn + add bytes (device count) with v.
__global__ void __launch_bounds__(BLK) k_fill_u8_dev(uint8_t* a, const int64_t* __restrict__ dn, int64_t add, uint8_t v) {
  const int64_t n = *dn + add;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) a[i] = v;
}



// ---------------------------------------------------------------------------
// find_avg_unu ray loop (ray_tracing.cc:701-938), one thread per moved ray.
//
// Each segment's deposit goes to its cell-sorted slot (so a cell's slots are
// contiguous and in ray order).  In the first passes the slot holds the
// {H, He} deposit summed over bins.  In the last pass it holds a photon
// record: the photons entering the segment in every bin and the segment
// length, negated on the segment where the ray is dumped (+inf behind a
// dump).  The cell kernel recomputes the per-bin deposits from the record
// and its own opacities (the same expressions, so the same bits), which
// keeps the per-bin Gamma and photon fields without storing 3 values per bin.

// Photon record of a segment: nin[0..F-1], the signed length at F, zeros up
// to the record length (a multiple of 4 values), written as 16-byte stores.
// (A record of F + 1 values rounded to even, 24 instead of 32 bytes at five
// bins, measured no faster: the cell kernel's walk then takes three 8-byte
// loads per record instead of two 16-byte ones, and each load of a wavefront
// touches one cache line per lane.)
#define FRT_MAX_REC 12  // record values at most: FRT_MAX_FREQ + 1 rounded up to 4
static_assert(FRT_MAX_FREQ + 1 <= FRT_MAX_REC, "record too short");

template <typename T>
__device__ __forceinline__ T rec_value(const T* nin, int F, T last, int k) {
  if (k < FRT_MAX_FREQ && k < F) return nin[k < FRT_MAX_FREQ ? k : 0];
  return (k == F) ? last : (T)0;
}

__device__ __forceinline__ void store_record_values(float* rec, const float* nin, int F, float last, int rec_len) {
#pragma unroll
  for (int k = 0; k < FRT_MAX_REC; k += 4) {
    if (k >= rec_len) break;
    *(float4*)(rec + k) = make_float4(rec_value(nin, F, last, k), rec_value(nin, F, last, k + 1),
                                      rec_value(nin, F, last, k + 2), rec_value(nin, F, last, k + 3));
  }
}

__device__ __forceinline__ void store_record_values(double* rec, const double* nin, int F, double last,
                                                    int rec_len) {
#pragma unroll
  for (int k = 0; k < FRT_MAX_REC; k += 2) {
    if (k >= rec_len) break;
    *(double2*)(rec + k) = make_double2(rec_value(nin, F, last, k), rec_value(nin, F, last, k + 1));
  }
}

template <typename T>
__device__ __forceinline__ void store_photon_record(T* rec, const T* nin, int F, float dist, bool dump,
                                                    int rec_len) {
  store_record_values(rec, nin, F, (T)(dump ? -dist : dist), rec_len);
}

template <typename T>
__device__ __forceinline__ void store_behind_dump(T* rec, int F, int rec_len) {
  const T zero[FRT_MAX_FREQ] = {};
  store_record_values(rec, zero, F, (T)INFINITY, rec_len);
}

// Photon record as registers: q[0..F-1] the photons in, d the signed length
// (16-byte loads; F + 1 <= rec_len <= FRT_MAX_REC).
__device__ __forceinline__ void load_photon_record(const float* rec, int F, int rec_len, float* q, float& d) {
  float v[FRT_MAX_REC];
#pragma unroll
  for (int k = 0; k < FRT_MAX_REC; k += 4) {
    float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < rec_len) t = *(const float4*)(rec + k);
    v[k] = t.x;
    v[k + 1] = t.y;
    v[k + 2] = t.z;
    v[k + 3] = t.w;
  }
  FRT_FOR_BINS(nu) q[nu] = v[nu];
  d = 0.f;
#pragma unroll
  for (int k = 1; k < FRT_MAX_REC; ++k)
    if (k == F) d = v[k];
}

// Per-bin deposit of one segment, FP64 form (ray_tracing.cc:764-840):
// HI/HeI split, absorbed photons, Gamma contributions.
__device__ __forceinline__ void bin_deposit_fp64(const GridConst& g, double dth, double dthe, double nin,
                                                 bool dump, float ah1, float ahe1, double& gh, double& ghe) {
  const double dtd = (double)g.dt;
  double fh = 0., fhe = 0., att = 0.;
  const bool any = (dth > 0.) || (dthe > 0.);
  double eh = 1., ehe = 1.;
  if (any || !dump) {
    eh = dm_exp(-dth);
    ehe = (dthe == 0.) ? 1. : dm_exp(-dthe);  // exp(-0) = 1 exactly
  }
  if (any) {
    const double a = (1. - eh) * ehe;
    const double b = (1. - ehe) * eh;
    fh = a / (a + b);
    fhe = b / (a + b);
    if (g.frac_limit && (a + b) == 0.) {
      fh = 1. / (1. + dm_exp(dthe - dth));
      fhe = 1. / (1. + dm_exp(dth - dthe));
    }
  }
  if (dump) {
    gh = fh * nin / dtd / g.vol / (double)ah1;
    ghe = fhe * nin / dtd / g.vol / (double)ahe1;
  } else {
    // exp(-dth - dthe): with dthe = 0 the argument is exactly -dth
    att = 1. - ((dthe == 0.) ? eh : dm_exp(-dth - dthe));
    gh = fh * nin * att / dtd / g.vol / (double)ah1;
    ghe = fhe * nin * att / dtd / g.vol / (double)ahe1;
  }
}

// Transmissions eh, ehe of one segment and bin, mixed form, with the
// absorbed fractions omh = 1 - eh, omhe = 1 - ehe from the same reduction
// (through expm1: in ionized cells dtau is ~1e-5 and 1 - exp(-dtau) would
// keep only ~3 significant digits).
__device__ __forceinline__ void bin_atten_mixed(float dth, float dthe, float& eh, float& ehe, float& omh,
                                                float& omhe) {
  float mh, mhe = 0.f;
  ehe = 1.f;
  frt_exp_expm1_neg(dth, eh, mh);
  if (dthe != 0.f) frt_exp_expm1_neg(dthe, ehe, mhe);
  omh = -mh;
  omhe = -mhe;
}

// Gamma contributions gh, ghe (in units of phot_norm photons) of one segment
// and bin from its transmissions (bin_atten_mixed), mixed form.
__device__ __forceinline__ void bin_share_mixed(const GridConst& g, float dth, float dthe, float eh, float ehe,
                                                float omh, float omhe, float nin, bool dump, float rh1,
                                                float rhe1, float& gh, float& ghe) {
  float fh = 0.f, fhe = 0.f;
  if ((dth > 0.f) || (dthe > 0.f)) {
    // HI share fh = a/(a+b), a = (1 - eh) ehe, b = (1 - ehe) eh
    // (ray_tracing.cc:770-771), with a and b divided by eh ehe:
    //   fh = omh / (omh + omhe e^-(dth - dthe))                for dth >= dthe,
    //   fh = omh e^-(dthe - dth) / (omh e^-(dthe - dth) + omhe) otherwise,
    // exact where eh or ehe leaves the float range (a float eh is 0 beyond a
    // depth of 87, while the FP64 form's double stays positive to 745).
    // Beyond 745 the FP64 form's exp underflows too, and its limits are kept:
    // fh = 1 if only dth is that deep, 0 if only dthe, the logistic
    // 1/(1 + e^(dthe - dth)) of frac_limit if both.
    constexpr float DEEP = 745.13322f;  // double exp(-x) is 0 beyond this
    const bool dh = dth > DEEP, dhe = dthe > DEEP;
    if (dh && dhe) {
      if (g.frac_limit) {
        fh = 1.f / (1.f + expf(dthe - dth));
        fhe = 1.f / (1.f + expf(dth - dthe));
      } else {
        fh = fhe = __int_as_float(0x7fc00000);
      }
    } else if (dh || dhe) {
      fh = dh ? 1.f : 0.f;
      fhe = dh ? 0.f : 1.f;
    } else {
      // shallower than 64 both exponentials and the products are normal
      // floats and a = omh ehe, b = omhe eh directly; deeper, the ratio form
      float a, b;
      if (fmaxf(dth, dthe) < 64.f) {
        a = omh * ehe;
        b = omhe * eh;
      } else {
        float w, unused;
        frt_exp_expm1_neg(fabsf(dth - dthe), w, unused);  // e^-|dth - dthe|
        a = dth >= dthe ? omh : omh * w;
        b = dth >= dthe ? omhe * w : omhe;
      }
      const float den = a + b;
      if (den > 0.f) {
        const float inv = __builtin_amdgcn_rcpf(den);  // 1 ulp
        fh = a * inv;
        fhe = b * inv;
      } else {  // depths so small that 1 - e^-d rounds to 0 in both
        fh = dth >= dthe ? 1.f : 0.f;
        fhe = 1.f - fh;
      }
    }
  }
  // 1 - exp(-(dth + dthe)) = omh + omhe - omh * omhe (exact identity)
  const float eabs = dump ? nin : nin * fmaf(-omh, omhe, omh + omhe);
  gh = fh * eabs * rh1;
  ghe = fhe * eabs * rhe1;
}

// Per-bin deposit of one segment, mixed form: transmissions eh, ehe and the
// Gamma contributions gh, ghe (in units of phot_norm photons).
__device__ __forceinline__ void bin_deposit_mixed(const GridConst& g, float dth, float dthe, float nin, bool dump,
                                                  float rh1, float rhe1, float& gh, float& ghe, float& eh,
                                                  float& ehe) {
  float omh, omhe;
  bin_atten_mixed(dth, dthe, eh, ehe, omh, omhe);
  bin_share_mixed(g, dth, dthe, eh, ehe, omh, omhe, nin, dump, rh1, rhe1, gh, ghe);
}

// Software pipeline over a ray's segments: the path entry (cell, length,
// slot) of segment p+2 and the cell opacities of segment p+1 are requested
// before segment p is computed, so the dependent loads overlap the math.
struct SegPipe {
  int64_t i, stride;
  int c;
  const int32_t* path_cell;
  const float* path_dist;
  const int32_t* path_slot;
  const float2* avg2;
  // stage registers: segment p (ready), p+1 (path + opacity), p+2 (path)
  int cell0, cell1, cell2, slot0, slot1, slot2;
  float dist0, dist1, dist2, ah0, ahe0, ah1, ahe1;

  __device__ __forceinline__ void load_path(int p, int& cell, float& dist, int& slot) const {
    cell = path_cell[(int64_t)p * stride + i];
    dist = path_dist[(int64_t)p * stride + i];
    slot = path_slot[(int64_t)p * stride + i];
  }
  __device__ __forceinline__ void start() {
    load_path(0, cell0, dist0, slot0);
    if (c > 1) load_path(1, cell1, dist1, slot1);
    const float2 a = avg2[cell0];
    ah0 = a.x;
    ahe0 = a.y;
  }
  // issue the loads of the next stages before segment p is computed
  __device__ __forceinline__ void prefetch(int p) {
    if (p + 2 < c) load_path(p + 2, cell2, dist2, slot2);
    if (p + 1 < c) {
      const float2 a = avg2[cell1];
      ah1 = a.x;
      ahe1 = a.y;
    }
  }
  __device__ __forceinline__ void advance() {
    cell0 = cell1; dist0 = dist1; slot0 = slot1; ah0 = ah1; ahe0 = ahe1;
    cell1 = cell2; dist1 = dist2; slot1 = slot2;
  }
};

// One segment of the FP64 deposit; returns true if the ray is dumped here.
// The last pass stores the segment's photon record instead of its deposit.
template <int NFC>
__device__ __forceinline__ bool deposit_segment_fp64(const GridConst& g, ContribPtrs& cp, bool last, int slot,
                                                     float dist, float ah1, float ahe1, const double* nph,
                                                     double* tauh, double* tauhe) {
  const int F = NFC > 0 ? NFC : g.nfreq;
  const double distd = (double)dist;
  double dth[FRT_MAX_FREQ], dthe[FRT_MAX_FREQ], nin[FRT_MAX_FREQ];
  double tot = 0.;
  FRT_FOR_BINS(nu) {
    dth[nu] = ((double)ah1 * g.sigH[nu]) * distd;  // avg_nh1*sigma*dist
    dthe[nu] = ((double)ahe1 * g.sigHe[nu]) * distd;
    nin[nu] = nph[nu] * dm_exp(-tauh[nu] - tauhe[nu]);
    tot += nin[nu] / FRT_PHOT_NORM;
  }
  const bool dump = tot < g.nphot_min;
  if (last) {
    store_photon_record((double*)cp.rec + (int64_t)slot * cp.rec_len, nin, F, dist, dump, cp.rec_len);
  } else {
    double gh_sum = 0., ghe_sum = 0.;
    FRT_FOR_BINS(nu) {
      double gh, ghe;
      bin_deposit_fp64(g, dth[nu], dthe[nu], nin[nu], dump, ah1, ahe1, gh, ghe);
      gh_sum += gh;
      ghe_sum += ghe;
    }
    ((double2*)cp.pair)[slot] = make_double2(gh_sum, ghe_sum);
  }
  FRT_FOR_BINS(nu) {
    tauh[nu] += dth[nu];
    tauhe[nu] += dthe[nu];
  }
  return dump;
}

// FP64 form: the reference's double locals and libm exp (through detmath),
// with the per-cell accumulation in double.
template <bool LAST, int NFC = 0, int CHUNK = 4096>
__global__ void __launch_bounds__(BLK) k_deposit_fp64(RaySoA rb, const int32_t* __restrict__ moved,
                                                      StepDev* sd, int64_t* head, int chunk, GridConst g,
                                                      const int32_t* __restrict__ path_cell,
                                                      const float* __restrict__ path_dist,
                                                      const int32_t* __restrict__ seg_cnt,
                                                      const int32_t* __restrict__ path_slot,
                                                      const float2* __restrict__ avg2,
                                                      ContribPtrs cp, uint8_t* __restrict__ alive,
                                                      DevCounters* cnt, uint8_t* __restrict__ tape_np,
                                                      float* __restrict__ psum) {
  unsigned long long ndumped = 0, nremoved = 0;
  (void)chunk;
  const int64_t nmoved = sd->nmoved;
  for_rays_by_count<CHUNK>(seg_cnt, moved, nmoved, head, [&](int64_t i, int c) {
    bool dumped_flag = false, removed = false;
    {
      const int F = NFC > 0 ? NFC : g.nfreq;
      double nph[FRT_MAX_FREQ], tauh[FRT_MAX_FREQ], tauhe[FRT_MAX_FREQ];
      float n0[8];
      phot_load(rb, i, F, g.cap, n0);
      FRT_FOR_BINS(nu) {
        nph[nu] = (double)n0[nu] * FRT_PHOT_NORM;
        tauh[nu] = 0.;
        tauhe[nu] = 0.;
      }
      SegPipe sp;
      sp.i = i;
      sp.stride = g.cap;
      sp.c = c;
      sp.path_cell = path_cell;
      sp.path_dist = path_dist;
      sp.path_slot = path_slot;
      sp.avg2 = avg2;
      sp.start();
      bool dumped = false;
      int p = 0;
      int last_cell = sp.cell0;
      for (; p < c; ++p) {
        sp.prefetch(p);
        last_cell = sp.cell0;
        if (deposit_segment_fp64<NFC>(g, cp, LAST, sp.slot0, sp.dist0, sp.ah0, sp.ahe0, nph, tauh, tauhe)) {
          dumped = true;
          ++p;
          break;
        }
        sp.advance();
      }
      // segments behind a dump deposit nothing this pass
      // adjoint tape: segments this pass processed, and whether it dumped
      if (tape_np) tape_np[i] = (uint8_t)((dumped ? 0x80 : 0) | p);
      for (; p < c; ++p) {
        const int slot = path_slot[(int64_t)p * g.cap + i];
        last_cell = path_cell[(int64_t)p * g.cap + i];
        if (LAST)
          store_behind_dump((double*)cp.rec + (int64_t)slot * cp.rec_len, F, cp.rec_len);
        else
          ((double2*)cp.pair)[slot] = make_double2(0., 0.);
      }
      if (LAST) {
        // photon update and removal (ray_tracing.cc:923-935)
        float sum = 0.f;
        float n1[8];
        FRT_FOR_BINS(nu) {
          const float v = (float)(nph[nu] * dm_exp(-tauh[nu] - tauhe[nu]) / FRT_PHOT_NORM);
          n1[nu] = v;
          sum += v;
        }
        phot_store(rb, i, F, g.cap, n1);
        psum[i] = sum;  // = phot_sum(n1): the same float sum in bin order
        dumped_flag = dumped;
        if ((double)sum > g.nphot_min) {
          ray_set_cell(rb, i, last_cell);
        } else {
          alive[i] = 0;
          removed = true;
        }
      }
    }
    ndumped += dumped_flag ? 1 : 0;
    nremoved += removed ? 1 : 0;
  });
  if (LAST) {
    block_add(&cnt->dumped, ndumped);
    block_add(&cnt->terminated, nremoved);
  }
}

// One segment of the mixed deposit; returns true if the ray is dumped here.
// The last pass stores the segment's photon record instead of its deposit.
template <int NFC>
__device__ __forceinline__ bool deposit_segment_mixed(const GridConst& g, ContribPtrs& cp, bool last, int slot,
                                                      float dist, float ah1, float ahe1, const float* nph,
                                                      float* trans, float inv_dtv, float nmin) {
  const int F = NFC > 0 ? NFC : g.nfreq;
  float nin[FRT_MAX_FREQ];
  float tot = 0.f;
  FRT_FOR_BINS(nu) {
    nin[nu] = nph[nu] * trans[nu];
    tot += nin[nu];
  }
  const bool dump = tot < nmin;
  if (last) {
    store_photon_record((float*)cp.rec + (int64_t)slot * cp.rec_len, nin, F, dist, dump, cp.rec_len);
    FRT_FOR_BINS(nu) {
      const float dth = ah1 * g.sigHf[nu] * dist;
      const float dthe = ahe1 * g.sigHef[nu] * dist;
      float eh, ehe = 1.f, m;
      frt_exp_expm1_neg(dth, eh, m);
      if (dthe != 0.f) frt_exp_expm1_neg(dthe, ehe, m);
      trans[nu] *= eh * ehe;
    }
  } else {
    // 1/(dt vol n) with the hardware reciprocal (1 ulp)
    const float rh1 = inv_dtv * __builtin_amdgcn_rcpf(ah1), rhe1 = inv_dtv * __builtin_amdgcn_rcpf(ahe1);
    float gh_sum = 0.f, ghe_sum = 0.f;
    FRT_FOR_BINS(nu) {
      const float dth = ah1 * g.sigHf[nu] * dist;
      const float dthe = ahe1 * g.sigHef[nu] * dist;
      float gh, ghe, eh, ehe;
      bin_deposit_mixed(g, dth, dthe, nin[nu], dump, rh1, rhe1, gh, ghe, eh, ehe);
      gh_sum += gh;
      ghe_sum += ghe;
      trans[nu] *= eh * ehe;
    }
    ((float2*)cp.pair)[slot] = make_float2(gh_sum, ghe_sum);
  }
  return dump;
}

// Mixed form: FP32 attenuation and photon bookkeeping; per-cell sums are
// compensated (Kahan) in the cell kernel.  The first passes are held to 72
// VGPRs (7 wavefronts per SIMD instead of 6, three registers spilled):
// the latency-bound pass gains from the extra wavefront; the last pass,
// which would spill thirty, keeps its registers.
template <bool LAST, int NFC = 0>
__global__ void __launch_bounds__(BLK, LAST ? 1 : 7) k_deposit_mixed(RaySoA rb, const int32_t* __restrict__ moved,
                                                       StepDev* sd, int64_t* head, int chunk, GridConst g,
                                                       const int32_t* __restrict__ path_cell,
                                                       const float* __restrict__ path_dist,
                                                       const int32_t* __restrict__ seg_cnt,
                                                       const int32_t* __restrict__ path_slot,
                                                       const float2* __restrict__ avg2,
                                                       ContribPtrs cp, uint8_t* __restrict__ alive,
                                                       DevCounters* cnt, uint8_t* __restrict__ tape_np,
                                                       float* __restrict__ psum) {
  unsigned long long ndumped = 0, nremoved = 0;
  // 1/(dt vol) in units of phot_norm photons
  const float inv_dtv = (float)(FRT_PHOT_NORM / ((double)g.dt * g.vol));
  const float nmin = (float)g.nphot_min;
  const int64_t nmoved = sd->nmoved;
  for_rays(seg_cnt, moved, nmoved, head, chunk, [&](int64_t i, int c) {
    bool dumped_flag = false, removed = false;
    {
      const int F = NFC > 0 ? NFC : g.nfreq;
      float nph[8], trans[FRT_MAX_FREQ];
      phot_load(rb, i, F, g.cap, nph);  // in units of phot_norm
      FRT_FOR_BINS(nu) trans[nu] = 1.f;  // exp(-tau) carried multiplicatively
      SegPipe sp;
      sp.i = i;
      sp.stride = g.cap;
      sp.c = c;
      sp.path_cell = path_cell;
      sp.path_dist = path_dist;
      sp.path_slot = path_slot;
      sp.avg2 = avg2;
      sp.start();
      bool dumped = false;
      int p = 0;
      int last_cell = sp.cell0;
      for (; p < c; ++p) {
        sp.prefetch(p);
        last_cell = sp.cell0;
        if (deposit_segment_mixed<NFC>(g, cp, LAST, sp.slot0, sp.dist0, sp.ah0, sp.ahe0, nph, trans, inv_dtv, nmin)) {
          dumped = true;
          ++p;
          break;
        }
        sp.advance();
      }
      // segments behind a dump deposit nothing this pass
      if (tape_np) tape_np[i] = (uint8_t)((dumped ? 0x80 : 0) | p);
      for (; p < c; ++p) {
        const int slot = path_slot[(int64_t)p * g.cap + i];
        last_cell = path_cell[(int64_t)p * g.cap + i];
        if (LAST)
          store_behind_dump((float*)cp.rec + (int64_t)slot * cp.rec_len, F, cp.rec_len);
        else
          ((float2*)cp.pair)[slot] = make_float2(0.f, 0.f);
      }
      if (LAST) {
        float sum = 0.f;
        float n1[8];
        FRT_FOR_BINS(nu) {
          const float v = nph[nu] * trans[nu];
          n1[nu] = v;
          sum += v;
        }
        phot_store(rb, i, F, g.cap, n1);
        psum[i] = sum;  // = phot_sum(n1): the same float sum in bin order
        dumped_flag = dumped;
        if ((double)sum > g.nphot_min) {
          ray_set_cell(rb, i, last_cell);
        } else {
          alive[i] = 0;
          removed = true;
        }
      }
    }
    ndumped += dumped_flag ? 1 : 0;
    nremoved += removed ? 1 : 0;
  });
  if (LAST) {
    block_add(&cnt->dumped, ndumped);
    block_add(&cnt->terminated, nremoved);
  }
}


#include "frt_thermal.h"

// find_avg_unu cell loop (ray_tracing.cc:941-1214, C2Ray branch 1077-1136,
// finalize 1139-1210), one thread per cell.  FIRST also performs the
// iteration-0 reset of :655-683; LAST finalizes and, with THERMAL, runs
// update_thermal in the same pass (merge and split touch no cell field).
// NFC > 0 fixes the bin count at compile time (the per-bin register arrays
// then hold NFC values instead of FRT_MAX_FREQ); NFC = 0 reads g.nfreq.
template <typename ACC, bool FIRST, bool LAST, bool THERMAL, int NFC = 0>
__global__ void __launch_bounds__(BLK) k_chem(CellPtrs cf, GridConst g, ContribPtrs cp,
                                              const int32_t* __restrict__ cstart,
                                              const int32_t* __restrict__ cend,
                                              DevCounters* cnt, double* __restrict__ tape_g,
                                              const StepDev* __restrict__ sd) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int nsub = 0;
  if (k < g.n3) {
    const int64_t n3 = g.n3;
    const int s0 = cstart[k], s1 = cend[k];
    const int F = NFC > 0 ? NFC : g.nfreq;
    // opacities this pass's deposit used (ray_tracing.cc:656-657,733-743)
    const float ah1 = FIRST ? cf.f[C_NH1][k] : cf.f[C_AVG_NH1][k];
    const float ahe1 = FIRST ? cf.f[C_NHE1][k] : cf.f[C_AVG_NHE1][k];
    // sums over the cell's segments in cell-sorted (= ray) order
    double GH, GHe;
    double GHnu[FRT_MAX_FREQ], GHenu[FRT_MAX_FREQ], NGnu[FRT_MAX_FREQ];
    if (sizeof(ACC) == 8) {
      double a = 0., b = 0.;
      double x[FRT_MAX_FREQ], y[FRT_MAX_FREQ], z[FRT_MAX_FREQ];
      FRT_FOR_BINS(nu) x[nu] = y[nu] = z[nu] = 0.;
      if (!LAST) {
        // pair loads issued four at a time; the sums keep the slot order
        const double2* P = (const double2*)cp.pair;
        int s = s0;
        for (; s + 4 <= s1; s += 4) {
          const double2 v0 = P[s], v1 = P[s + 1], v2 = P[s + 2], v3 = P[s + 3];
          a += v0.x;
          b += v0.y;
          a += v1.x;
          b += v1.y;
          a += v2.x;
          b += v2.y;
          a += v3.x;
          b += v3.y;
        }
        for (; s < s1; ++s) {
          const double2 v = P[s];
          a += v.x;
          b += v.y;
        }
      }
      // photon records: the next slot's record is requested before the
      // current one is processed
      const double* R = (const double*)cp.rec;
      double qn[FRT_MAX_FREQ], dn = 0.;
      if (LAST && s0 < s1) {
        const double* r = R + (int64_t)s0 * cp.rec_len;
        FRT_FOR_BINS(nu) qn[nu] = r[nu];
        dn = r[F];
      }
      for (int s = s0; LAST && s < s1; ++s) {
        {
          // per-bin deposits of the segment from its photon record
          double r[FRT_MAX_FREQ];
          FRT_FOR_BINS(nu) r[nu] = qn[nu];
          const double d = dn;
          if (s + 1 < s1) {
            const double* rn = R + (int64_t)(s + 1) * cp.rec_len;
            FRT_FOR_BINS(nu) qn[nu] = rn[nu];
            dn = rn[F];
          }
          if (isinf(d)) continue;  // behind a dump: no deposit
          const bool dump = d < 0.;
          const double distd = dump ? -d : d;
          double gs = 0., gse = 0.;
          FRT_FOR_BINS(nu) {
            const double nin = r[nu];
            const double dth = ((double)ah1 * g.sigH[nu]) * distd;
            const double dthe = ((double)ahe1 * g.sigHe[nu]) * distd;
            double gh, ghe;
            bin_deposit_fp64(g, dth, dthe, nin, dump, ah1, ahe1, gh, ghe);
            gs += gh;
            gse += ghe;
            x[nu] += gh;
            y[nu] += ghe;
            z[nu] += nin / FRT_PHOT_NORM;
          }
          a += gs;
          b += gse;
        }
      }
      GH = a;
      GHe = b;
      FRT_FOR_BINS(nu) {
        GHnu[nu] = x[nu];
        GHenu[nu] = y[nu];
        NGnu[nu] = z[nu];
      }
    } else {
      // compensated float sums in ray order
      float a = 0.f, ca = 0.f, b = 0.f, cb = 0.f;
      float x[FRT_MAX_FREQ], cx[FRT_MAX_FREQ], y[FRT_MAX_FREQ], cy[FRT_MAX_FREQ], z[FRT_MAX_FREQ];
      FRT_FOR_BINS(nu) x[nu] = cx[nu] = y[nu] = cy[nu] = z[nu] = 0.f;
      const float inv_dtv = (float)(FRT_PHOT_NORM / ((double)g.dt * g.vol));
      const float rh1 = inv_dtv / ah1, rhe1 = inv_dtv / ahe1;
      // one compensated step of each running sum
#define FRT_KAHAN2(VA, VB)      \
  {                             \
    float ya = (VA) - ca;       \
    float ta = a + ya;          \
    ca = (ta - a) - ya;         \
    a = ta;                     \
    float yb = (VB) - cb;       \
    float tb = b + yb;          \
    cb = (tb - b) - yb;         \
    b = tb;                     \
  }
      if (!LAST) {
        // pair loads issued four at a time; the sums keep the slot order
        const float2* P = (const float2*)cp.pair;
        int s = s0;
        for (; s + 4 <= s1; s += 4) {
          const float2 v0 = P[s], v1 = P[s + 1], v2 = P[s + 2], v3 = P[s + 3];
          FRT_KAHAN2(v0.x, v0.y);
          FRT_KAHAN2(v1.x, v1.y);
          FRT_KAHAN2(v2.x, v2.y);
          FRT_KAHAN2(v3.x, v3.y);
        }
        for (; s < s1; ++s) {
          const float2 v = P[s];
          FRT_KAHAN2(v.x, v.y);
        }
      }
      // photon records: the next slot's record is requested before the
      // current one is processed
      const float* R = (const float*)cp.rec;
      float qn[FRT_MAX_FREQ], dn = 0.f;
      // Five bins and whole wavefronts of cells: the wavefront's records
      // (its 64 cells' contiguous slots) are dealt to its lanes in equal
      // consecutive shares, each lane keeping compensated partial sums per
      // cell, and each cell's partials are then added in slot order by the
      // cell's lane -- a lane per cell would leave about half the lanes idle
      // (cells hold from none to well over the mean number of records).  The
      // sums are grouped differently from the one-lane walk below, so the
      // bits differ from it, and they are the same on every run.
      const bool balanced = LAST && NFC == 5 && (g.n3 & 63) == 0;
      if (balanced) {
        __shared__ float s_part[BLK / 64][128][17];  // per (cell + lane): H, He, H_nu, He_nu, photons_nu
        __shared__ int s_end[BLK / 64][64];          // prefix maximum of the cells' slot ends
        __shared__ float s_op[BLK / 64][64][2];      // the cells' opacities
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        int e = s1 > s0 ? s1 : 0;
        int m = s1 > s0 ? s0 : INT_MAX;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int u = __shfl_up(e, off, 64);
          if (lane >= off) e = max(e, u);
          m = min(m, __shfl_xor(m, off, 64));
        }
        const int r1 = __shfl(e, 63, 64), r0 = m;
        s_end[w][lane] = e;
        s_op[w][lane][0] = ah1;
        s_op[w][lane][1] = ahe1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        const int nrec = r0 < r1 ? r1 - r0 : 0;
        const int L = (nrec + 63) >> 6;
        const int q0 = r0 + lane * L, q1 = min(q0 + L, r1);
        if (nrec > 0 && q0 < q1) {
          int c = 0, hi = 63;  // the cell of record q0: the first whose end exceeds it
          while (c < hi) {
            const int mid = (c + hi) >> 1;
            if (s_end[w][mid] > q0) hi = mid;
            else c = mid + 1;
          }
          bool has = false;
          auto flush = [&](int cell) {
            float* o = s_part[w][cell + lane];
            o[0] = a - ca;
            o[1] = b - cb;
            FRT_FOR_BINS(nu) {
              o[2 + nu] = x[nu] - cx[nu];
              o[7 + nu] = y[nu] - cy[nu];
              o[12 + nu] = z[nu];
            }
            a = ca = b = cb = 0.f;
            FRT_FOR_BINS(nu) x[nu] = cx[nu] = y[nu] = cy[nu] = z[nu] = 0.f;
          };
          load_photon_record(R + (int64_t)q0 * cp.rec_len, F, cp.rec_len, qn, dn);
          for (int r = q0; r < q1; ++r) {
            while (s_end[w][c] <= r) {
              if (has) flush(c);
              has = false;
              ++c;
            }
            has = true;
            const float cah1 = s_op[w][c][0], cahe1 = s_op[w][c][1];
            const float crh1 = inv_dtv / cah1, crhe1 = inv_dtv / cahe1;
            float q[FRT_MAX_FREQ];
            FRT_FOR_BINS(nu) q[nu] = qn[nu];
            const float d = dn;
            if (r + 1 < q1) load_photon_record(R + (int64_t)(r + 1) * cp.rec_len, F, cp.rec_len, qn, dn);
            if (isinf(d)) continue;  // behind a dump: no deposit
            const bool dump = d < 0.f;
            const float dist = dump ? -d : d;
            float va = 0.f, vb = 0.f;
            FRT_FOR_BINS(nu) {
              const float nin = q[nu];
              const float dth = cah1 * g.sigHf[nu] * dist;
              const float dthe = cahe1 * g.sigHef[nu] * dist;
              float rh, rhe, eh, ehe;
              bin_deposit_mixed(g, dth, dthe, nin, dump, crh1, crhe1, rh, rhe, eh, ehe);
              va += rh;
              vb += rhe;
              float yx = rh - cx[nu];
              float tx = x[nu] + yx;
              cx[nu] = (tx - x[nu]) - yx;
              x[nu] = tx;
              float yy = rhe - cy[nu];
              float ty = y[nu] + yy;
              cy[nu] = (ty - y[nu]) - yy;
              y[nu] = ty;
              z[nu] += nin;
            }
            FRT_KAHAN2(va, vb);
          }
          if (has) flush(c);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        // this lane's cell: its partials, lanes j0..j1 in order, slot cell + j
        a = ca = b = cb = 0.f;
        FRT_FOR_BINS(nu) x[nu] = cx[nu] = y[nu] = cy[nu] = z[nu] = 0.f;
        if (s1 > s0) {
          const int j0 = (s0 - r0) / L, j1 = (s1 - 1 - r0) / L;
          for (int j = j0; j <= j1; ++j) {
            const float* o = s_part[w][lane + j];
            FRT_KAHAN2(o[0], o[1]);
            FRT_FOR_BINS(nu) {
              float yx = o[2 + nu] - cx[nu];
              float tx = x[nu] + yx;
              cx[nu] = (tx - x[nu]) - yx;
              x[nu] = tx;
              float yy = o[7 + nu] - cy[nu];
              float ty = y[nu] + yy;
              cy[nu] = (ty - y[nu]) - yy;
              y[nu] = ty;
              z[nu] += o[12 + nu];
            }
          }
        }
      }
      if (LAST && !balanced && s0 < s1) load_photon_record(R + (int64_t)s0 * cp.rec_len, F, cp.rec_len, qn, dn);
      for (int s = s0; LAST && !balanced && s < s1; ++s) {
        float va, vb;
        {
          float q[FRT_MAX_FREQ];
          FRT_FOR_BINS(nu) q[nu] = qn[nu];
          const float d = dn;
          if (s + 1 < s1) load_photon_record(R + (int64_t)(s + 1) * cp.rec_len, F, cp.rec_len, qn, dn);
          if (isinf(d)) continue;  // behind a dump: no deposit
          const bool dump = d < 0.f;
          const float dist = dump ? -d : d;
          va = 0.f;
          vb = 0.f;
          FRT_FOR_BINS(nu) {
            const float nin = q[nu];
            const float dth = ah1 * g.sigHf[nu] * dist;
            const float dthe = ahe1 * g.sigHef[nu] * dist;
            float rh, rhe, eh, ehe;
            bin_deposit_mixed(g, dth, dthe, nin, dump, rh1, rhe1, rh, rhe, eh, ehe);
            va += rh;
            vb += rhe;
            float yx = rh - cx[nu];
            float tx = x[nu] + yx;
            cx[nu] = (tx - x[nu]) - yx;
            x[nu] = tx;
            float yy = rhe - cy[nu];
            float ty = y[nu] + yy;
            cy[nu] = (ty - y[nu]) - yy;
            y[nu] = ty;
            z[nu] += nin;
          }
        }
        FRT_KAHAN2(va, vb);
      }
#undef FRT_KAHAN2
      GH = (double)a;
      GHe = (double)b;
      FRT_FOR_BINS(nu) {
        GHnu[nu] = (double)x[nu];
        GHenu[nu] = (double)y[nu];
        NGnu[nu] = (double)z[nu];
      }
    }

    if (tape_g) {  // adjoint tape: this pass's Gamma sums
      tape_g[k] = GH;
      tape_g[n3 + k] = GHe;
    }
    if (LAST) {
      // per-bin rates depend on the sums only: stored here, so they are not
      // held in registers through the chemistry below
      FRT_FOR_BINS(nu) {
        cf.gamma_h1_nu[nu * n3 + k] = (float)GHnu[nu];
        cf.gamma_he1_nu[nu * n3 + k] = (float)GHenu[nu];
        cf.ngam_nu[nu * n3 + k] = (float)NGnu[nu];
      }
    }
    const float nH = cf.f[C_NH][k];
    const float nHe = cf.f[C_NHE][k];
    const float clump = cf.f[C_CLUMP][k];
    float rH2, rHe2, dne;
    if (FIRST) {
      const float T = cf.f[C_TEMP][k];
      const double Tm = (double)frt_maxd(g.temp_min, T);
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
      dne = -cf.f[C_NE][k];
      if (!LAST) cf.f[C_DNE_DT][k] = dne;
    } else {
      rH2 = cf.f[C_RECOMB_H2][k];
      rHe2 = cf.f[C_RECOMB_HE2][k];
      dne = cf.f[C_DNE_DT][k];
    }
    const double dtd = (double)g.dt;
    const double nelec = (double)(nH - ah1 + nHe - ahe1);  // float sum, then double
    const double x0 = 1. - (double)cf.f[C_FH1][k];
    const double pr = GH + (double)(clump * rH2) * nelec;
    const double xeq = GH / pr;
    const double ti = 1. / pr;
    const double e_h = dm_exp(-dtd / ti);
    float avg_h1 = (float)((double)nH * (1. - (xeq + ti / dtd * (x0 - xeq) * (1. - e_h))));
    const double x0he = 1. - (double)cf.f[C_FHE1][k];
    const double prhe = GHe + (double)(clump * rHe2) * nelec;
    const double xeqhe = GHe / prhe;
    const double tihe = 1. / prhe;
    const double e_he = dm_exp(-dtd / tihe);
    float avg_he1 = (float)((double)nHe * (1. - (xeqhe + tihe / dtd * (x0he - xeqhe) * (1. - e_he))));
    if (g.dither_seed) {  // rounding-sensitivity measurement (frt_rng.h)
      const uint64_t r = (uint64_t)frt_cell_to_rowmajor(g.lay, (int)k);
      const uint64_t st = (uint64_t)sd->step;  // completed steps
      avg_h1 = frt_ulp_step(avg_h1, frt_dither_up(g.dither_seed, st, g.pass, r, 0));
      avg_he1 = frt_ulp_step(avg_he1, frt_dither_up(g.dither_seed, st, g.pass, r, 1));
    }
    cf.f[C_AVG_NH1][k] = avg_h1;
    cf.f[C_AVG_NHE1][k] = avg_he1;
    cf.avg2[k] = make_float2(avg_h1, avg_he1);

    if (LAST) {
      const float nH1n = (float)((double)nH * (1. - (xeq + (x0 - xeq) * e_h)));
      const float nH2n = nH - nH1n;
      const float nHe1n = (float)((double)nHe * (1. - (xeqhe + (x0he - xeqhe) * e_he)));
      const float nHe2n = nHe - nHe1n;
      if (((double)(avg_h1 / nH) < 0.5) && (cf.f[C_TREION_FLAG][k] == 0.f)) {
        cf.f[C_TREION][k] = sd->t;
        cf.f[C_TREION_FLAG][k] = 1.f;
      }
      cf.f[C_GAMMA_H1_TOT][k] = (float)GH;
      cf.f[C_GAMMA_HE1_TOT][k] = (float)GHe;
      cf.f[C_RECRATE_H2][k] = (float)((double)(clump * rH2) * nelec * (double)nH2n);
      cf.f[C_RECRATE_HE2][k] = (float)((double)(clump * rHe2) * nelec * (double)nHe2n);
      const float ne_n = nH2n + nHe2n;  // + 2 nHe3, nHe3 = 0
      float dne_n = dne + ne_n;
      dne_n = dne_n / g.dt;
      const float ntot_n = nH + nHe + ne_n;
      const float fH2n = nH2n / nH;
      cf.f[C_NH1][k] = nH1n;
      cf.f[C_NH2][k] = nH2n;
      cf.f[C_NHE1][k] = nHe1n;
      cf.f[C_NHE2][k] = nHe2n;
      cf.f[C_NE][k] = ne_n;
      cf.f[C_DNE_DT][k] = dne_n;
      cf.f[C_NTOT][k] = ntot_n;
      cf.f[C_FH2][k] = fH2n;
      cf.f[C_FHE2][k] = nHe2n / nHe;
      cf.f[C_FH1][k] = nH1n / nH;
      cf.f[C_FHE1][k] = nHe1n / nHe;
      cf.f[C_FHE3][k] = 0.f / nHe;  // nHe3 = 0 (ray_tracing.cc:1128,1194)
      // mean free paths from the opacity this pass used (ray_tracing.cc:1078-1081)
      cf.f[C_MFP][k] = (float)(1. / g.sigH[0] / (double)ah1);
      cf.f[C_MFP912][k] = (float)(1. / g.sig912 / (double)ah1);
      if (THERMAL) {
        if (g.temp_ev) {
          ThermalIn in;
          in.nH = nH;
          in.nHe = nHe;
          in.nH1 = nH1n;
          in.nH2 = nH2n;
          in.nHe2 = nHe2n;
          in.ne = ne_n;
          in.ntot = ntot_n;
          in.dne = dne_n;
          in.fH2 = fH2n;
          in.ah1 = avg_h1;
          in.ahe1 = avg_he1;
          in.nH1_prev = cf.f[C_NH1_PREV][k];
          thermal_density_terms(cf, g, k, in);
          float temp = cf.f[C_TEMP][k];
          float heat = cf.f[C_HEAT][k];
          float cool = cf.f[C_COOL][k];
          float ifs, inc;
          nsub = thermal_cell(g, in, cf.gamma_h1_nu, cf.gamma_he1_nu, n3, k, temp, heat, cool, ifs, inc);
          cf.f[C_TEMP][k] = temp;
          cf.f[C_HEAT][k] = heat;
          cf.f[C_COOL][k] = cool;
          cf.f[C_IF_SPEED][k] = ifs;
          cf.f[C_INC_FLUX][k] = inc;
          cf.f[C_NH1_PREV][k] = nH1n;
        } else {
          cf.f[C_TEMP][k] = ((double)(nH1n / nH) > 0.5) ? 100.f : 1e4f;
          cf.f[C_IF_SPEED][k] = 0.f;
          cf.f[C_INC_FLUX][k] = 0.f;
        }
      }
    }
  }
  if (LAST && THERMAL) block_add(&cnt->substeps, (unsigned long long)nsub);
}

// Standalone update_thermal (used when the phases are called one by one).
__global__ void __launch_bounds__(BLK) k_thermal(CellPtrs cf, GridConst g, DevCounters* cnt) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int nsub = 0;
  if (k < g.n3) {
    if (g.temp_ev) {
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
      in.nH1_prev = cf.f[C_NH1_PREV][k];
      thermal_density_terms(cf, g, k, in);
      float temp = cf.f[C_TEMP][k];
      float heat = cf.f[C_HEAT][k];
      float cool = cf.f[C_COOL][k];
      float ifs, inc;
      nsub = thermal_cell(g, in, cf.gamma_h1_nu, cf.gamma_he1_nu, g.n3, k, temp, heat, cool, ifs, inc);
      cf.f[C_TEMP][k] = temp;
      cf.f[C_HEAT][k] = heat;
      cf.f[C_COOL][k] = cool;
      cf.f[C_IF_SPEED][k] = ifs;
      cf.f[C_INC_FLUX][k] = inc;
      cf.f[C_NH1_PREV][k] = in.nH1;
    } else {
      cf.f[C_TEMP][k] = ((double)(cf.f[C_NH1][k] / cf.f[C_NH][k]) > 0.5) ? 100.f : 1e4f;
      cf.f[C_IF_SPEED][k] = 0.f;
      cf.f[C_INC_FLUX][k] = 0.f;
    }
  }
  block_add(&cnt->substeps, (unsigned long long)nsub);
}

// ---------------------------------------------------------------------------
// c2ray_iter = FALSE (FP64 form): the march deposits into u_nu segment by
// segment and attenuates the ray by the HI opacity alone (ray_tracing.cc:
// 1329-1355 in the crossing loop, 1443-1470 for the last segment, the same
// expressions), then re-labels the ray's cell (:1507-1509).  One thread per
// moved ray writes each segment's per-bin u_nu addition (float, as the
// reference adds it) to the segment's cell-sorted slot; k_unu_sum adds them
// per cell in slot (= ray list) order.
__global__ void __launch_bounds__(BLK) k_deposit_inline(RaySoA rb, const int32_t* __restrict__ moved,
                                                        const StepDev* __restrict__ sd, GridConst g,
                                                        const int32_t* __restrict__ path_cell,
                                                        const float* __restrict__ path_dist,
                                                        const int32_t* __restrict__ seg_cnt,
                                                        const int32_t* __restrict__ path_slot,
                                                        const float* __restrict__ nH1, ContribPtrs cp,
                                                        float* __restrict__ psum) {
  const int F = g.nfreq;
  const double vol = g.vol;  // ((double)dx)^3 as the reference's ddx*ddy*ddz
  const int64_t nmoved = sd->nmoved;
  // a slot holds F float additions in the first F floats of its record
  const int64_t stride = 2 * (int64_t)cp.rec_len;
  for_rays(seg_cnt, moved, nmoved, nullptr, 0, [&](int64_t i, int c) {
    float nph[8];
    phot_load(rb, i, F, g.cap, nph);
    int last_cell = 0;
    for (int p = 0; p < c; ++p) {
      const int cell = path_cell[(int64_t)p * g.cap + i];
      const float dist = path_dist[(int64_t)p * g.cap + i];
      const int slot = path_slot[(int64_t)p * g.cap + i];
      const float n1 = nH1[cell];
      float* rec = (float*)cp.rec + (int64_t)slot * stride;
      last_cell = cell;
      FRT_FOR_BINS(nu) {
        const double nphot = (double)nph[nu] * FRT_PHOT_NORM;
        const double dtau = (double)n1 * g.sigH[nu] * (double)dist;
        const double att = dm_exp(-dtau);
        const double nphot_add = nphot * (1. - att);
        const double gamma = nphot_add / (double)g.dt / vol / (double)n1;
        rec[nu] = (float)((double)g.spectrum[nu] * gamma * FRT_H_PLANCK * g.nu[nu] / (double)g.clight / g.sigH[nu]);
        nph[nu] *= (float)att;
      }
    }
    phot_store(rb, i, F, g.cap, nph);
    psum[i] = phot_sum(nph, F);
    ray_set_cell(rb, i, last_cell);
  });
}

// u_nu of every cell: zero (rt_funcs.cc:174-189), plus the additions of the
// cell's segments in slot order, in float.
__global__ void __launch_bounds__(BLK) k_unu_sum(CellPtrs cf, GridConst g, ContribPtrs cp,
                                                 const int32_t* __restrict__ cstart,
                                                 const int32_t* __restrict__ cend) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= g.n3) return;
  const int F = g.nfreq;
  const int64_t stride = 2 * (int64_t)cp.rec_len;
  float u[8];
  FRT_FOR_BINS(nu) u[nu] = 0.f;
  const int s0 = cstart[k], s1 = cend[k];
  for (int sl = s0; sl < s1; ++sl) {
    const float* rec = (const float*)cp.rec + (int64_t)sl * stride;
    FRT_FOR_BINS(nu) u[nu] += rec[nu];
  }
  FRT_FOR_BINS(nu) cf.u_nu[(int64_t)nu * g.n3 + k] = u[nu];
}

// update_gamma (gas_funcs.cc:74-99, collisional ionization off) and
// update_chem (:299-351) with solve_ion (:243-296) iterated five times, one
// thread per cell.  In solve_ion's first helium branch (He III the most
// abundant ion at the start of the step) the reference's (nHe[i] - nHe1[i])
// is a difference of row pointers of two arrays (undefined); the cell values
// are used.  Photons below 54.4 eV leave gamma_He2_tot = 0, so with the
// usual initial state (no He III) that branch is not reached.
__global__ void __launch_bounds__(BLK) k_gamma_chem(CellPtrs cf, GridConst g) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= g.n3) return;
  const int F = g.nfreq;
  const int64_t n3 = g.n3;
  // update_gamma: float fields += double expressions
  float gH1 = 0.f, gHe1 = 0.f, gHe2 = 0.f;
  FRT_FOR_BINS(nu) {
    const double u = (double)cf.u_nu[(int64_t)nu * n3 + k];
    const double cl = (double)g.clight;
    gH1 = (float)((double)gH1 + g.sigH[nu] * cl * u / FRT_H_PLANCK / g.nu[nu]);
    gHe1 = (float)((double)gHe1 + g.sigHe[nu] * cl * u / FRT_H_PLANCK / g.nu[nu]);
    gHe2 = (float)((double)gHe2 + g.sigHe2[nu] * cl * u / FRT_H_PLANCK / g.nu[nu]);
  }
  // update_chem
  const float dt = g.dt;
  const float C = cf.f[C_CLUMP][k];
  const float nH = cf.f[C_NH][k], nHe = cf.f[C_NHE][k];
  const float nH1p = cf.f[C_NH1_PREV][k], nH2p = cf.f[C_NH2_PREV][k];
  const float nHe1p = cf.f[C_NHE1_PREV][k], nHe2p = cf.f[C_NHE2_PREV][k], nHe3p = cf.f[C_NHE3_PREV][k];
  float nH1 = cf.f[C_NH1][k], nH2 = cf.f[C_NH2][k];
  float nHe1 = cf.f[C_NHE1][k], nHe2 = cf.f[C_NHE2][k], nHe3 = cf.f[C_NHE3][k];
  float ne = cf.f[C_NE][k];
  const float temp = cf.f[C_TEMP][k];
  float dne_dt = -ne;
  const double T = (double)temp;
  float rH2, rHe2, rHe3;
  if (!g.rec_case_b) {
    rH2 = (float)frt_alphaA_H2(T);
    rHe2 = (float)frt_alphaA_He2(T);
    rHe3 = (float)frt_alphaA_He3(T);
  } else {
    rH2 = (float)frt_alphaB_H2(T);
    rHe2 = (float)frt_alphaB_He2(T);
    rHe3 = (float)frt_alphaB_He3(T);
  }
  if (((double)temp >= 3e4) && ((double)temp <= 1e6)) rHe2 += (float)frt_Dalpha_He2(T);
  for (int m = 0; m < 5; ++m) {
    if (nH1p >= nH2p) {
      nH2 = (float)((double)(nH2p + gH1 * nH * dt) / (1. + (double)(gH1 * dt) + (double)(C * rH2 * ne * dt)));
      nH1 = nH - nH2;
    } else {
      nH1 = (float)((double)(nH1p + C * rH2 * ne * nH * dt) / (1. + (double)(gH1 * dt) + (double)(C * rH2 * ne * dt)));
      nH2 = nH - nH1;
    }
    if ((nHe1p <= nHe3p) && (nHe2p <= nHe3p)) {
      nHe1 = (float)((double)(nHe1p + C * rHe2 * ne * (nHe - nHe3) * dt) /
                     (1. + (double)(gHe1 * dt) + (double)(C * rHe2 * ne * dt)));
      nHe2 = (float)((double)(nHe2p + gHe1 * (nHe - nHe3) * dt + C * rHe3 * ne * (nHe - nHe1) * dt) /
                     (1. + (double)((gHe1 + gHe2) * dt) + (double)(C * (rHe2 + rHe3) * ne * dt)));
      nHe3 = nHe - nHe1 - nHe2;
    } else if ((nHe1p <= nHe2p) && (nHe3p <= nHe2p)) {
      nHe1 = (float)((double)(nHe1p + C * rHe2 * ne * (nHe - nHe3) * dt) /
                     (1. + (double)(gHe1 * dt) + (double)(C * rHe2 * ne * dt)));
      nHe3 = (float)((double)(nHe3p + gHe2 * (nHe - nHe1) * dt) / (1. + (double)(gHe2 * dt) + (double)(C * rHe3 * ne * dt)));
      nHe2 = nHe - nHe1 - nHe3;
    } else {
      nHe2 = (float)((double)(nHe2p + gHe1 * (nHe - nHe3) * dt + C * rHe3 * ne * (nHe - nHe1) * dt) /
                     (1. + (double)((gHe1 + gHe2) * dt) + (double)(C * (rHe2 + rHe3) * ne * dt)));
      nHe3 = (float)((double)(nHe3p + gHe2 * (nHe - nHe1) * dt) / (1. + (double)(gHe2 * dt) + (double)(C * rHe3 * ne * dt)));
      nHe1 = nHe - nHe2 - nHe3;
    }
    ne = (float)((double)(nH2 + nHe2) + 2. * (double)nHe3);
  }
  const float ntot = nH + nHe + ne;
  dne_dt += ne;
  dne_dt /= dt;
  cf.f[C_GAMMA_H1_TOT][k] = gH1;
  cf.f[C_GAMMA_HE1_TOT][k] = gHe1;
  cf.f[C_RECOMB_H2][k] = rH2;
  cf.f[C_RECOMB_HE2][k] = rHe2;
  cf.f[C_RECOMB_HE3][k] = rHe3;
  cf.f[C_NH1][k] = nH1;
  cf.f[C_NH2][k] = nH2;
  cf.f[C_NHE1][k] = nHe1;
  cf.f[C_NHE2][k] = nHe2;
  cf.f[C_NHE3][k] = nHe3;
  cf.f[C_NE][k] = ne;
  cf.f[C_NTOT][k] = ntot;
  cf.f[C_DNE_DT][k] = dne_dt;
  cf.f[C_NH1_PREV][k] = nH1;
  cf.f[C_NH2_PREV][k] = nH2;
  cf.f[C_NHE1_PREV][k] = nHe1;
  cf.f[C_NHE2_PREV][k] = nHe2;
  cf.f[C_NHE3_PREV][k] = nHe3;
  cf.f[C_FH1][k] = nH1 / nH;
  cf.f[C_FH2][k] = nH2 / nH;
  cf.f[C_FHE1][k] = nHe1 / nHe;
  cf.f[C_FHE2][k] = nHe2 / nHe;
  cf.f[C_FHE3][k] = nHe3 / nHe;
}

// ---------------------------------------------------------------------------
// Merging (ray_tracing.cc:176-227 calc_nphot_merge, :324-477 merge).
__global__ void __launch_bounds__(BLK) k_count_alive(const uint8_t* __restrict__ alive, const StepDev* __restrict__ sd,
                                                     DevCounters* cnt) {
  const int64_t n = sd->nrays;
  unsigned long long c = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK)
    c += alive[i] ? 1 : 0;
  block_add(&cnt->alive, c);
}

// Sum of one counter over the copies (the host sums them the same way).
__device__ __forceinline__ unsigned long long counter_total(const DevCounters* cnt, const unsigned long long* field0) {
  constexpr unsigned stride = sizeof(DevCounters) / sizeof(unsigned long long);
  unsigned long long t = 0;
  for (int r = 0; r < FRT_CTR_COPIES; ++r) t += field0[r * stride];
  return t;
}

// count_rays against the merge threshold (ray_tracing.cc:157-170): whether
// calc_nphot_merge runs, and the cleared histogram it fills.
__global__ void __launch_bounds__(BLK) k_merge_setup(StepDev* sd, const DevCounters* cnt, int64_t num_thresh,
                                                     int32_t* __restrict__ pdf) {
  __shared__ int on;
  if (threadIdx.x == 0) {
    const int64_t count = (int64_t)counter_total(cnt, &cnt->alive);
    on = count > num_thresh ? 1 : 0;
    sd->pdf_on = on;
  }
  __syncthreads();
  if (on)
    for (int b = threadIdx.x; b < FRT_PDF_BINS; b += BLK) pdf[b] = 0;
}

// Photon-count histogram of calc_nphot_merge: privatized per block in LDS,
// non-empty bins flushed with one global atomic each (integer sums: the
// result does not depend on order).
static constexpr unsigned PDF_GRID = 1024;

__global__ void __launch_bounds__(BLK) k_pdf(const float* __restrict__ psum, const StepDev* __restrict__ sd,
                                             const uint8_t* __restrict__ alive, double lg_min,
                                             double del, int32_t* __restrict__ pdf) {
  if (!sd->pdf_on) return;
  const int64_t nrays = sd->nrays;
  __shared__ int32_t h[FRT_PDF_BINS];
  for (int b = threadIdx.x; b < FRT_PDF_BINS; b += BLK) h[b] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < nrays; i += (int64_t)gridDim.x * BLK) {
    if (!alive[i]) continue;
    const float s = psum[i];
    // (int) maxd((log10(sum) - lg_min)/del, 0.): maxd narrows to float
    int ind = (int)frt_maxd((float)((dm_log10((double)s) - lg_min) / del), 0.f);
    if (ind > FRT_PDF_BINS - 1) ind = FRT_PDF_BINS - 1;
    if (ind < 0) ind = 0;  // only reachable for a NaN photon count
    atomicAdd(&h[ind], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < FRT_PDF_BINS; b += BLK)
    if (h[b]) atomicAdd(&pdf[b], h[b]);
}

// calc_nphot_merge's top-down walk (ray_tracing.cc:205-219): nphot_merge =
// 10^(b del + lg_min) for the highest bin b whose cumulative count from the
// top reaches num_thresh (kept from the last time when count_rays did not
// exceed the threshold, 0 if no bin reaches it).  One block: per-thread sums
// of consecutive bins, a suffix scan of those, then the last qualifying bin
// of each thread's range; the largest wins.  Then this step's merge switch
// and rotation index.
__global__ void __launch_bounds__(BLK) k_pdf_walk(StepDev* sd, const int32_t* __restrict__ pdf, int64_t num_thresh,
                                                  double lg_min, double del, uint64_t seed) {
  constexpr int PER = (FRT_PDF_BINS + BLK - 1) / BLK;
  __shared__ int64_t tsum[BLK];
  __shared__ int best;
  const int t = threadIdx.x;
  if (sd->pdf_on) {
    const int b0 = t * PER, b1 = min(b0 + PER, FRT_PDF_BINS);
    int64_t own = 0;
    for (int b = b0; b < b1; ++b) own += pdf[b];
    tsum[t] = own;
    if (t == 0) best = -1;
    __syncthreads();
    // counts in the bins above this thread's range
    int64_t above = 0;
    for (int q = t + 1; q < BLK; ++q) above += tsum[q];
    int64_t cum = above;
    int found = -1;
    for (int b = b1 - 1; b >= b0; --b) {
      cum += pdf[b];
      if (cum >= num_thresh) {
        found = b;
        break;
      }
    }
    if (found >= 0) atomicMax(&best, found);
    __syncthreads();
    if (t == 0) sd->nphot_merge = best >= 0 ? (float)dm_pow(10., best * del + lg_min) : 0.f;
  }
  if (t == 0) {
    const int on = (sd->nphot_merge > 0.f) && (sd->nrays > 0);
    sd->merge_on = on;
    sd->mcount = on ? sd->nrays + 1 : 0;
    sd->itm = frt_itheta_hash(seed, FRT_STREAM_MERGE + (uint64_t)sd->step, 0);
  }
}

__global__ void __launch_bounds__(BLK) k_merge_flag(RaySoA rb, const StepDev* __restrict__ sd, GridConst g,
                                                    const uint8_t* __restrict__ alive,
                                                    const float* __restrict__ psum,
                                                    const float* __restrict__ rot_inv,
                                                    int32_t* __restrict__ cand,
                                                    uint32_t* __restrict__ key) {
  if (!sd->merge_on) return;
  constexpr int WIN = 128;  // rays per window of a wavefront
  __shared__ int s_cand[BLK / 64][WIN];
  const int64_t nrays = sd->nrays;
  const float nphot_merge = sd->nphot_merge;
  const float* inv = rot_inv + 9 * (int64_t)sd->itm;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t nwaves = (int64_t)gridDim.x * (BLK / 64);
  // each wavefront flags a window of 128 rays, lists its candidates (ballots)
  // and computes their keys one candidate per lane: the HEALPix bin of the
  // direction is the costly part, and a lane per ray would compute it with
  // most lanes idle
  for (int64_t base = ((int64_t)blockIdx.x * (BLK / 64) + w) * WIN; base <= nrays; base += nwaves * WIN) {
    int nc = 0;
#pragma unroll
    for (int j = 0; j < WIN / 64; ++j) {
      const int64_t i = base + j * 64 + lane;
      const bool c = i < nrays && alive[i] && psum[i] < nphot_merge;
      if (i <= nrays) cand[i] = c ? 1 : 0;
      const uint64_t m = __ballot(c);
      if (c) s_cand[w][nc + __popcll(m & lt)] = (int)(i - base);
      nc += __popcll(m);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    for (int t = lane; t < nc; t += 64) {
      const int64_t i = base + s_cand[w][t];
      const float4 d = rb.dir[i];
      float vx, vy, vz;
      rotate3(inv, d.x, d.y, d.z, &vx, &vy, &vz);  // get_unit_vector(.., 6)
      const int bin = hpx_vec2pix_nest(g.hpx_lvl, (double)vx, (double)vy, (double)vz);
      const int nb = 12 << (2 * g.hpx_lvl);
      key[i] = (uint32_t)ray_cell(rb, i) * (uint32_t)nb + (uint32_t)bin;
    }
    __builtin_amdgcn_wave_barrier();  // the list is read before the next window writes it
  }
}

// l_from_r (ray_tracing.cc:87-95).
__device__ __forceinline__ int l_from_r_dev(const GridConst& g, float r, float safety) {
  const float q = r / g.dx0;
  if ((double)q < 1.) return g.hpx_lvl;
  const double arg = FRT_PI_REF * (double)safety / 3. * ((double)q * (double)q);
  int l = (int)ceil(dm_log(arg) / g.log4);  // g.log4 = dm_log(4.)
  if (l > g.l_max) l = g.l_max;
  if (l < g.hpx_lvl) l = g.hpx_lvl;
  return l;
}

// One thread per occupied (cell, direction) bin: accumulates its members in
// list order exactly as the serial phase 1, then normalizes (phase 2).  Each
// wavefront takes windows of 128 sorted candidates and lists the group heads
// among them (ballots), then gives one head to each lane: groups hold about
// two members, so a lane per candidate would leave half the lanes idle.
__global__ void __launch_bounds__(BLK) k_merge_groups(RaySoA rb, const StepDev* __restrict__ sd, GridConst g,
                                                      const uint32_t* __restrict__ key,
                                                      const int32_t* __restrict__ val,
                                                      uint8_t* __restrict__ alive, const float* __restrict__ rot_fwd,
                                                      float safety, DevCounters* cnt, float* __restrict__ psum) {
  constexpr int WIN = 128;  // candidates per window
  __shared__ int s_head[BLK / 64][WIN];
  const int64_t ncand = sd->ncand;
  const uint32_t itheta_m = sd->itm;
  const float* fwd = rot_fwd + 9 * (int64_t)itheta_m;
  unsigned long long nheads = 0, merged = 0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t nwaves = (int64_t)gridDim.x * (BLK / 64);
  for (int64_t base = ((int64_t)blockIdx.x * (BLK / 64) + w) * WIN; base < ncand; base += nwaves * WIN) {
  int nh = 0;
#pragma unroll
  for (int j = 0; j < WIN / 64; ++j) {
    const int64_t c = base + j * 64 + lane;
    const bool hd = c < ncand && (c == 0 || key[c - 1] != key[c]);
    const uint64_t m = __ballot(hd);
    if (hd) s_head[w][nh + __popcll(m & lt)] = (int)(c - base);
    nh += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  if (lane == 0) nheads += (unsigned long long)nh;
  for (int t = lane; t < nh; t += 64) {
    const int64_t s = base + s_head[w][t];
    const int F = g.nfreq;
    const int64_t cap = g.cap;
    int64_t e = s + 1;
    while (e < ncand && key[e] == key[s]) ++e;
    const int b = val[s];
    Ray rbase;
    ray_load(rb, b, rbase);
    float NP[8];
    phot_load(rb, b, F, cap, NP);
    const float nb = phot_sum(NP, F);
    const float rrb = rbase.r / g.dx0;
    float R = (float)((double)rrb * (double)rrb * (double)nb);
    float DR = rbase.dr * nb;
    float X = rbase.x * nb;
    float Y = rbase.y * nb;
    float Z = rbase.z * nb;
    float T = rbase.time * nb;
    for (int64_t j = s + 1; j < e; ++j) {
      const int m = val[j];
      float nmv[8];
      phot_load(rb, m, F, cap, nmv);
      const float nm = phot_sum(nmv, F);
      FRT_FOR_BINS(nu) NP[nu] += nmv[nu];
      const float4 pm = rb.pos[m];
      const float rrm = rb.dir[m].w / g.dx0;
      R = (float)((double)R + (double)rrm * (double)rrm * (double)nm);
      DR += __int_as_float(rb.meta[m].x) * nm;
      X += pm.x * nm;
      Y += pm.y * nm;
      Z += pm.z * nm;
      T += pm.w * nm;
      alive[m] = 0;
      ++merged;
    }
    float ntot = 0.f;
    FRT_FOR_BINS(nu) ntot += NP[nu];
    R = R / ntot;
    R = (float)sqrt((double)R);
    R = R * g.dx0;
    DR = DR / ntot;
    X = X / ntot;
    Y = Y / ntot;
    Z = Z / ntot;
    T = T / ntot;
    if (e - s > 1) {
      const int nbins = 12 << (2 * g.hpx_lvl);
      const int bin = (int)(key[s] % (uint32_t)nbins);
      double vx, vy, vz;
      hpx_pix2vec_nest(g.hpx_lvl, bin, &vx, &vy, &vz);
      const int order = l_from_r_dev(g, R, safety);
      const float rmin = g.rmin_tab[order];
      const float rmax = g.rmax_tab[order];
      R = frt_maxd(rmax - DR, rmin);
      DR = rmax - R;
      const int pixel = hpx_vec2pix_nest(order, vx, vy, vz);
      hpx_pix2vec_nest(order, pixel, &vx, &vy, &vz);
      float ox, oy, oz;
      rotate3(fwd, (float)vx, (float)vy, (float)vz, &ox, &oy, &oz);
      rbase.order = order;
      rbase.pixel = pixel;
      rbase.itheta = (int)itheta_m;
      rbase.xhat = ox;
      rbase.yhat = oy;
      rbase.zhat = oz;
    }
    phot_store(rb, b, F, cap, NP);
    psum[b] = ntot;  // = phot_sum(NP)
    rbase.r = R;
    rbase.dr = DR;
    rbase.x = X;
    rbase.y = Y;
    rbase.z = Z;
    rbase.time = T;
    ray_store(rb, b, rbase);
  }
  __builtin_amdgcn_wave_barrier();  // the window's head list is read before the next is written
  }
  block_add(&cnt->groups, nheads);
  block_add(&cnt->merged_away, merged);
}

// ---------------------------------------------------------------------------
// split_rays decision (rt_funcs.cc:211-237): rays out per input ray.
__global__ void __launch_bounds__(BLK) k_split_flag(RaySoA rb, const StepDev* __restrict__ sd, GridConst g,
                                                    const uint8_t* __restrict__ alive,
                                                    const float* __restrict__ psum,
                                                    int nosplit, int32_t* __restrict__ n_out,
                                                    DevCounters* cnt) {
  const int64_t nrays = sd->nrays;
  // maxd(4.*nphot_merge, 4.*nphot_min): both narrowed to float
  const float split_thresh =
      nosplit ? INFINITY : frt_maxd((float)(4. * (double)sd->nphot_merge), (float)(4. * g.nphot_min));
  unsigned long long nsplit = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i <= nrays; i += (int64_t)gridDim.x * BLK) {
    int out = 0;
    if (i < nrays && alive[i]) {
      const int l = rb.meta[i].w & 0xffff;
      const float r = rb.dir[i].w;
      // (a compaction alone splits nothing and reads no photon sum)
      const bool split = !nosplit && (r > g.rmax_tab[l]) && (l < g.l_max) && (psum[i] > split_thresh);
      out = split ? 4 : 1;
      nsplit += split ? 1 : 0;
    }
    n_out[i] = out;
  }
  block_add(&cnt->splits, nsplit);
}

// Tile of the cell of every output ray of the compaction (key of the stable
// sort that fixes the list order); values are the compacted positions, src
// records (input ray << 2 | child).  One thread per input ray.  The four
// children of the split rays of a wavefront are spread over its lanes (a
// split is rare, and its child geometry costs hundreds of instructions), so
// a wavefront with k splits computes 4k children in ceil(4k / 64) rounds
// instead of four rounds per split lane.
// Capacity check of the split's output (ray_tracing.cc:275-287): the rays
// after the split plus the next new_rays must fit; otherwise nothing more of
// the step is written and the host reports the error.
__global__ void k_split_check(StepDev* sd, int64_t cap, int64_t nnew) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sd->err) {
    sd->ncompact = 0;
    return;
  }
  if (sd->nsplit + nnew > cap) {
    sd->err |= FRT_SDERR_SPLIT_CAPACITY;
    sd->ncompact = 0;
  } else {
    sd->ncompact = sd->nsplit;
  }
}

__global__ void k_split_done(StepDev* sd) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !sd->err) sd->nrays = sd->ncompact;
}

__global__ void __launch_bounds__(BLK) k_out_keys(RaySoA src, const StepDev* __restrict__ sd, GridConst g,
                                                  const int32_t* __restrict__ n_out,
                                                  const int32_t* __restrict__ out_off,
                                                  const float* __restrict__ rot_fwd,
                                                  uint32_t* __restrict__ key, int32_t* __restrict__ val) {
  __shared__ int s_par[BLK];  // split rays of each wavefront, in lane order
  if (sd->ncompact == 0) return;
  const int64_t nrays = sd->nrays;
  const int lane = threadIdx.x & 63;
  const int wbase = threadIdx.x & ~63;
  // block-uniform trip count: the ballot and the barrier see every lane
  for (int64_t base = (int64_t)blockIdx.x * BLK; base < nrays; base += (int64_t)gridDim.x * BLK) {
  const int64_t i = base + threadIdx.x;
  const int no = i < nrays ? n_out[i] : 0;
  if (no == 1) {
    const int o = out_off[i];
    key[o] = frt_tile_key_internal(g.lay, ray_cell(src, i));
    val[o] = (int32_t)((uint32_t)i << 2);  // the source: ray i (ch 0), carried by the sort
  }
  const unsigned long long mask = __ballot(no == 4);
  if (no == 4) s_par[wbase + __popcll(mask & ((1ull << lane) - 1ull))] = (int)i;
  __syncthreads();
  const int nsplit = __popcll(mask);
  for (int t = lane; t < 4 * nsplit; t += 64) {
    const int ip = s_par[wbase + (t >> 2)];
    const int ch = t & 3;
    const int o = out_off[ip];
    Ray par;
    ray_load(src, ip, par);
    key[o + ch] = frt_tile_key_internal(g.lay, child_geometry(par, ch, g, rot_fwd).cell);
    val[o + ch] = (int32_t)(((uint32_t)ip << 2) | (uint32_t)ch);
  }
  __syncthreads();  // s_par is rewritten by the next round
  }
}

// Compaction into the other buffer in stable tile order, as a gather: output q
// takes the ray (or split child) that sorted to position q.
__global__ void __launch_bounds__(BLK) k_compact(RaySoA src, RaySoA dst, const StepDev* __restrict__ sd, GridConst g,
                                                 const int32_t* __restrict__ n_out,
                                                 const uint32_t* __restrict__ sorted_src,
                                                 const float* __restrict__ rot_fwd) {
  const int64_t total = sd->ncompact;
  const int F = g.nfreq;
  const int nplanes = frt_phot_planes(F);
  const int64_t cap = g.cap;
  for (int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x; q < total; q += (int64_t)gridDim.x * BLK) {
    const uint32_t packed = sorted_src[q];
    const int64_t i = (int64_t)(packed >> 2);
    const int ch = (int)(packed & 3u);
    if (n_out[i] == 1) {
      dst.pos[q] = src.pos[i];
      dst.dir[q] = src.dir[i];
      dst.meta[q] = src.meta[i];
      dst.phot[q] = src.phot[i];
      const float* shi = reinterpret_cast<const float*>(src.phot + cap);
      float* dhi = reinterpret_cast<float*>(dst.phot + cap);
      for (int k = 0; k < nplanes; ++k) dhi[(int64_t)k * cap + q] = shi[(int64_t)k * cap + i];
      continue;
    }
    Ray par;
    ray_load(src, i, par);
    const int l = par.order;
    const float r0 = par.r;
    const ChildGeo c = child_geometry(par, ch, g, rot_fwd);
    Ray kid;
    kid.cell = c.cell;
    kid.pixel = c.pix;
    kid.order = l + 1;
    kid.itheta = par.itheta;
    kid.xhat = c.vx;
    kid.yhat = c.vy;
    kid.zhat = c.vz;
    kid.r = r0;
    kid.dr = g.rmax_tab[l + 1] - r0;
    kid.x = c.x;
    kid.y = c.y;
    kid.z = c.z;
    kid.time = par.time;
    ray_store(dst, q, kid);
    float n[8];
    phot_load(src, i, F, cap, n);
    FRT_FOR_BINS(nu) n[nu] = (float)((double)n[nu] / 4.);
    phot_store(dst, q, F, cap, n);
  }
}

// new_rays/init_rays (rt_funcs.cc:240-261, ray_tracing.cc:230-321), hpx_lvl 0:
// 12 rays per source cell, appended in source order.
__global__ void __launch_bounds__(BLK) k_cast(RaySoA dst, StepDev* sd, int nsrc, GridConst g,
                                              const int32_t* __restrict__ src_cell,
                                              const double* __restrict__ src_lum,
                                              float* __restrict__ src_ray_dt,
                                              const float* __restrict__ rot_fwd, uint64_t seed) {
  if (sd->err) return;
  const int64_t base = sd->nrays;
  const uint64_t stream = sd->cast_calls;
  if (base + (int64_t)nsrc * 12 > g.cap) {  // new_rays would overflow the buffer
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&sd->err, FRT_SDERR_CAST_CAPACITY);
    return;
  }
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)nsrc * 12) return;
  const int s = (int)(t / 12);
  const int pix = (int)(t - (int64_t)s * 12);
  const uint32_t it = frt_itheta_hash(seed, stream, (uint64_t)s);
  double px, py, pz;
  hpx_pix2vec_nest(0, pix, &px, &py, &pz);
  float vx, vy, vz;
  rotate3(rot_fwd + 9 * (int64_t)it, (float)px, (float)py, (float)pz, &vx, &vy, &vz);
  const int64_t q = base + t;
  const double lum = src_lum[s];
  const double rdt = (double)src_ray_dt[s];
  const int64_t cap = g.cap;
  const int F = g.nfreq;
  float n[8];
  FRT_FOR_BINS(nu) n[nu] = (float)((double)g.spectrum[nu] * lum * rdt / 12. / 1. / FRT_PHOT_NORM);
  phot_store(dst, q, F, cap, n);
  Ray ry;
  ry.cell = src_cell[s];
  ry.pixel = pix;
  ry.order = 0;
  ry.itheta = (int)it;
  ry.xhat = vx;
  ry.yhat = vy;
  ry.zhat = vz;
  ry.r = 0.f;
  ry.dr = g.rmax_tab[0] - 0.f;
  const float half = (float)((double)g.dx / 2.);
  ry.x = half;
  ry.y = half;
  ry.z = half;
  ry.time = 0.f;
  ray_store(dst, q, ry);
}

// After a cast: the sources' ray_dt restart from 0 and the list grows by
// the new rays (skipped after a capacity error).
__global__ void __launch_bounds__(BLK) k_cast_done(StepDev* sd, float* __restrict__ ray_dt, int nsrc) {
  if (sd->err) return;
  const int s = blockIdx.x * BLK + threadIdx.x;
  if (s < nsrc) ray_dt[s] = 0.f;
  if (s == 0) {
    sd->nrays += (int64_t)nsrc * 12;
    sd->cast_calls += 1;
  }
}

// Start of frt_step: time advanced (main.cc:123), counters and queue heads
// cleared, error bits reset.
__global__ void __launch_bounds__(BLK) k_step_begin(StepDev* sd, DevCounters* cnt, float dt, int advance) {
  constexpr int nw = FRT_CTR_COPIES * sizeof(DevCounters) / sizeof(unsigned long long);
  unsigned long long* w = (unsigned long long*)cnt;
  for (int q = threadIdx.x; q < nw; q += BLK) w[q] = 0ull;
  if (threadIdx.x < 16) sd->queue[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    if (advance) sd->t = sd->t + dt;
    sd->err = 0u;
    sd->pdf_on = 0;
    sd->merge_on = 0;
    sd->mcount = 0;
    sd->ncand = 0;
  }
}

__global__ void k_queue_reset(StepDev* sd) {
  if (threadIdx.x < 16 && blockIdx.x == 0) sd->queue[threadIdx.x] = 0;
}

__global__ void k_step_end(StepDev* sd) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !sd->err) sd->step += 1;
}

__global__ void __launch_bounds__(BLK) k_add_ray_dt(float* __restrict__ ray_dt, int nsrc, float dt) {
  const int s = blockIdx.x * BLK + threadIdx.x;
  if (s < nsrc) ray_dt[s] += dt;
}

__global__ void __launch_bounds__(BLK) k_copy(const int4* __restrict__ a, int4* __restrict__ b,
                                              int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n16; i += (int64_t)gridDim.x * BLK)
    b[i] = a[i];
}

// ===========================================================================
// Host-side launchers.

#define FRT_LAUNCH(ctx, name, grid, kern, ...)                                   \
  do {                                                                          \
    if ((grid) > 0) {                                                           \
      ProfRec _rec;                                                             \
      prof_begin(ctx, name, &_rec);                                             \
      hipLaunchKernelGGL(kern, dim3(grid), dim3(BLK), 0, (ctx)->stream, __VA_ARGS__); \
      prof_end(ctx, &_rec);                                                     \
      hipError_t _e = hipGetLastError();                                        \
      if (_e != hipSuccess) return hip_check(ctx, _e, name);                    \
    }                                                                           \
  } while (0)

static RaySoA& cur_buf(frt_ctx* ctx) { return ctx->buf[ctx->cur]; }

// Host-side sizes of the launches: the ray count known at the start of the
// entry point, or, while a step graph is captured, the ray-count bucket the
// graph serves (a power of two; replays with more rays recapture).  Every
// kernel reads its actual count from the step scalars and strides over it,
// so a size only shapes the grid.
static int64_t bound_rays(const frt_ctx* ctx) { return ctx->capturing ? ctx->capture_rays : ctx->nrays; }
static int64_t bound_segs(const frt_ctx* ctx) {
  return std::min<int64_t>(ctx->seg_cap, bound_rays(ctx) * (int64_t)ctx->g.P);
}

// Blocks of a grid-stride launch: one thread per item up to the size, at
// most grid_cap blocks.
static unsigned ray_grid(const frt_ctx* ctx, int64_t bound) {
  const int64_t g = (std::max<int64_t>(bound, 1) + BLK - 1) / BLK;
  return (unsigned)std::min<int64_t>(g, std::max(ctx->tune_grid_cap, 1));
}

// Persistent grid of the queue kernels.
static unsigned queue_grid(const frt_ctx* ctx) {
  return (unsigned)std::max(1, ctx->num_cu * std::max(ctx->tune_blocks_per_cu, 1));
}

int launch_add_ray_dt(frt_ctx* ctx) {
  if (ctx->nsrc > 0)
    FRT_LAUNCH(ctx, "add_ray_dt", grid_for(ctx->nsrc), k_add_ray_dt, ctx->src_ray_dt, ctx->nsrc, ctx->g.dt);
  return FRT_OK;
}

// Clears the counters, queue heads and error bits; with advance, also the
// time step of frt_step.
int launch_step_begin(frt_ctx* ctx, bool advance) {
  FRT_LAUNCH(ctx, "step_begin", 1, k_step_begin, ctx->sd, ctx->dcnt, ctx->g.dt, advance ? 1 : 0);
  return FRT_OK;
}

int launch_step_end(frt_ctx* ctx) {
  FRT_LAUNCH(ctx, "step_end", 1, k_step_end, ctx->sd);
  return FRT_OK;
}

int launch_march(frt_ctx* ctx) {
  ctx->adj_valid = false;  // new paths: a recorded adjoint tape no longer matches
  const bool q = ctx->tune_march_queue != 0;
  if (q) FRT_LAUNCH(ctx, "queue_reset", 1, k_queue_reset, ctx->sd);
  const unsigned grid = q ? queue_grid(ctx) : ray_grid(ctx, bound_rays(ctx));
  const int chunk = q ? ctx->tune_queue_chunk : 0;
  if (ctx->deferred) {  // the previous split's compaction, gathered by the march
    FRT_LAUNCH(ctx, "march", grid, k_march<true>, cur_buf(ctx), ctx->sd, ctx->g, ctx->path_cell, ctx->path_dist,
               ctx->seg_cnt, ctx->alive, ctx->psum, ctx->dcnt, chunk, ctx->buf[1 - ctx->cur], (const int32_t*)ctx->n_out,
               (const uint32_t*)ctx->split_sorted, (const float*)ctx->rot_fwd,
               (const float*)(ctx->subgrid ? ctx->sg.fion : nullptr));
    ctx->deferred = false;
  } else {
    FRT_LAUNCH(ctx, "march", grid, k_march<false>, cur_buf(ctx), ctx->sd, ctx->g, ctx->path_cell, ctx->path_dist,
               ctx->seg_cnt, ctx->alive, ctx->psum, ctx->dcnt, chunk, ctx->buf[1 - ctx->cur], (const int32_t*)nullptr,
               (const uint32_t*)nullptr, (const float*)ctx->rot_fwd,
               (const float*)(ctx->subgrid ? ctx->sg.fion : nullptr));
  }
  ctx->psum_state = FRT_PSUM_UNMOVED;
  return FRT_OK;
}

// The compaction a split left to the next march, run on its own: whatever
// reads or rewrites the rays before that march sees the compacted list.
int materialize_rays(frt_ctx* ctx) {
  if (!ctx->deferred) return FRT_OK;
  const int64_t nb = std::min<int64_t>(ctx->g.cap, 4 * ctx->nrays + 1);
  FRT_LAUNCH(ctx, "compact", ray_grid(ctx, nb), k_compact, ctx->buf[1 - ctx->cur], cur_buf(ctx), ctx->sd, ctx->g,
             ctx->n_out, (const uint32_t*)ctx->split_sorted, ctx->rot_fwd);
  ctx->deferred = false;
  ctx->psum_state = FRT_PSUM_NONE;
  return FRT_OK;
}

// Photon sums of all rays from their photons, where the march and the
// deposit passes have not left them (entry points called out of step order).
__global__ void __launch_bounds__(BLK) k_psum(RaySoA rb, const StepDev* __restrict__ sd, GridConst g,
                                              float* __restrict__ psum) {
  const int64_t nrays = sd->nrays;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < nrays; i += (int64_t)gridDim.x * BLK)
    psum[i] = sum_nphot(rb, i, g.nfreq, g.cap);
}

int ensure_psum(frt_ctx* ctx) {
  if (ctx->psum_state == FRT_PSUM_VALID) return FRT_OK;
  FRT_LAUNCH(ctx, "psum", std::min(ray_grid(ctx, bound_rays(ctx)), MAX_GRID), k_psum, cur_buf(ctx), ctx->sd,
             ctx->g, ctx->psum);
  ctx->psum_state = FRT_PSUM_VALID;
  return FRT_OK;
}

int launch_fill_alive(frt_ctx* ctx) {
  FRT_LAUNCH(ctx, "fill_alive", ray_grid(ctx, bound_rays(ctx)), k_fill_u8_dev, ctx->alive,
             (const int64_t*)&ctx->sd->nrays, (int64_t)0, (uint8_t)1);
  return FRT_OK;
}

static int grow_segments(frt_ctx* ctx, int64_t need) {
  if (need <= ctx->seg_cap) return FRT_OK;
  int64_t cap = ctx->seg_cap ? ctx->seg_cap : 1024;
  while (cap < need) cap += std::max<int64_t>(cap / 4, 1024);
  hipStreamSynchronize(ctx->stream);  // the stream may still use the old buffers
  graph_invalidate(ctx);              // captured graphs hold the old addresses
  for (int b = 0; b < 2; ++b) {
    hipFree(ctx->seg_key[b]);
    hipFree(ctx->seg_val[b]);
    ctx->seg_key[b] = nullptr;
    ctx->seg_val[b] = nullptr;
  }
  hipFree(ctx->contrib.rec);  // base of the record + pair block
  ctx->contrib.pair = nullptr;
  ctx->contrib.rec = nullptr;
  const int F = ctx->g.nfreq;
  const size_t acc = (ctx->cfg.precision == FRT_PREC_MIXED) ? 4 : 8;
  // last-pass photon record: F photon counts and the signed length, rounded
  // up to a multiple of 4 values (whole 16-byte stores)
  const int rec_len = (F + 4) & ~3;
  const size_t per = acc * (size_t)(2 + rec_len);
  hipError_t e = hipSuccess;
  for (int b = 0; b < 2 && e == hipSuccess; ++b) {
    e = hipMalloc((void**)&ctx->seg_key[b], (size_t)cap * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->seg_val[b], (size_t)cap * 4);
  }
  void* blk = nullptr;
  if (e == hipSuccess) e = hipMalloc(&blk, per * (size_t)cap);
  if (e != hipSuccess) {
    ctx->seg_cap = 0;
    return set_error(ctx, FRT_E_NOMEM, "segment buffers: " + std::string(hipGetErrorString(e)));
  }
  char* c = (char*)blk;
  ctx->contrib.rec = c;                        // cap x rec_len values
  ctx->contrib.pair = c + acc * rec_len * cap;  // cap x 2 values
  ctx->contrib.rec_len = rec_len;
  ctx->seg_cap = cap;
  return FRT_OK;
}

// Segment buffers for the paths of the current rays: at most path_length
// segments per ray, so a step never meets more than nrays * path_length
// (while capturing, the bound must hold for every replay: the graph is
// recaptured whenever the buffers grow).
int ensure_segment_capacity(frt_ctx* ctx) {
  return grow_segments(ctx, std::max<int64_t>(ctx->nrays * (int64_t)ctx->g.P, 1));
}

// Copy the counter copies and the step scalars to the host and sum the
// counters into hcnt[0].
int read_device_counters(frt_ctx* ctx) {
  int rc = hip_check(ctx, hipMemcpyAsync(ctx->hcnt, ctx->dcnt, FRT_CTR_COPIES * sizeof(DevCounters),
                                         hipMemcpyDeviceToHost, ctx->stream), "counters");
  if (rc) return rc;
  rc = hip_check(ctx, hipMemcpyAsync(ctx->hsd, ctx->sd, sizeof(StepDev), hipMemcpyDeviceToHost, ctx->stream),
                 "step scalars");
  if (rc) return rc;
  rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (rc) return rc;
  DevCounters* h = ctx->hcnt;
  for (int r = 1; r < FRT_CTR_COPIES; ++r) {
    h[0].moved += h[r].moved;
    h[0].dumped += h[r].dumped;
    h[0].terminated += h[r].terminated;
    h[0].alive += h[r].alive;
    h[0].merged_away += h[r].merged_away;
    h[0].groups += h[r].groups;
    h[0].splits += h[r].splits;
    h[0].substeps += h[r].substeps;
    h[0].path_overflow |= h[r].path_overflow;
  }
  return FRT_OK;
}

// Up-front sizing of the segment buffers and of the scan/sort workspace, so
// that the step loop does not reallocate (each reallocation waits for the
// stream) while the ray and segment counts grow: two segments per ray at full
// ray capacity, capped at 2^27 segments.  Larger counts grow the buffers
// before the step that could need them.
int reserve_workspace(frt_ctx* ctx) {
  const int64_t cap = ctx->g.cap;
  const int64_t segs = std::min<int64_t>(cap * std::min(ctx->cfg.path_length, 2), (int64_t)1 << 27);
  int rc = grow_segments(ctx, segs);
  if (rc) return rc;
  return alloc_scan_sort(ctx);
}

// Segment offsets, moved-ray list, cell sort and cell ranges of the current paths.
int launch_segments(frt_ctx* ctx) {
  const int64_t nb = bound_rays(ctx);
  StepDev* sd = ctx->sd;
  // segment offsets and, in the same scan, the list of moved rays (for the
  // deposit passes)
  int rc = dev_scan_moved(ctx, ctx->seg_cnt, ctx->seg_off, ctx->moved_idx, &sd->nrays, 1, nb + 1, &sd->nseg,
                          &sd->nmoved, "scan_segments");
  if (rc) return rc;
  FRT_LAUNCH(ctx, "seg_keys", ray_grid(ctx, nb), k_seg_keys, sd, ctx->g.cap, ctx->seg_cnt, ctx->seg_off,
             ctx->path_cell, ctx->seg_key[0], ctx->seg_val[0]);
  // the cell-sorted slot of every path entry and the cells' slot ranges
  return dev_sort_cells(ctx, ctx->seg_key, ctx->seg_val, &sd->nseg, bound_segs(ctx), ctx->key_bits, ctx->g.n3,
                        ctx->path_slot, ctx->cell_start, ctx->cell_end, "sort_segments");
}

__global__ void __launch_bounds__(BLK) k_tape_nph(RaySoA rb, int64_t n, int F, int64_t cap, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  float v[8];
  phot_load(rb, i, F, cap, v);
  FRT_FOR_BINS(nu) out[i * F + nu] = v[nu];
}

int launch_find_avg_unu(frt_ctx* ctx, bool fuse_thermal) {
  if (ctx->subgrid) return launch_find_avg_unu_subgrid(ctx, fuse_thermal);
  if (int rc0 = materialize_rays(ctx)) return rc0;
  const int I = ctx->cfg.itercount;
  const int64_t n = ctx->nrays;
  const bool mixed = ctx->cfg.precision == FRT_PREC_MIXED;
  const bool q = ctx->tune_deposit_queue != 0 && I <= 16 - FRT_Q_DEPOSIT;
  GridConst g = ctx->g;
  RaySoA rb = cur_buf(ctx);
  const unsigned gc = grid_for(g.n3);
  const int F = g.nfreq;
  StepDev* sd = ctx->sd;
  ctx->deposit_launches_last = 0;
  if (q) FRT_LAUNCH(ctx, "queue_reset", 1, k_queue_reset, sd);
  // adjoint tape: per pass [opacity HI, HeI, Gamma_HI, Gamma_HeI][n3], then
  // the starting f_H1, f_He1, and the rays' starting photons
  double* tape = ctx->adj_on ? ctx->adj_tape : nullptr;
  if (tape) {
    FRT_LAUNCH(ctx, "adj_tape", gc, k_tape_pair, ctx->c.f[C_FH1], ctx->c.f[C_FHE1], tape + (size_t)4 * I * g.n3,
               (int64_t)g.n3);
    FRT_LAUNCH(ctx, "adj_tape", grid_for(n), k_tape_nph, rb, n, F, g.cap, ctx->adj_nph);
    // the cell state the fused thermal update reads and then overwrites
    const int tf[5] = {C_TEMP, C_HEAT, C_COOL, C_NH1_PREV, C_NE};
    for (int qq = 0; qq < 5; ++qq) {
      hipError_t e = hipMemcpyAsync(ctx->adj_tstate + (size_t)qq * g.n3, ctx->c.f[tf[qq]], 4 * (size_t)g.n3,
                                    hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) return hip_check(ctx, e, "adjoint tape");
    }
    ctx->adj_cast_base = -1;
    ctx->adj_split_n = ctx->adj_split_total = -1;
  }
  const unsigned dgrid_mixed = q ? queue_grid(ctx) : ray_grid(ctx, bound_rays(ctx));
  const int fchunk = ctx->tune_fp64_chunk == 2048 ? 2048 : 4096;
  const unsigned dgrid_fp64 =
      q ? queue_grid(ctx) : (unsigned)std::max<int64_t>(1, (bound_rays(ctx) + fchunk - 1) / fchunk);
  for (int it = 0; it < I; ++it) {
    const bool first = (it == 0), last = (it == I - 1);
    g.pass = it;
    double* tk = tape ? tape + (size_t)4 * it * g.n3 : nullptr;
    uint8_t* tnp = tape ? ctx->adj_np + (size_t)it * g.cap : nullptr;
    int64_t* head = q ? &sd->queue[FRT_Q_DEPOSIT + it] : nullptr;
    const int chunk = ctx->tune_queue_chunk;
    if (tk)
      FRT_LAUNCH(ctx, "adj_tape", gc, k_tape_pair, ctx->c.f[first ? C_NH1 : C_AVG_NH1],
                 ctx->c.f[first ? C_NHE1 : C_AVG_NHE1], tk, (int64_t)g.n3);
    if (first)  // the first pass attenuates with n_HI, n_HeI themselves
      FRT_LAUNCH(ctx, "pack_avg", gc, k_pack_avg, ctx->c.f[C_NH1], ctx->c.f[C_NHE1], ctx->c.avg2, g.n3);
    // five-bin specializations for the FP64 form (198 -> 167 and 141 VGPRs,
    // 2 -> 3 waves per SIMD); the mixed kernels were not faster specialized
    // (first passes 71 -> 75 VGPRs; last pass 84 -> 69 but 3 % slower)
#define FRT_DEP_ARGS rb, ctx->moved_idx, sd, head, chunk, g, ctx->path_cell, ctx->path_dist, ctx->seg_cnt, \
                     ctx->path_slot, ctx->c.avg2, ctx->contrib, ctx->alive, ctx->dcnt, tnp, ctx->psum
    if (mixed) {
      if (last)
        FRT_LAUNCH(ctx, "deposit_mixed", dgrid_mixed, (k_deposit_mixed<true>), FRT_DEP_ARGS);
      else
        FRT_LAUNCH(ctx, "deposit_mixed", dgrid_mixed, (k_deposit_mixed<false>), FRT_DEP_ARGS);
    } else if (F == 5) {
      if (fchunk == 2048) {
        if (last)
          FRT_LAUNCH(ctx, "deposit_fp64", dgrid_fp64, (k_deposit_fp64<true, 5, 2048>), FRT_DEP_ARGS);
        else
          FRT_LAUNCH(ctx, "deposit_fp64", dgrid_fp64, (k_deposit_fp64<false, 5, 2048>), FRT_DEP_ARGS);
      } else {
        if (last)
          FRT_LAUNCH(ctx, "deposit_fp64", dgrid_fp64, (k_deposit_fp64<true, 5, 4096>), FRT_DEP_ARGS);
        else
          FRT_LAUNCH(ctx, "deposit_fp64", dgrid_fp64, (k_deposit_fp64<false, 5, 4096>), FRT_DEP_ARGS);
      }
    } else {
      if (last)
        FRT_LAUNCH(ctx, "deposit_fp64", dgrid_fp64, (k_deposit_fp64<true, 0, 4096>), FRT_DEP_ARGS);
      else
        FRT_LAUNCH(ctx, "deposit_fp64", dgrid_fp64, (k_deposit_fp64<false, 0, 4096>), FRT_DEP_ARGS);
    }
#undef FRT_DEP_ARGS
    ctx->deposit_launches_last += 1;
    // the last pass's record loop is specialized for five bins
#define FRT_CHEM(ACC, F1, L1, TH)                                                                 \
  do {                                                                                            \
    if (L1 && F == 5)                                                                             \
      FRT_LAUNCH(ctx, "chem", gc, (k_chem<ACC, F1, L1, TH, 5>), ctx->c, g, ctx->contrib,          \
                 ctx->cell_start, ctx->cell_end, ctx->dcnt, tk ? tk + 2 * g.n3 : nullptr, sd);    \
    else                                                                                          \
      FRT_LAUNCH(ctx, "chem", gc, (k_chem<ACC, F1, L1, TH>), ctx->c, g, ctx->contrib,             \
                 ctx->cell_start, ctx->cell_end, ctx->dcnt, tk ? tk + 2 * g.n3 : nullptr, sd);    \
  } while (0)
    if (mixed) {
      if (first && last) {
        if (fuse_thermal) FRT_CHEM(float, true, true, true); else FRT_CHEM(float, true, true, false);
      } else if (first) {
        FRT_CHEM(float, true, false, false);
      } else if (last) {
        if (fuse_thermal) FRT_CHEM(float, false, true, true); else FRT_CHEM(float, false, true, false);
      } else {
        FRT_CHEM(float, false, false, false);
      }
    } else {
      if (first && last) {
        if (fuse_thermal) FRT_CHEM(double, true, true, true); else FRT_CHEM(double, true, true, false);
      } else if (first) {
        FRT_CHEM(double, true, false, false);
      } else if (last) {
        if (fuse_thermal) FRT_CHEM(double, false, true, true); else FRT_CHEM(double, false, true, false);
      } else {
        FRT_CHEM(double, false, false, false);
      }
    }
#undef FRT_CHEM
  }
  if (tape) {
    ctx->adj_valid = true;
    ctx->adj_iters = I;
    ctx->adj_nrays = n;
  }
  // the last pass wrote the photon sums of the moved rays; the march's hold
  // for the others
  if (I > 0 && ctx->psum_state != FRT_PSUM_NONE) ctx->psum_state = FRT_PSUM_VALID;
  return FRT_OK;
}

// c2ray_iter = FALSE: the march's deposit and u_nu (after launch_segments).
int launch_deposit_inline(frt_ctx* ctx) {
  if (int rc0 = materialize_rays(ctx)) return rc0;
  GridConst g = ctx->g;
  FRT_LAUNCH(ctx, "deposit_inline", ray_grid(ctx, bound_rays(ctx)), k_deposit_inline, cur_buf(ctx), ctx->moved_idx,
             ctx->sd, g, ctx->path_cell, ctx->path_dist, ctx->seg_cnt, ctx->path_slot, ctx->c.f[C_NH1],
             ctx->contrib, ctx->psum);
  FRT_LAUNCH(ctx, "unu_sum", grid_for(g.n3), k_unu_sum, ctx->c, g, ctx->contrib, ctx->cell_start, ctx->cell_end);
  // the deposit wrote the photon sums of the moved rays; the march's hold for the others
  if (ctx->psum_state != FRT_PSUM_NONE) ctx->psum_state = FRT_PSUM_VALID;
  return FRT_OK;
}

// c2ray_iter = FALSE: update_gamma and update_chem (main.cc:186-191).
int launch_gamma_chem(frt_ctx* ctx) {
  FRT_LAUNCH(ctx, "gamma_chem", grid_for(ctx->g.n3), k_gamma_chem, ctx->c, ctx->g);
  return FRT_OK;
}

int launch_thermal(frt_ctx* ctx) {
  if (ctx->subgrid) return launch_thermal_subgrid(ctx);
  FRT_LAUNCH(ctx, "thermal", grid_for(ctx->g.n3), k_thermal, ctx->c, ctx->g, ctx->dcnt);
  return FRT_OK;
}

// count_rays, calc_nphot_merge (if needed) and merge, with every decision
// taken on the device.
int launch_merge(frt_ctx* ctx) {
  if (int rc0 = materialize_rays(ctx)) return rc0;
  const int64_t nb = bound_rays(ctx);
  GridConst g = ctx->g;
  RaySoA rb = cur_buf(ctx);
  StepDev* sd = ctx->sd;
  int rc;
  FRT_LAUNCH(ctx, "count_alive", std::min(ray_grid(ctx, nb), MAX_GRID), k_count_alive, ctx->alive, sd, ctx->dcnt);
  const int64_t nbins = 12ll << (2 * ctx->cfg.hpx_lvl);
  const int64_t num_thresh = (int64_t)g.n3 * (int64_t)(ctx->cfg.ray_per_cell_max - nbins);
  ctx->stats.num_thresh = num_thresh;
  // calc_nphot_merge's log-spaced photon-count bins (ray_tracing.cc:181-186)
  const double lg_max = dm_log10(ctx->nphot_max);
  const double lg_min = dm_log10(ctx->nphot_min);
  const double del = (lg_max - lg_min) / (10000 - 1);
  FRT_LAUNCH(ctx, "merge_setup", 1, k_merge_setup, sd, (const DevCounters*)ctx->dcnt, num_thresh, ctx->pdf);
  if ((rc = ensure_psum(ctx))) return rc;
  FRT_LAUNCH(ctx, "pdf", std::min(ray_grid(ctx, nb), PDF_GRID), k_pdf, (const float*)ctx->psum, sd, ctx->alive, lg_min,
             del, ctx->pdf);
  FRT_LAUNCH(ctx, "pdf_walk", 1, k_pdf_walk, sd, (const int32_t*)ctx->pdf, num_thresh, lg_min, del, ctx->cfg.seed);
  // merge(): rotation index drawn every pass (ray_tracing.cc:327)
  FRT_LAUNCH(ctx, "merge_flag", ray_grid(ctx, nb / 2 + 1), k_merge_flag, rb, sd, g, ctx->alive, (const float*)ctx->psum,
             ctx->rot_inv,
             ctx->cand_flag, ctx->cand_key_tmp);
  // the candidates in list order with their keys, for the sort
  rc = dev_compact(ctx, ctx->cand_flag, ctx->cand_key_tmp, ctx->m_val[0], ctx->m_key[0], &sd->mcount, 0, nb + 1,
                   &sd->ncand, "scan_merge");
  if (rc) return rc;
  int bits = 1;
  while ((1ll << bits) < (int64_t)g.n3 * nbins) ++bits;
  int which = 0;
  rc = dev_sort_pairs(ctx, ctx->m_key, ctx->m_val, &sd->ncand, 0, nb, bits, &which, "sort_merge");
  if (rc) return rc;
  FRT_LAUNCH(ctx, "merge_groups", std::min(ray_grid(ctx, nb), 4 * MAX_GRID), k_merge_groups, rb, sd, g,
             ctx->m_key[which], ctx->m_val[which],
             ctx->alive, ctx->rot_fwd, (float)ctx->cfg.safety_factor, ctx->dcnt, ctx->psum);
  return FRT_OK;
}

// split_rays into the other buffer (nosplit: compaction only).
int launch_split_compact(frt_ctx* ctx, bool nosplit) {
  if (int rc0 = materialize_rays(ctx)) return rc0;
  const int64_t nb = bound_rays(ctx);
  GridConst g = ctx->g;
  StepDev* sd = ctx->sd;
  if (!nosplit) {
    if (int rc0 = ensure_psum(ctx)) return rc0;
  }
  FRT_LAUNCH(ctx, "split_flag", std::min(ray_grid(ctx, nb + 1), MAX_GRID), k_split_flag, cur_buf(ctx), sd, g,
             ctx->alive, (const float*)ctx->psum, nosplit ? 1 : 0, ctx->n_out, ctx->dcnt);
  int rc = dev_scan(ctx, ctx->n_out, ctx->out_off, &sd->nrays, 1, nb + 1, &sd->nsplit, "scan_split");
  if (rc) return rc;
  FRT_LAUNCH(ctx, "split_check", 1, k_split_check, sd, g.cap, (int64_t)ctx->nsrc * 12);
  // stable sort of the output rays by cell tile: the list order of every step
  FRT_LAUNCH(ctx, "out_keys", ray_grid(ctx, nb), k_out_keys, cur_buf(ctx), sd, g, ctx->n_out, ctx->out_off,
             ctx->rot_fwd, ctx->m_key[0], ctx->m_val[0]);
  int which = 0;
  const int64_t nout = std::min<int64_t>(g.cap, 4 * nb);
  rc = dev_sort_pairs(ctx, ctx->m_key, ctx->m_val, &sd->ncompact, 0, nout, ctx->tile_bits, &which, "sort_rays");
  if (rc) return rc;
  ctx->split_sorted = ctx->m_val[which];
  if (nosplit) {
    // one thread per input ray; the grid-stride loop takes the split children beyond
    FRT_LAUNCH(ctx, "compact", ray_grid(ctx, nb), k_compact, cur_buf(ctx), ctx->buf[1 - ctx->cur], sd, g, ctx->n_out,
               (const uint32_t*)ctx->m_val[which], ctx->rot_fwd);
  } else {
    ctx->deferred = true;  // the next march gathers the list through the permutation
  }
  FRT_LAUNCH(ctx, "split_done", 1, k_split_done, sd);
  ctx->psum_state = FRT_PSUM_NONE;  // a new list order
  ctx->cur = 1 - ctx->cur;
  ctx->split_gen += 1;
  if (ctx->adj_on && !nosplit) {  // adjoint tape: the split's input count (its output count after the pull)
    ctx->adj_split_n = ctx->nrays;
    ctx->adj_split_total = -2;
    ctx->adj_split_gen = ctx->split_gen;
  }
  return FRT_OK;
}

int launch_cast(frt_ctx* ctx) {
  if (ctx->nsrc == 0) return FRT_OK;
  const int64_t nnew = (int64_t)ctx->nsrc * 12;
  FRT_LAUNCH(ctx, "cast", grid_for(nnew), k_cast, cur_buf(ctx), ctx->sd, ctx->nsrc, ctx->g, ctx->src_cell,
             ctx->src_lum, ctx->src_ray_dt, ctx->rot_fwd, ctx->cfg.seed);
  ctx->psum_state = FRT_PSUM_NONE;  // new rays without photon sums
  if (ctx->adj_on && ctx->adj_raydt && ctx->adj_raydt_cap >= ctx->nsrc) {  // adjoint tape: the ray_dt the cast used
    hipError_t e = hipMemcpyAsync(ctx->adj_raydt, ctx->src_ray_dt, 4 * (size_t)ctx->nsrc, hipMemcpyDeviceToDevice,
                                  ctx->stream);
    if (e != hipSuccess) return hip_check(ctx, e, "adjoint tape");
    ctx->adj_cast_base = -2;  // the list length before the cast, after the pull
  }
  FRT_LAUNCH(ctx, "cast_done", grid_for(ctx->nsrc), k_cast_done, ctx->sd, ctx->src_ray_dt, ctx->nsrc);
  return FRT_OK;
}

// Writes the host mirror of the step scalars to the device (a kernel with the
// values as arguments, so no host buffer stays in flight) and clears the
// error bits.
__global__ void k_set_scalars(StepDev* sd, int64_t nrays, float t, int64_t step, uint64_t cast_calls,
                              float nphot_merge) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sd->nrays = nrays;
  sd->t = t;
  sd->step = step;
  sd->cast_calls = cast_calls;
  sd->nphot_merge = nphot_merge;
  sd->err = 0u;
}

int push_step_scalars(frt_ctx* ctx) {
  FRT_LAUNCH(ctx, "set_scalars", 1, k_set_scalars, ctx->sd, ctx->nrays, ctx->t, ctx->step,
             (uint64_t)ctx->cast_calls, ctx->nphot_merge);
  return FRT_OK;
}

// ---------------------------------------------------------------------------
// Restart files (io_funcs.cc), packed on the GPU.

template <typename T>
__device__ __forceinline__ void put_le(uint8_t* dst, T v) {
  __builtin_memcpy(dst, &v, sizeof(T));  // byte stores: records are not aligned
}

// write_rays_binary (io_funcs.cc:274-303): one record of
// FRT_RAY_FILE_FIXED + 4 nfreq bytes per ray, fields in the order and widths
// of the reference's ray_type; the slot index is the list position and
// active is TRUE (every ray between steps is ready to march).
__global__ void __launch_bounds__(BLK) k_pack_ray_file(RaySoA rb, int64_t n, GridConst g, uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int F = g.nfreq;
  const int64_t rec = FRT_RAY_FILE_FIXED + 4 * F;
  Ray r;
  ray_load(rb, i, r);
  float nph[8];
  phot_load(rb, i, F, g.cap, nph);
  uint8_t* o = out + i * rec;
  put_le<int64_t>(o, i);
  put_le<uint8_t>(o + 8, (uint8_t)1);
  put_le<int16_t>(o + 9, (int16_t)r.order);
  put_le<int32_t>(o + 11, (int32_t)r.pixel);
  put_le<uint16_t>(o + 15, (uint16_t)r.itheta);
  put_le<float>(o + 17, r.xhat);
  put_le<float>(o + 21, r.yhat);
  put_le<float>(o + 25, r.zhat);
  FRT_FOR_BINS(nu) put_le<float>(o + 29 + 4 * nu, nph[nu]);
  uint8_t* q = o + 29 + 4 * F;
  put_le<float>(q, r.r);
  put_le<float>(q + 4, r.dr);
  put_le<float>(q + 8, r.x);
  put_le<float>(q + 12, r.y);
  put_le<float>(q + 16, r.z);
  put_le<float>(q + 20, r.time);
  put_le<int32_t>(q + 24, (int32_t)frt_cell_to_rowmajor(g.lay, r.cell));
}

// write_gas_binary (io_funcs.cc:358-393): per row-major cell FRT_GAS_FILE_FIXED
// floats and ngam_nu[nfreq]; out[0] is the time, written by the host.  The
// fields this library does not evolve come from ext ([7][n3], row-major) when
// bit j of have is set, else from their C2Ray-mode values.
__global__ void __launch_bounds__(BLK) k_pack_gas_file(CellPtrs cf, GridConst g, const float* __restrict__ ext,
                                                       unsigned have, int held_rho, float* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (r >= g.n3) return;
  const int F = g.nfreq;
  const int64_t n3 = g.n3;
  const int64_t k = frt_cell_from_rowmajor(g.lay, (int)r);
  float* o = out + 1 + r * (FRT_GAS_FILE_FIXED + F);
  auto extra = [&](int j, float dflt) { return (have >> j & 1u) ? ext[(int64_t)j * n3 + r] : dflt; };
  // rho: given, held, or nH m_H / (1 - Y), the inverse of init (init_funcs.cc:64)
  o[0] = extra(0, held_rho ? cf.f[C_RHO][k] : (float)((double)cf.f[C_NH][k] * FRT_M_H / (1. - FRT_Y_HE)));
  o[1] = cf.f[C_TEMP][k];
  o[2] = cf.f[C_NH1][k];
  o[3] = cf.f[C_FH1][k];
  o[4] = cf.f[C_FHE1][k];
  o[5] = cf.f[C_FH2][k];
  o[6] = cf.f[C_FHE2][k];
  o[7] = extra(1, cf.f[C_FHE3][k]);
  o[8] = cf.f[C_GAMMA_H1_TOT][k];
  o[9] = cf.f[C_GAMMA_HE1_TOT][k];
  o[10] = extra(2, cf.f[C_MFP][k]);
  o[11] = extra(3, cf.f[C_MFP912][k]);
  o[12] = cf.f[C_HEAT][k];
  o[13] = cf.f[C_COOL][k];
  o[14] = cf.f[C_AVG_NH1][k];
  o[15] = extra(4, 1.f);  // fion = 1 with subgrid off (rt_funcs.cc:50-57)
  o[16] = extra(5, 0.f);  // tcross
  o[17] = extra(6, cf.f[C_TREION][k]);  // treion_out is set with treion (ray_tracing.cc:1129-1132)
  FRT_FOR_BINS(nu) o[FRT_GAS_FILE_FIXED + nu] = cf.ngam_nu[(int64_t)nu * n3 + k];
}

// write_vIF_binary (io_funcs.cc:396-430): per row-major cell IF_speed,
// Inc_Flux, fion and Delta_fion; without I-front tracking fion = 1
// (rt_funcs.cc:55) and Delta_fion = 0 (ray_tracing.cc:645).
__global__ void __launch_bounds__(BLK) k_pack_vif_file(CellPtrs cf, GridConst g, float4* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (r >= g.n3) return;
  const int64_t k = frt_cell_from_rowmajor(g.lay, (int)r);
  out[r] = make_float4(cf.f[C_IF_SPEED][k], cf.f[C_INC_FLUX][k], 1.f, 0.f);
}

int launch_pack_vif_file(frt_ctx* ctx, float4* out) {
  FRT_LAUNCH(ctx, "pack_vif_file", grid_for(ctx->g.n3), k_pack_vif_file, ctx->c, ctx->g, out);
  return FRT_OK;
}

// ---- write_otf reductions (io_funcs.cc:500-657) ------------------------------
// Each thread accumulates FRT_OTF_NV values over a grid-stride range of
// cells: double sums (calc_vol_avg / calc_mass_avg, data_funcs.cc:75-112)
// and float extremes (maxarr from 0, minarr from 1e32, data_funcs.cc:22-51);
// the block reduces them with a fixed butterfly and writes one partial per
// value; k_otf_final combines the partials in a fixed order.  The result does
// not depend on the schedule; it differs from a serial double sum only by the
// order of the additions.
__device__ __forceinline__ bool otf_is_max(int i) { return i == FRT_OTF_NSUM || i == FRT_OTF_NSUM + 2; }
__device__ __forceinline__ bool otf_is_min(int i) { return i == FRT_OTF_NSUM + 1 || i == FRT_OTF_NSUM + 3; }

__device__ __forceinline__ void otf_block_store(double (&v)[FRT_OTF_NV], double* __restrict__ part) {
  __shared__ double red[BLK / 64][FRT_OTF_NV];
#pragma unroll
  for (int i = 0; i < FRT_OTF_NV; ++i) {
    double x = v[i];
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_xor(x, off, 64);
      x = otf_is_max(i) ? fmax(x, y) : otf_is_min(i) ? fmin(x, y) : x + y;
    }
    v[i] = x;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int i = 0; i < FRT_OTF_NV; ++i) red[w][i] = v[i];
  }
  __syncthreads();
  const int i = threadIdx.x;
  if (i < FRT_OTF_NV) {
    double x = red[0][i];
    for (int q = 1; q < BLK / 64; ++q)
      x = otf_is_max(i) ? fmax(x, red[q][i]) : otf_is_min(i) ? fmin(x, red[q][i]) : x + red[q][i];
    part[(int64_t)blockIdx.x * FRT_OTF_NV + i] = x;
  }
}

__global__ void __launch_bounds__(BLK) k_otf_cells(CellPtrs cf, int64_t n3, float avg_rho, float temp_0,
                                                   double* __restrict__ part) {
  double v[FRT_OTF_NV];
#pragma unroll
  for (int i = 0; i < FRT_OTF_NV; ++i) v[i] = 0.;
  v[FRT_OTF_NSUM + 1] = v[FRT_OTF_NSUM + 3] = (double)1e32f;
  for (int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x; k < n3; k += (int64_t)gridDim.x * BLK) {
    const double fH1 = cf.f[C_FH1][k], fH2 = cf.f[C_FH2][k], fHe1 = cf.f[C_FHE1][k];
    const double fHe2 = cf.f[C_FHE2][k], fHe3 = cf.f[C_FHE3][k];
    const double nH = cf.f[C_NH][k], nHe = cf.f[C_NHE][k], nH1 = cf.f[C_NH1][k], nH2 = cf.f[C_NH2][k];
    const float Gf = cf.f[C_GAMMA_H1_TOT][k], Gp = cf.f[C_GAMMA_PREV][k];
    const float Tf = cf.f[C_TEMP][k], mfpf = cf.f[C_MFP][k];
    const double G = Gf, rec = cf.f[C_RECRATE_H2][k], T = Tf, mfp = mfpf, mfp912 = cf.f[C_MFP912][k];
    const double heat = cf.f[C_HEAT][k], cool = cf.f[C_COOL][k];
    // binary (io_funcs.cc:528-533) with fion = 1
    const bool bin = (Gf >= 3e-14f) && (Gp >= 3e-14f) && (fabs(dm_log10(Gf) - dm_log10(Gp)) < 1.);
    const double b = bin ? 1. : 0.;
    // mean-density cells and their temperature (:535-549, subgrid off)
    const float rr = cf.f[C_RHO][k] / avg_rho;
    const bool rm = ((double)rr > 0.9) && ((double)rr < 1.1);
    const float fH2f = cf.f[C_FH2][k];
    const double temp0 = rm ? (double)(float)((double)(Tf * fH2f) + (double)temp_0 * (1. - (double)fH2f))
                            : (double)temp_0;
    const double r = rm ? 1. : 0.;
    v[0] += fH1; v[1] += fH2; v[2] += fHe1; v[3] += fHe2; v[4] += fHe3;
    v[5] += (double)cf.f[C_AVG_NH1][k]; v[6] += (double)cf.f[C_AVG_NHE1][k];
    v[7] += nH; v[8] += nHe; v[9] += (double)cf.f[C_NE][k];
    v[10] += G; v[11] += rec; v[12] += T; v[13] += mfp; v[14] += mfp912; v[15] += heat; v[16] += cool;
    v[17] += fH1 * nH; v[18] += fH2 * nH; v[19] += fHe1 * nHe; v[20] += fHe2 * nHe; v[21] += fHe3 * nHe;
    v[22] += b; v[23] += nH * b; v[24] += nHe * b; v[25] += G * b; v[26] += rec * b; v[27] += T * b;
    v[28] += mfp * b; v[29] += mfp912 * b; v[30] += heat * b; v[31] += cool * b;
    v[32] += G * nH; v[33] += nH1; v[34] += G * nH1; v[35] += rec * nH; v[36] += nH2; v[37] += rec * nH2;
    v[38] += T * nH; v[39] += r; v[40] += temp0 * r; v[41] += mfp * nH; v[42] += mfp912 * nH;
    v[43] += heat * nH; v[44] += cool * nH;
    if (Gf > (float)v[45]) v[45] = Gf;
    if (Gf < (float)v[46]) v[46] = Gf;
    if (mfpf > (float)v[47]) v[47] = mfpf;
    if (mfpf < (float)v[48]) v[48] = mfpf;
  }
  otf_block_store(v, part);
}

// count_photons_io (io_funcs.cc:18-31): per ray the float sum over bins
// (sum_nphot, :10-16), added in double.
__global__ void __launch_bounds__(BLK) k_otf_rays(RaySoA rb, int64_t n, int F, int64_t cap,
                                                  double* __restrict__ part) {
  __shared__ double red[BLK / 64];
  double acc = 0.;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    float nph[8];
    phot_load(rb, i, F, cap, nph);
    float s = 0.f;
    FRT_FOR_BINS(nu) s += nph[nu];
    acc += (double)s;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = red[0];
    for (int q = 1; q < BLK / 64; ++q) x += red[q];
    part[blockIdx.x] = x;
  }
}

// One block per value: each thread combines a fixed strided subset of the
// block partials in order, then a fixed butterfly and the four wave results
// in order (the value FRT_OTF_NV is the photon count of k_otf_rays).
__global__ void __launch_bounds__(BLK) k_otf_final(const double* __restrict__ part, int nb, const double* __restrict__ rpart,
                                                   int nbr, double* __restrict__ out) {
  __shared__ double red[BLK / 64];
  const int i = blockIdx.x;
  const bool rays = i == FRT_OTF_NV;
  const bool mx = otf_is_max(i), mn = otf_is_min(i);
  const int n = rays ? nbr : nb;
  double x = mn ? (double)1e32f : 0.;
  for (int b = threadIdx.x; b < n; b += BLK) {
    const double y = rays ? rpart[b] : part[(int64_t)b * FRT_OTF_NV + i];
    x = mx ? fmax(x, y) : mn ? fmin(x, y) : x + y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double y = __shfl_xor(x, off, 64);
    x = mx ? fmax(x, y) : mn ? fmin(x, y) : x + y;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int q = 1; q < BLK / 64; ++q) t = mx ? fmax(t, red[q]) : mn ? fmin(t, red[q]) : t + red[q];
    out[i] = t;
  }
}

int launch_otf(frt_ctx* ctx, float avg_rho, float temp_0, double* part, double* out) {
  if (int rc0 = materialize_rays(ctx)) return rc0;
  const unsigned nb = (unsigned)std::min<int64_t>((ctx->g.n3 + BLK - 1) / BLK, FRT_OTF_BLOCKS);
  const unsigned nbr = (unsigned)std::max<int64_t>(1, std::min<int64_t>((ctx->nrays + BLK - 1) / BLK, FRT_OTF_BLOCKS));
  double* rpart = part + (size_t)FRT_OTF_BLOCKS * FRT_OTF_NV;
  FRT_LAUNCH(ctx, "otf_cells", nb, k_otf_cells, ctx->c, (int64_t)ctx->g.n3, avg_rho, temp_0, part);
  FRT_LAUNCH(ctx, "otf_rays", nbr, k_otf_rays, cur_buf(ctx), ctx->nrays, ctx->g.nfreq, ctx->g.cap, rpart);
  FRT_LAUNCH(ctx, "otf_final", FRT_OTF_NV + 1, k_otf_final, part, (int)nb, rpart, (int)nbr, out);
  return FRT_OK;
}

// MP-Gadget coupling (io_funcs.cc:434-496).  mode 0: update_gadget_var's
// sums (float, gadget_ini_heat = 0); mode 1: write_gadget_binary's division
// by the step time, in place, and the two row-major cubes into out; mode 2:
// reset_gadget_var.
__global__ void __launch_bounds__(BLK) k_gadget(CellPtrs cf, GridConst g, int mode, float step_time,
                                                float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= g.n3) return;
  float gg = cf.f[C_GADGET_GAMMA][k], gh = cf.f[C_GADGET_HEAT][k];
  if (mode == 0) {
    gg += cf.f[C_GAMMA_H1_TOT][k] * g.dt;
    gh += 0.f + cf.f[C_HEAT][k] * g.dt / cf.f[C_RHO][k];
  } else if (mode == 1) {
    gg /= step_time;
    gh /= step_time;
    const int64_t r = frt_cell_to_rowmajor(g.lay, (int)k);
    out[r] = gg;
    out[g.n3 + r] = gh;
  } else {
    gg = 0.f;
    gh = 0.f;
  }
  cf.f[C_GADGET_GAMMA][k] = gg;
  cf.f[C_GADGET_HEAT][k] = gh;
}

int launch_gadget(frt_ctx* ctx, int mode, float* out) {
  FRT_LAUNCH(ctx, "gadget", grid_for(ctx->g.n3), k_gadget, ctx->c, ctx->g, mode, ctx->gadget_step_time, out);
  return FRT_OK;
}

// update_cosmo_expansion's cell loop (gas_funcs.cc:483-512), and
// update_hydro_step's (:688-744) for a new density field, with fion = 1
// (rt_funcs.cc:55): rho scaled (by (a_prev/a_current)^3, or a_rt^-3 for the
// comoving density just read), the number densities recomputed from it and the
// ionized fractions, in the reference's float / double promotions.  The cell
// order does not matter here.
__global__ void __launch_bounds__(BLK) k_cosmo_cells(CellPtrs cf, int64_t n3, double rho_scale) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= n3) return;
  const float rho = (float)((double)cf.f[C_RHO][k] * rho_scale);
  const float nH = (float)((1. - FRT_Y_HE) * (double)rho / FRT_M_H);
  const float nHe = (float)(FRT_Y_HE * (double)rho / (4. * FRT_M_H));
  const float nH2 = cf.f[C_FH2][k] * nH;  // / fion, fion = 1
  const float nHe2 = cf.f[C_FHE2][k] * nHe;
  const float nHe3 = cf.f[C_FHE3][k] * nHe;
  const float nH1 = nH - nH2;
  const float nHe1 = nHe - nHe2 - nHe3;
  const float ne = (float)((double)(nH2 + nHe2) + 2. * (double)nHe3);
  cf.f[C_RHO][k] = rho;
  cf.f[C_NH][k] = nH;
  cf.f[C_NHE][k] = nHe;
  cf.f[C_NH2][k] = nH2;
  cf.f[C_NHE2][k] = nHe2;
  cf.f[C_NH1][k] = nH1;
  cf.f[C_AVG_NH1][k] = nH1;
  cf.f[C_NHE1][k] = nHe1;
  cf.f[C_AVG_NHE1][k] = nHe1;
  cf.f[C_NE][k] = ne;
  cf.f[C_NTOT][k] = nH + nHe + ne;
  cf.f[C_NH1_PREV][k] = nH1;
  cf.avg2[k] = make_float2(nH1, nHe1);
}

// Ray loop of update_cosmo_expansion (gas_funcs.cc:519-550): positions within
// the cell scaled by a_current/a_prev and moved off the faces, with g.dx
// already the new cell size.
__device__ __forceinline__ float cosmo_clamp(float x, float dx, float min_dist) {
  if (x <= 0.f) x = min_dist * dx;
  if ((double)x >= (double)dx * (1. - (double)min_dist / 2.)) x = (float)((double)dx * (1. - (double)min_dist));
  return x;
}

__global__ void __launch_bounds__(BLK) k_cosmo_rays(float4* __restrict__ pos, int64_t n, float ratio, float dx,
                                                    float min_dist) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  p.x = cosmo_clamp(p.x * ratio, dx, min_dist);
  p.y = cosmo_clamp(p.y * ratio, dx, min_dist);
  p.z = cosmo_clamp(p.z * ratio, dx, min_dist);
  pos[i] = p;
}

int launch_cosmo_expansion(frt_ctx* ctx, float ratio, double rho_scale, bool rays) {
  if (rays)
    if (int rc0 = materialize_rays(ctx)) return rc0;
  FRT_LAUNCH(ctx, "cosmo_cells", grid_for(ctx->g.n3), k_cosmo_cells, ctx->c, (int64_t)ctx->g.n3, rho_scale);
  if (rays && ctx->nrays > 0)
    FRT_LAUNCH(ctx, "cosmo_rays", grid_for(ctx->nrays), k_cosmo_rays, cur_buf(ctx).pos, ctx->nrays, ratio,
               ctx->g.dx, ctx->g.min_dist);
  return FRT_OK;
}

// calc_vol_avg (data_funcs.cc:75-92) of a field over the row-major cells in
// a fixed order: thread t of one block sums rows t, t + VAVG_THREADS, ... in
// double, then the partial sums are added pairwise (stride halving); the
// result is (float)((double)(float)sum / n^3).  The CPU checker sums in the
// same order, so the two agree bit for bit.
static constexpr int VAVG_THREADS = 1024;

__global__ void __launch_bounds__(VAVG_THREADS) k_vol_avg(const float* __restrict__ f, FrtCellLayout lay,
                                                          int64_t n3, float* __restrict__ out) {
  __shared__ double part[VAVG_THREADS];
  double t = 0.;
  for (int64_t r = threadIdx.x; r < n3; r += VAVG_THREADS) t += (double)f[frt_cell_from_rowmajor(lay, (int)r)];
  part[threadIdx.x] = t;
  __syncthreads();
  for (int h = VAVG_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (float)((double)(float)part[0] / (double)n3);
}

// ddelta_dt and rho_prev of update_hydro_step (gas_funcs.cc:732-737), after
// rho was rescaled: float arithmetic as in the reference.
__global__ void __launch_bounds__(BLK) k_hydro_ddelta(CellPtrs cf, int64_t n3, float avg_rho,
                                                      const float* __restrict__ avg_rho_prev, float dt_cosmic) {
  const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (k >= n3) return;
  const float rho = cf.f[C_RHO][k];
  const float arp = *avg_rho_prev;
  if (arp > 0.f) cf.f[C_DDELTA_DT][k] = (rho / avg_rho - cf.f[C_RHO_PREV][k] / arp) / dt_cosmic;
  cf.f[C_RHO_PREV][k] = rho;
}

int launch_hydro_reload(frt_ctx* ctx, double rho_scale, float dt_cosmic) {
  float* arp = (float*)ctx->dsmall;
  {
    ProfRec rec;
    prof_begin(ctx, "vol_avg", &rec);
    hipLaunchKernelGGL(k_vol_avg, dim3(1), dim3(VAVG_THREADS), 0, ctx->stream, ctx->c.f[C_RHO_PREV], ctx->g.lay,
                       (int64_t)ctx->g.n3, arp);
    prof_end(ctx, &rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_check(ctx, e, "vol_avg");
  }
  FRT_LAUNCH(ctx, "cosmo_cells", grid_for(ctx->g.n3), k_cosmo_cells, ctx->c, (int64_t)ctx->g.n3, rho_scale);
  FRT_LAUNCH(ctx, "hydro_ddelta", grid_for(ctx->g.n3), k_hydro_ddelta, ctx->c, (int64_t)ctx->g.n3, ctx->g.avg_rho,
             (const float*)arp, dt_cosmic);
  return FRT_OK;
}

int launch_pack_ray_file(frt_ctx* ctx, uint8_t* out) {
  if (int rc0 = materialize_rays(ctx)) return rc0;
  FRT_LAUNCH(ctx, "pack_ray_file", grid_for(ctx->nrays), k_pack_ray_file, cur_buf(ctx), ctx->nrays, ctx->g, out);
  return FRT_OK;
}

int launch_pack_gas_file(frt_ctx* ctx, const float* ext, unsigned have, int held_rho, float* out) {
  FRT_LAUNCH(ctx, "pack_gas_file", grid_for(ctx->g.n3), k_pack_gas_file, ctx->c, ctx->g, ext, have, held_rho, out);
  return FRT_OK;
}

int launch_copy(frt_ctx* ctx, int64_t nbytes, int repeats) {
  const int64_t n16 = nbytes / 16;
  unsigned grid = (unsigned)std::min<int64_t>((n16 + BLK - 1) / BLK, 256 * 64);
  for (int r = 0; r < repeats; ++r)
    FRT_LAUNCH(ctx, "copy", grid, k_copy, (const int4*)ctx->copy_src, (int4*)ctx->copy_dst, n16);
  return FRT_OK;
}

// Rotation matrices of get_unit_vector for every rotation index
// (healpix_funcs.cc:21-39 angle grid, :112-133 matrices), float arithmetic.
void build_rotation_tables(float* fwd, float* inv) {
  for (int i = 0; i < FRT_NALPHA; ++i)
    for (int j = 0; j < FRT_NALPHA; ++j)
      for (int k = 0; k < FRT_NALPHA; ++k) {
        const int ind = i * FRT_NALPHA * FRT_NALPHA + j * FRT_NALPHA + k;
        const float a0 = (float)(2. * FRT_PI_REF / FRT_NALPHA * i);
        const float b0 = (float)(2. * FRT_PI_REF / FRT_NALPHA * j);
        const float g0 = (float)(2. * FRT_PI_REF / FRT_NALPHA * k);
        const float ca = dm_cosf(a0), sa = dm_sinf(a0);
        const float cb = dm_cosf(b0), sb = dm_sinf(b0);
        const float cg = dm_cosf(g0), sg = dm_sinf(g0);
        float* f = fwd + 9 * ind;
        f[0] = cb * cg;
        f[1] = -cb * sg;
        f[2] = sb;
        f[3] = sa * sb * cg + ca * sg;
        f[4] = -sa * sb * sg + ca * cg;
        f[5] = -sa * cb;
        f[6] = -ca * sb * cg + sa * sg;
        f[7] = ca * sb * sg + sa * cg;
        f[8] = ca * cb;
        float* v = inv + 9 * ind;
        v[0] = cg * cb;
        v[1] = sg * ca + cg * sb * sa;
        v[2] = sg * sa - cg * sb * ca;
        v[3] = -sg * cb;
        v[4] = cg * ca - sg * sb * sa;
        v[5] = cg * sa + sg * sb * ca;
        v[6] = sb;
        v[7] = -cb * sa;
        v[8] = cb * ca;
      }
}

#include "frt_subgrid.h"

}  // namespace frt

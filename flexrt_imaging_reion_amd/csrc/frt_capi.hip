// frt_capi.hip -- context management and the extern "C" entry points.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "frt_internal.h"
#include "frt_order.h"

namespace frt {

int set_error(frt_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_check(frt_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return FRT_OK;
  return set_error(ctx, FRT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Host <-> device copies run on the library's stream (created non-blocking,
// so it is not ordered with the null stream) and wait for it: a copy is then
// ordered after the kernels queued before it and before those queued after.
hipError_t stream_copy(frt_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e;
}

hipError_t stream_zero(frt_ctx* ctx, void* dst, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(dst, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e;
}

int ensure_cub_tmp(frt_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->cub_tmp_bytes) return FRT_OK;
  hipStreamSynchronize(ctx->stream);  // earlier scans/sorts may still use the workspace
  hipFree(ctx->cub_tmp);
  ctx->cub_tmp = nullptr;
  size_t want = std::max(bytes, ctx->cub_tmp_bytes * 2);
  hipError_t e = hipMalloc(&ctx->cub_tmp, want);
  if (e != hipSuccess) {
    ctx->cub_tmp_bytes = 0;
    return set_error(ctx, FRT_E_NOMEM, "scan/sort workspace");
  }
  ctx->cub_tmp_bytes = want;
  return FRT_OK;
}

void prof_begin(frt_ctx* ctx, const char* name, ProfRec* rec) {
  rec->name = nullptr;
  if (!ctx->prof) return;
  if (ctx->event_pool.size() < 2) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t ev;
      if (hipEventCreate(&ev) == hipSuccess) ctx->event_pool.push_back(ev);
    }
  }
  if (ctx->event_pool.size() < 2) return;
  rec->a = ctx->event_pool.back();
  ctx->event_pool.pop_back();
  rec->b = ctx->event_pool.back();
  ctx->event_pool.pop_back();
  rec->name = name;
  hipEventRecord(rec->a, ctx->stream);
}

void prof_end(frt_ctx* ctx, ProfRec* rec) {
  if (!rec->name) return;
  hipEventRecord(rec->b, ctx->stream);
  ctx->prof_pending.push_back(*rec);
}

void prof_resolve(frt_ctx* ctx) {
  for (auto& r : ctx->prof_pending) {
    hipEventSynchronize(r.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.a, r.b);
    auto& acc = ctx->prof_acc[r.name];
    acc.first += (double)ms;
    acc.second += 1;
    ctx->event_pool.push_back(r.a);
    ctx->event_pool.push_back(r.b);
  }
  ctx->prof_pending.clear();
}

// Drops the captured step graphs (they are captured again on the next step).
void graph_invalidate(frt_ctx* ctx) {
  for (int c = 0; c < 2; ++c) {
    if (ctx->graph_valid[c]) hipGraphExecDestroy(ctx->graph_exec[c]);
    ctx->graph_valid[c] = false;
  }
}

}  // namespace frt

using namespace frt;

// Device scratch of frt_otf_averages_noalloc: the per-block partial sums of
// the write_otf reductions and their combination, FRT_OTF_SCRATCH doubles
// (about 400 KiB), held by every device the library runs on until the process
// ends.  Host code never uses the name `scratch` as an address: the device
// address of each GPU's copy comes from hipGetSymbolAddress at frt_create and
// is kept per device.
#define FRT_OTF_SCRATCH ((size_t)FRT_OTF_BLOCKS * (FRT_OTF_NV + 1) + 64)
static __device__ __align__(256) double scratch[FRT_OTF_SCRATCH];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_address(frt_ctx* ctx, void** out) {
  *out = nullptr;
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return set_error(ctx, FRT_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[ctx->device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(ctx, FRT_E_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = ptr;
  }
  *out = g_scratch_ptr[ctx->device];
  return FRT_OK;
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

int alloc_rays(frt_ctx* ctx, RaySoA& b, int64_t cap, int nfreq) {
  hipError_t e = dalloc(&b.pos, cap);
  if (e == hipSuccess) e = dalloc(&b.dir, cap);
  if (e == hipSuccess) e = dalloc(&b.meta, cap);
  if (e == hipSuccess) e = hipMalloc((void**)&b.phot, std::max<size_t>(frt_phot_bytes(nfreq, cap), 16));
  if (e != hipSuccess) return set_error(ctx, FRT_E_NOMEM, std::string("ray buffer: ") + hipGetErrorString(e));
  return FRT_OK;
}

void free_rays(RaySoA& b) {
  hipFree(b.pos); hipFree(b.dir); hipFree(b.meta); hipFree(b.phot);
  memset(&b, 0, sizeof(b));
}

int read_counters(frt_ctx* ctx) {
  return read_device_counters(ctx);
}

// Clears the counters, queue heads and error bits on the device.
int reset_counters(frt_ctx* ctx) { return launch_step_begin(ctx, false); }

// Takes the device step scalars into the host mirror after a pull, resolves
// the adjoint tape's pending counts, and turns the device error bits into an
// error code.  A split that would overflow the buffer leaves the rays as
// they were before it (the compaction and everything after it are skipped).
int apply_pull(frt_ctx* ctx, bool ran_march) {
  const StepDev* h = ctx->hsd;
  ctx->nrays = h->nrays;
  ctx->nseg = h->nseg;
  ctx->nmoved = h->nmoved;
  ctx->nphot_merge = h->nphot_merge;
  ctx->t = h->t;
  ctx->step = h->step;
  ctx->cast_calls = h->cast_calls;
  if (ctx->adj_split_total == -2) ctx->adj_split_total = h->nsplit;
  if (ctx->adj_cast_base == -2) ctx->adj_cast_base = h->nrays - (int64_t)ctx->nsrc * 12;
  if (h->err & FRT_SDERR_SPLIT_CAPACITY) {
    ctx->cur = 1 - ctx->cur;  // the compaction did not run
    ctx->deferred = false;
    ctx->psum_state = FRT_PSUM_NONE;
    ctx->split_gen -= 1;
    ctx->adj_split_n = ctx->adj_split_total = -1;
    return set_error(ctx, FRT_E_CAPACITY,
                     "ray buffer full: " + std::to_string(h->nsplit + (int64_t)ctx->nsrc * 12) +
                         " rays needed, capacity " + std::to_string(ctx->g.cap));
  }
  if (h->err & FRT_SDERR_CAST_CAPACITY) {
    ctx->adj_cast_base = -1;
    return set_error(ctx, FRT_E_CAPACITY, "ray buffer full in new_rays");
  }
  if (ran_march && ctx->hcnt->path_overflow)
    return set_error(ctx, FRT_E_PATH_OVERFLOW,
                     "a ray crossed more than path_length cells in one step; raise path_length");
  return FRT_OK;
}

// Rays held after removals: drop dead rays (no splits) so the buffer holds
// live rays only.  Ends with a pull, so the host counts are current.
int ensure_compact(frt_ctx* ctx) {
  if (int rc0 = materialize_rays(ctx)) return rc0;
  if (!ctx->have_alive) return FRT_OK;
  int rc = launch_split_compact(ctx, true);
  ctx->have_alive = false;
  if (rc) return rc;
  if ((rc = read_counters(ctx))) return rc;
  return apply_pull(ctx, false);
}

int check_ready(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  if (!ctx->spectrum_set) return set_error(ctx, FRT_E_STATE, "frt_set_spectrum has not been called");
  return FRT_OK;
}

// update_rays: ray_dt, march, segment bookkeeping (launches only).
int do_update_rays(frt_ctx* ctx) {
  ctx->stats.rays_start = ctx->nrays;
  int rc;
  if ((rc = launch_add_ray_dt(ctx))) return rc;
  if ((rc = launch_march(ctx))) return rc;
  if ((rc = launch_segments(ctx))) return rc;
  // c2ray_iter = FALSE: the march deposits into u_nu (ray_tracing.cc:1329-1355,1443-1470)
  if (ctx->c2ray_off && (rc = launch_deposit_inline(ctx))) return rc;
  ctx->have_segments = true;
  ctx->have_alive = true;
  return FRT_OK;
}

int do_find_avg_unu(frt_ctx* ctx, bool fuse_thermal) {
  if (!ctx->have_segments) return set_error(ctx, FRT_E_STATE, "find_avg_unu needs update_rays first");
  ctx->g.nphot_min = ctx->nphot_min;
  int rc;
  if (ctx->c2ray_off) {
    // set_to_true only (main.cc:142-144); in a fused step the cell updates
    // of main.cc:186-193 run here, as merge and split touch no cell field
    if (fuse_thermal && ((rc = launch_gamma_chem(ctx)) || (rc = launch_thermal(ctx)))) return rc;
    ctx->have_segments = false;
    ctx->have_alive = true;
    return FRT_OK;
  }
  rc = launch_find_avg_unu(ctx, fuse_thermal);
  if (rc) return rc;
  ctx->have_segments = false;
  ctx->have_alive = true;
  return FRT_OK;
}

int fill_alive(frt_ctx* ctx) {
  if (ctx->have_alive) return FRT_OK;
  int rc = launch_fill_alive(ctx);
  if (rc) return rc;
  ctx->have_alive = true;
  return FRT_OK;
}

int do_merge(frt_ctx* ctx) {
  int rc = fill_alive(ctx);
  if (rc) return rc;
  ctx->g.nphot_min = ctx->nphot_min;
  return launch_merge(ctx);
}

int do_split(frt_ctx* ctx) {
  int rc = fill_alive(ctx);
  if (rc) return rc;
  rc = launch_split_compact(ctx, false);
  ctx->have_alive = false;
  return rc;
}

// Algorithmic bytes of the last find_avg_unu's deposit launches (the
// roofline model): per ray seg_cnt (4) + nphot read (4F); per segment path
// cell + dist (8) + slot (4) + avg_nh1/avg_nhe1 (8) + the 2-value deposit
// (first passes) or the (F+1)-value photon record (last pass); the last
// pass adds the ray write-back (4F + 4 + 1).
static void deposit_model(frt_ctx* ctx) {
  const int64_t I = ctx->deposit_launches_last;
  ctx->deposit_bytes_last = 0.;
  if (I <= 0) return;
  const int F = ctx->g.nfreq;
  const double acc = ctx->cfg.precision == FRT_PREC_MIXED ? 4. : 8.;
  const double n = (double)ctx->stats.rays_start, moved = (double)ctx->stats.rays_moved, nseg = (double)ctx->nseg;
  for (int64_t it = 0; it < I; ++it) {
    const bool last = it == I - 1;
    double b = n * 4. + moved * (4. * F) + nseg * 20.;
    b += nseg * (last ? (F + 1.) : 2.) * acc;
    if (last) b += moved * (4. * F + 5.);
    ctx->deposit_bytes_last += b;
  }
}

// Pull after an entry point: host mirror, errors, statistics.
int finish_stats(frt_ctx* ctx, bool ran_march = false) {
  int rc = read_counters(ctx);
  if (rc) return rc;
  rc = apply_pull(ctx, ran_march);
  DevCounters* h = ctx->hcnt;
  ctx->stats.rays_moved = (int64_t)h->moved;
  ctx->stats.segments = ctx->nseg;
  ctx->stats.rays_dumped = (int64_t)h->dumped;
  ctx->stats.rays_terminated = (int64_t)h->terminated;
  ctx->stats.rays_counted = (int64_t)h->alive;
  ctx->stats.merge_candidates = ctx->hsd->ncand;
  ctx->stats.merge_groups = (int64_t)h->groups;
  ctx->stats.rays_merged_away = (int64_t)h->merged_away;
  ctx->stats.rays_split = (int64_t)h->splits;
  ctx->stats.thermal_substeps = (int64_t)h->substeps;
  ctx->stats.rays_end = ctx->nrays;
  ctx->stats.nphot_min = ctx->nphot_min;
  ctx->stats.nphot_max = ctx->nphot_max;
  ctx->stats.nphot_merge = ctx->nphot_merge;
  ctx->stats.t = ctx->t;
  ctx->stats.step = ctx->step;
  deposit_model(ctx);
  if (ctx->prof) prof_resolve(ctx);
  return rc;
}

// ---- step graph -----------------------------------------------------------------
static void graph_key(const frt_ctx* ctx, GraphKey* k) {
  memset(k, 0, sizeof(*k));
  memcpy(&k->g, &ctx->g, sizeof(GridConst));
  k->nsrc = ctx->nsrc;
  k->ptrs[0] = ctx->src_cell;
  k->ptrs[1] = ctx->src_lum;
  k->ptrs[2] = ctx->src_ray_dt;
  k->ptrs[3] = ctx->seg_key[0];
  k->ptrs[4] = ctx->seg_val[0];
  k->ptrs[5] = ctx->contrib.rec;
  k->seg_cap = ctx->seg_cap;
  k->tune[0] = ctx->tune_deposit_queue;
  k->tune[1] = ctx->tune_march_queue;
  k->tune[2] = ctx->tune_queue_chunk;
  k->tune[3] = ctx->tune_blocks_per_cu;
  k->tune[4] = ctx->tune_grid_cap;
  k->tune[5] = ctx->tune_fp64_chunk;
  k->itercount = ctx->cfg.itercount;
  // ray-count bucket: the smallest power of two (at least 4096) holding the
  // rays at the start of the step, within the capacity
  int64_t b = 4096;
  while (b < ctx->nrays) b *= 2;
  k->capture_rays = std::min<int64_t>(b, ctx->g.cap);
  k->deferred = ctx->deferred ? 1 : 0;
}

// The launches of one frt_step after its step_begin.
static int step_body(frt_ctx* ctx) {
  int rc;
  if ((rc = do_update_rays(ctx))) return rc;
  if ((rc = do_find_avg_unu(ctx, true))) return rc;
  if ((rc = do_merge(ctx))) return rc;
  if ((rc = do_split(ctx))) return rc;
  if ((rc = launch_cast(ctx))) return rc;
  ctx->stats.rays_new = (int64_t)ctx->nsrc * 12;
  return launch_step_end(ctx);
}

static bool graph_eligible(const frt_ctx* ctx) {
  // (the subgrid mode's I-front rounds are sized by the host as they run)
  return ctx->graph_on && !ctx->prof && !ctx->adj_on && !ctx->have_alive && !ctx->have_segments && !ctx->subgrid;
}

// One frt_step as a replay of the graph captured for this ray-buffer parity
// (captured first if absent or stale).  Capture records the launches without
// running them; the host-side effects of the recorded calls are undone and
// then applied once per replay.
static int graph_step(frt_ctx* ctx) {
  const int c = ctx->cur;
  GraphKey key;
  graph_key(ctx, &key);
  if (ctx->graph_valid[c] && memcmp(&key, &ctx->graph_key[c], sizeof(GraphKey)) != 0) {
    hipGraphExecDestroy(ctx->graph_exec[c]);
    ctx->graph_valid[c] = false;
  }
  if (!ctx->graph_valid[c]) {
    const int64_t split_gen = ctx->split_gen;
    const frt_step_stats stats = ctx->stats;
    const bool deferred = ctx->deferred;
    int32_t* const split_sorted = ctx->split_sorted;
    hipGraph_t graph = nullptr;
    ctx->capturing = true;
    ctx->capture_rays = key.capture_rays;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    int rc = (e == hipSuccess) ? launch_step_begin(ctx, true) : hip_check(ctx, e, "begin capture");
    if (rc == FRT_OK) rc = step_body(ctx);
    if (e == hipSuccess) {
      const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
      if (rc == FRT_OK && e2 != hipSuccess) rc = hip_check(ctx, e2, "end capture");
    }
    ctx->capturing = false;
    ctx->cur = c;
    ctx->split_gen = split_gen;
    ctx->stats = stats;
    ctx->deferred_after_graph[c] = ctx->deferred;  // the captured step's effect, applied per replay
    ctx->sorted_after_graph[c] = ctx->split_sorted;
    ctx->deferred = deferred;
    ctx->split_sorted = split_sorted;
    ctx->have_alive = ctx->have_segments = false;
    if (rc == FRT_OK) {
      e = hipGraphInstantiate(&ctx->graph_exec[c], graph, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = hip_check(ctx, e, "graph instantiate");
    }
    if (graph) hipGraphDestroy(graph);
    if (rc) return rc;
    ctx->graph_valid[c] = true;
    ctx->graph_key[c] = key;
  }
  int rc = push_step_scalars(ctx);
  if (rc) return rc;
  ctx->stats.rays_start = ctx->nrays;
  rc = hip_check(ctx, hipGraphLaunch(ctx->graph_exec[c], ctx->stream), "graph launch");
  if (rc) return rc;
  ctx->cur = 1 - c;
  ctx->split_gen += 1;
  ctx->deferred = ctx->deferred_after_graph[c];
  ctx->split_sorted = ctx->sorted_after_graph[c];
  ctx->psum_state = FRT_PSUM_NONE;  // a step ends with the cast
  ctx->deposit_launches_last = ctx->cfg.itercount;
  ctx->stats.rays_new = (int64_t)ctx->nsrc * 12;
  return FRT_OK;
}

// The C ABI's ids 0..FRT_F_INC_FLUX are the internal scalar slots of the
// same number; the per-bin fields follow, then the later scalar fields.
float* field_ptr(frt_ctx* ctx, int field) {
  if (field >= 0 && field <= FRT_F_INC_FLUX) return ctx->c.f[field];
  switch (field) {
    case FRT_F_GAMMA_H1_NU: return ctx->c.gamma_h1_nu;
    case FRT_F_GAMMA_HE1_NU: return ctx->c.gamma_he1_nu;
    case FRT_F_NGAM_NU: return ctx->c.ngam_nu;
    case FRT_F_RHO: return ctx->c.f[C_RHO];
    case FRT_F_FHE3: return ctx->c.f[C_FHE3];
    case FRT_F_MFP: return ctx->c.f[C_MFP];
    case FRT_F_MFP_912: return ctx->c.f[C_MFP912];
    case FRT_F_GAMMA_PREV: return ctx->c.f[C_GAMMA_PREV];
    case FRT_F_GADGET_GAMMA: return ctx->c.f[C_GADGET_GAMMA];
    case FRT_F_GADGET_HEAT: return ctx->c.f[C_GADGET_HEAT];
    case FRT_F_DDELTA_DT: return ctx->c.f[C_DDELTA_DT];
    case FRT_F_RHO_PREV: return ctx->c.f[C_RHO_PREV];
    case FRT_F_NHE3: return ctx->c.f[C_NHE3];
    case FRT_F_NH2_PREV: return ctx->c.f[C_NH2_PREV];
    case FRT_F_NHE1_PREV: return ctx->c.f[C_NHE1_PREV];
    case FRT_F_NHE2_PREV: return ctx->c.f[C_NHE2_PREV];
    case FRT_F_NHE3_PREV: return ctx->c.f[C_NHE3_PREV];
    case FRT_F_RECOMB_HE3: return ctx->c.f[C_RECOMB_HE3];
    case FRT_F_U_NU: return ctx->c.u_nu;
    default: return nullptr;
  }
}

static bool per_bin(int field) {
  return field == FRT_F_GAMMA_H1_NU || field == FRT_F_GAMMA_HE1_NU || field == FRT_F_NGAM_NU || field == FRT_F_U_NU;
}

// Whole-file reads and writes of the restart files.
int write_file(frt_ctx* ctx, const char* path, const void* data, size_t bytes) {
  FILE* f = fopen(path, "wb");
  if (!f) return set_error(ctx, FRT_E_IO, std::string("cannot open ") + path + " for writing");
  const size_t w = bytes ? fwrite(data, 1, bytes, f) : 0;
  const int closed = fclose(f);
  if (w != bytes || closed != 0) return set_error(ctx, FRT_E_IO, std::string("short write to ") + path);
  return FRT_OK;
}

int read_file(frt_ctx* ctx, const char* path, std::vector<uint8_t>& data) {
  FILE* f = fopen(path, "rb");
  if (!f) return set_error(ctx, FRT_E_IO, std::string("cannot open ") + path);
  fseek(f, 0, SEEK_END);
  const long size = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (size < 0) {
    fclose(f);
    return set_error(ctx, FRT_E_IO, std::string("cannot size ") + path);
  }
  data.resize((size_t)size);
  const size_t r = size ? fread(data.data(), 1, (size_t)size, f) : 0;
  fclose(f);
  if (r != (size_t)size) return set_error(ctx, FRT_E_IO, std::string("short read of ") + path);
  return FRT_OK;
}

template <typename T>
T get_le(const uint8_t* p) {
  T v;
  memcpy(&v, p, sizeof(T));
  return v;
}

}  // namespace

static void otf_fill(const frt_ctx* ctx, const double* S, frt_otf* out);

extern "C" {

const char* frt_version(void) { return "flexrt-mi355x 0.1 (gfx950)"; }

int frt_config_default(frt_config* cfg, int32_t n) {
  if (!cfg || n <= 0) return FRT_E_ARG;
  memset(cfg, 0, sizeof(*cfg));
  cfg->n = n;
  cfg->nfreq = 1;
  cfg->path_length = 4;
  cfg->itercount = 5;
  cfg->precision = FRT_PREC_FP64;
  cfg->hpx_lvl = 0;
  cfg->l_max = 12;
  cfg->ray_per_cell_max = 22;
  cfg->rec_case_b = 1;
  cfg->temp_ev = 1;
  cfg->recomb_cool = 1;
  cfg->free_free = 1;
  cfg->compton = 1;
  cfg->frac_limit = 1;
  cfg->ray_capacity = 0;
  cfg->seed = 1;
  cfg->box_ckpc = 40 * 1e3 / 0.68;
  cfg->zz = 12.0038;
  cfg->a_rt = 1. / (1. + 12.0038);
  cfg->dt = 0.;  // 0: dt = dx / (cl_factor c) as update_dt (rt_funcs.cc:9-12)
  cfg->cl_factor = 0.3;
  cfg->nl_const = 0.5;
  cfg->safety_factor = 2.;
  cfg->min_dist = 1e-4;
  cfg->temp_min = 1e2;
  cfg->hh = 0.68;
  cfg->omega_m = 0.305147;
  cfg->omega_l = 0.6948;
  return FRT_OK;
}

// dx = (Lx0*a)/Nx*kpc_to_cm with float Lx0, a (gas_funcs.cc:592,710,
// init_funcs.cc:22-24), and what follows from it.
static void set_cell_size(GridConst& g, const frt_config& c, float a_rt) {
  const float Lx = (float)c.box_ckpc * a_rt;
  g.dx = (float)((double)(Lx / (float)c.n) * FRT_KPC_TO_CM);
  g.sf = g.dx / g.dx0;
  g.vol = (double)g.dx * (double)g.dx * (double)g.dx;
}

// Redshift-dependent constants of update_thermal (gas_funcs.cc:362-458).
static void set_redshift(GridConst& g, const frt_config& c, float zz) {
  g.zz = zz;
  g.hubble = frt_hubble((float)(1. / (1. + (double)zz)), c.hh, c.omega_m, c.omega_l);
  g.compton_coef = frt_compton_rate((double)zz);
  g.t_cmb = 2.726 * (double)(1.f + zz);
}

int frt_create(const frt_config* cfg_in, frt_ctx** out) {
  if (!cfg_in || !out) return FRT_E_ARG;
  *out = nullptr;
  const frt_config& c = *cfg_in;
  if (c.n < 2 || c.n > 1024 || c.nfreq < 1 || c.nfreq > FRT_MAX_FREQ || c.path_length < 1 ||
      c.itercount < 1 || c.l_max < 0 || c.l_max >= FRT_MAX_LEVEL - 1 || c.hpx_lvl != 0 ||
      (c.precision != FRT_PREC_FP64 && c.precision != FRT_PREC_MIXED) || !(c.a_rt > 0.) ||
      !(c.box_ckpc > 0.))
    return FRT_E_ARG;
  // the sub-grid model is built in the FP64 form with temperature evolution
  // (temp_ev = FALSE would take interp_temp, subgrid.cc:637-649)
  if (c.subgrid != 0 && (c.subgrid != 1 || c.precision != FRT_PREC_FP64 || !c.temp_ev)) return FRT_E_ARG;
  // the c2ray_iter = FALSE mode is built in the FP64 form, without the
  // sub-grid model (its march deposit ignores it, ray_tracing.cc:1336-1337)
  if (c.c2ray_off != 0 && (c.c2ray_off != 1 || c.precision != FRT_PREC_FP64 || c.subgrid != 0)) return FRT_E_ARG;
  const int64_t n3 = (int64_t)c.n * c.n * c.n;
  const int64_t cap = c.ray_capacity > 0 ? c.ray_capacity : 30 * n3;
  // segment indices are int32; split bookkeeping packs (ray << 2 | child) in 32 bits
  if (cap * (int64_t)c.path_length >= (1ll << 31) || cap >= (1ll << 30)) return FRT_E_ARG;

  frt_ctx* ctx = new frt_ctx();
  ctx->cfg = c;
  ctx->cur = 0;
  ctx->nrays = 0;
  hipGetDevice(&ctx->device);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return FRT_E_HIP;
  }

  GridConst& g = ctx->g;
  memset(&g, 0, sizeof(g));
  g.n = c.n;
  g.lay = frt_cell_layout(c.n);
  g.n3 = (int)n3;
  g.nfreq = c.nfreq;
  g.P = c.path_length;
  g.cap = cap;
  // dx0 = Lx0/Nx*kpc_to_cm and dx = (Lx0*a)/Nx*kpc_to_cm with float Lx0, a
  // (global_variables.cc:69, gas_funcs.cc:592,710, init_funcs.cc:22-24)
  const float Lx0 = (float)c.box_ckpc;
  const float a_rt = (float)c.a_rt;
  g.dx0 = (float)((double)(Lx0 / (float)c.n) * FRT_KPC_TO_CM);
  set_cell_size(g, c, a_rt);
  const float clf = (float)c.cl_factor;
  g.clight = (float)((double)clf * FRT_CL);
  g.nc = (float)c.nl_const;
  if (c.dt > 0.)
    g.dt = (float)c.dt;
  else
    g.dt = (float)((double)(1.0f * g.dx / clf) / FRT_CL);
  g.min_dist = (float)c.min_dist;
  g.temp_min = (float)c.temp_min;
  g.l_max = c.l_max;
  g.hpx_lvl = c.hpx_lvl;
  g.frac_limit = c.frac_limit;
  g.rec_case_b = c.rec_case_b;
  g.temp_ev = c.temp_ev;
  g.recomb_cool = c.recomb_cool;
  g.free_free = c.free_free;
  g.compton = c.compton;
  set_redshift(g, c, (float)c.zz);
  g.nhe_to_nh = (float)((FRT_Y_HE / 4.) / (1. - FRT_Y_HE));
  g.sig912 = frt_sigmapi_H1(13.6 / FRT_H_EV);  // mfp_912 (ray_tracing.cc:1080)
  g.log4 = dm_log(4.);
  const float safety = (float)c.safety_factor;
  for (int l = 0; l < FRT_MAX_LEVEL; ++l) {
    // rmax_from_l / rmin_from_l (ray_tracing.cc:66-79)
    g.rmax_tab[l] = (float)((double)g.dx0 * dm_pow(2, l) * dm_pow(3. / FRT_PI_REF / (double)safety, 0.5));
    g.rmin_tab[l] = (l == 0) ? 0.f : (float)((double)g.dx0 * dm_pow(2, l - 1) * dm_pow(3. / FRT_PI_REF / (double)safety, 0.5));
  }
  ctx->key_bits = 1;
  while ((1ll << ctx->key_bits) < n3) ctx->key_bits++;
  {
    const int64_t t = frt_tiles_per_side(c.n);
    ctx->tile_bits = 1;
    while ((1ll << ctx->tile_bits) < t * t * t) ctx->tile_bits++;
  }
  ctx->t = 0.f;
  ctx->nphot_merge = 0.f;
  ctx->nphot_min = 0.;
  ctx->nphot_max = 0.;
  ctx->step = 0;
  ctx->cast_calls = 0;
  ctx->spectrum_set = 0;
  memset(&ctx->stats, 0, sizeof(ctx->stats));
  ctx->prof = 0;
  ctx->seg_cap = 0;
  ctx->mcap = cap;
  ctx->have_segments = false;
  ctx->have_alive = false;
  ctx->have_rho = false;
  ctx->otf_buf = nullptr;
  ctx->otf_scratch = nullptr;
  ctx->gadget_step_time = 0.f;
  ctx->adj_on = ctx->adj_valid = false;
  ctx->adj_iters = 0;
  ctx->adj_nrays = 0;
  ctx->adj_tape = nullptr;
  ctx->adj_nph = nullptr;
  ctx->adj_np = nullptr;
  ctx->adj_tstate = nullptr;
  ctx->adj_raydt = nullptr;
  ctx->adj_raydt_cap = 0;
  ctx->split_gen = 0;
  ctx->adj_split_gen = -1;
  ctx->adj_cast_base = -1;
  ctx->adj_split_n = ctx->adj_split_total = -1;
  ctx->adjr_cell = ctx->adjr_ray = ctx->adjr_ray2 = ctx->adjr_lum = ctx->adjr_work = nullptr;
  ctx->adjr_inv = nullptr;
  ctx->adjr_nray = 0;
  ctx->adjr_step = 0;
  ctx->adjr_on = false;
  ctx->split_sorted = nullptr;
  ctx->deferred = false;
  ctx->psum_state = FRT_PSUM_NONE;
  ctx->graph_on = 0;
  ctx->capturing = false;
  ctx->graph_valid[0] = ctx->graph_valid[1] = false;
  // launch tuning defaults: one thread per ray, no queues
  ctx->tune_deposit_queue = 0;
  ctx->tune_march_queue = 0;
  ctx->tune_queue_chunk = 256;
  ctx->tune_blocks_per_cu = 8;
  ctx->tune_grid_cap = 1 << 20;
  ctx->tune_fp64_chunk = 4096;
  ctx->num_cu = 256;
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && ncu > 0)
      ctx->num_cu = ncu;
  }

  int rc = FRT_OK;
  for (int b = 0; b < 2 && rc == FRT_OK; ++b) rc = alloc_rays(ctx, ctx->buf[b], cap, c.nfreq);
  hipError_t e = hipSuccess;
  const int P = c.path_length;
  if (rc == FRT_OK) {
    if (e == hipSuccess) e = dalloc(&ctx->path_cell, (size_t)cap * P);
    if (e == hipSuccess) e = dalloc(&ctx->path_dist, (size_t)cap * P);
    if (e == hipSuccess) e = dalloc(&ctx->path_slot, (size_t)cap * P);
    if (e == hipSuccess) e = dalloc(&ctx->seg_cnt, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->seg_off, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->alive, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->psum, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->n_out, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->out_off, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->cand_flag, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->cand_key_tmp, cap + 1);
    if (e == hipSuccess) e = dalloc(&ctx->moved_idx, cap + 1);
    for (int b = 0; b < 2; ++b) {
      if (e == hipSuccess) e = dalloc(&ctx->m_key[b], cap + 1);
      if (e == hipSuccess) e = dalloc(&ctx->m_val[b], cap + 1);
    }
    if (e == hipSuccess) e = dalloc(&ctx->pdf, FRT_PDF_BINS);
    if (e == hipSuccess) e = dalloc(&ctx->cell_start, n3);
    if (e == hipSuccess) e = dalloc(&ctx->cell_end, n3);
    const size_t nf = (size_t)(C_NSCALAR + 3 * c.nfreq) * (size_t)n3;
    if (e == hipSuccess) e = dalloc(&ctx->cell_block, nf);
    if (e == hipSuccess) e = dalloc(&ctx->c.avg2, (size_t)n3);
    if (e == hipSuccess) e = dalloc(&ctx->c.u_nu, (size_t)c.nfreq * (size_t)n3);
    if (e == hipSuccess) e = dalloc(&ctx->rot_fwd, 9 * FRT_NTHETA);
    if (e == hipSuccess) e = dalloc(&ctx->rot_inv, 9 * FRT_NTHETA);
    if (e == hipSuccess) e = dalloc(&ctx->dcnt, FRT_CTR_COPIES);
    if (e == hipSuccess) e = dalloc(&ctx->dsmall, 64);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->hcnt, FRT_CTR_COPIES * sizeof(DevCounters), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->hscalar, 512, hipHostMallocDefault);
    if (e == hipSuccess) e = dalloc(&ctx->sd, 1);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->hsd, sizeof(StepDev), hipHostMallocDefault);
    if (e != hipSuccess) rc = set_error(ctx, FRT_E_NOMEM, std::string("allocation: ") + hipGetErrorString(e));
  }
  if (rc == FRT_OK) {
    for (int f = 0; f < C_NSCALAR; ++f) ctx->c.f[f] = ctx->cell_block + (size_t)f * n3;
    ctx->c.gamma_h1_nu = ctx->cell_block + (size_t)C_NSCALAR * n3;
    ctx->c.gamma_he1_nu = ctx->c.gamma_h1_nu + (size_t)c.nfreq * n3;
    ctx->c.ngam_nu = ctx->c.gamma_he1_nu + (size_t)c.nfreq * n3;
    e = hipMemsetAsync(ctx->cell_block, 0, sizeof(float) * (size_t)(C_NSCALAR + 3 * c.nfreq) * n3, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->c.u_nu, 0, sizeof(float) * (size_t)c.nfreq * n3, ctx->stream);
    // clumping factor defaults to const_clump = 1 (user_inputs.h:87, set_clump subgrid.cc:799)
    if (e == hipSuccess) {
      std::vector<float> ones((size_t)n3, 1.0f);
      e = stream_copy(ctx, ctx->c.f[C_CLUMP], ones.data(), sizeof(float) * n3, hipMemcpyHostToDevice);
    }
    std::vector<float> fwd(9 * FRT_NTHETA), inv(9 * FRT_NTHETA);
    build_rotation_tables(fwd.data(), inv.data());
    if (e == hipSuccess) e = stream_copy(ctx, ctx->rot_fwd, fwd.data(), 4 * fwd.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = stream_copy(ctx, ctx->rot_inv, inv.data(), 4 * inv.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->dcnt, 0, FRT_CTR_COPIES * sizeof(DevCounters), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->sd, 0, sizeof(StepDev), ctx->stream);
    if (e == hipSuccess) memset(ctx->hsd, 0, sizeof(StepDev));
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = set_error(ctx, FRT_E_HIP, std::string("init: ") + hipGetErrorString(e));
  }
  if (rc == FRT_OK) rc = reserve_workspace(ctx);
  if (rc == FRT_OK) rc = scratch_address(ctx, &ctx->otf_scratch);
  ctx->subgrid = c.subgrid == 1;
  ctx->c2ray_off = c.c2ray_off == 1;
  if (rc == FRT_OK && ctx->subgrid) rc = alloc_subgrid(ctx);
  if (rc != FRT_OK) {
    std::string msg = ctx->err;
    frt_destroy(ctx);
    fprintf(stderr, "frt_create: %s\n", msg.c_str());
    return rc;
  }
  *out = ctx;
  return FRT_OK;
}

int frt_destroy(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  hipStreamSynchronize(ctx->stream);
  graph_invalidate(ctx);
  free_scan_sort(ctx);
  free_subgrid(ctx);
  hipFree(ctx->sd);
  if (ctx->hsd) hipHostFree(ctx->hsd);
  for (int b = 0; b < 2; ++b) free_rays(ctx->buf[b]);
  hipFree(ctx->path_cell); hipFree(ctx->path_dist); hipFree(ctx->seg_cnt); hipFree(ctx->seg_off);
  hipFree(ctx->alive); hipFree(ctx->psum); hipFree(ctx->n_out); hipFree(ctx->out_off); hipFree(ctx->cand_flag);
  hipFree(ctx->cand_key_tmp); hipFree(ctx->moved_idx);
  for (int b = 0; b < 2; ++b) {
    hipFree(ctx->seg_key[b]); hipFree(ctx->seg_val[b]); hipFree(ctx->m_key[b]); hipFree(ctx->m_val[b]);
  }
  hipFree(ctx->contrib.rec); hipFree(ctx->path_slot); hipFree(ctx->pdf);  // rec is the block base
  hipFree(ctx->cell_start); hipFree(ctx->cell_end); hipFree(ctx->cell_block); hipFree(ctx->c.avg2);
  hipFree(ctx->c.u_nu);
  hipFree(ctx->otf_buf);
  hipFree(ctx->adj_tape);
  hipFree(ctx->adj_nph);
  hipFree(ctx->adj_np);
  hipFree(ctx->adj_tstate); hipFree(ctx->adj_raydt);
  hipFree(ctx->adjr_cell); hipFree(ctx->adjr_ray); hipFree(ctx->adjr_ray2); hipFree(ctx->adjr_lum);
  hipFree(ctx->adjr_work); hipFree(ctx->adjr_inv);
  hipFree(ctx->src_cell); hipFree(ctx->src_lum); hipFree(ctx->src_ray_dt);
  hipFree(ctx->rot_fwd); hipFree(ctx->rot_inv); hipFree(ctx->dcnt); hipFree(ctx->dsmall);
  if (ctx->hcnt) hipHostFree(ctx->hcnt);
  if (ctx->hscalar) hipHostFree(ctx->hscalar);
  hipFree(ctx->cub_tmp); hipFree(ctx->copy_src); hipFree(ctx->copy_dst);
  for (auto& r : ctx->prof_pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
  for (auto ev : ctx->event_pool) hipEventDestroy(ev);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return FRT_OK;
}

const char* frt_last_error(const frt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int frt_set_spectrum(frt_ctx* ctx, const double* nu_phot, const float* spectrum) {
  if (!ctx || !nu_phot || !spectrum) return FRT_E_ARG;
  for (int nu = 0; nu < ctx->g.nfreq; ++nu) {
    ctx->g.nu[nu] = nu_phot[nu];
    ctx->g.spectrum[nu] = spectrum[nu];
    ctx->g.sigH[nu] = frt_sigmapi_H1(nu_phot[nu]);
    ctx->g.sigHe[nu] = frt_sigmapi_He1(nu_phot[nu]);
    ctx->g.sigHe2[nu] = frt_sigmapi_He2(nu_phot[nu]);
    ctx->g.sigHf[nu] = (float)ctx->g.sigH[nu];
    ctx->g.sigHef[nu] = (float)ctx->g.sigHe[nu];
  }
  ctx->spectrum_set = 1;
  return FRT_OK;
}

int frt_set_sources(frt_ctx* ctx, int32_t nsrc, const int32_t* cell, const double* lum) {
  if (!ctx || nsrc < 0 || (nsrc > 0 && (!cell || !lum))) return FRT_E_ARG;
  for (int s = 0; s < nsrc; ++s)
    if (cell[s] < 0 || cell[s] >= ctx->g.n3) return set_error(ctx, FRT_E_ARG, "source cell out of range");
  hipStreamSynchronize(ctx->stream);  // the stream may still read the old arrays
  hipFree(ctx->src_cell); hipFree(ctx->src_lum); hipFree(ctx->src_ray_dt);
  ctx->src_cell = nullptr; ctx->src_lum = nullptr; ctx->src_ray_dt = nullptr;
  ctx->nsrc = nsrc;
  ctx->h_src_cell.assign(cell, cell + nsrc);
  ctx->h_src_lum.assign(lum, lum + nsrc);
  hipError_t e = dalloc(&ctx->src_cell, nsrc);
  if (e == hipSuccess) e = dalloc(&ctx->src_lum, nsrc);
  if (e == hipSuccess) e = dalloc(&ctx->src_ray_dt, nsrc);
  std::vector<int32_t> icell(cell, cell + nsrc);
  for (auto& c : icell) c = frt_cell_from_rowmajor(ctx->g.lay, c);
  if (e == hipSuccess && nsrc > 0) e = stream_copy(ctx, ctx->src_cell, icell.data(), 4 * (size_t)nsrc, hipMemcpyHostToDevice);
  if (e == hipSuccess && nsrc > 0) e = stream_copy(ctx, ctx->src_lum, lum, 8 * (size_t)nsrc, hipMemcpyHostToDevice);
  if (e == hipSuccess && nsrc > 0) e = stream_zero(ctx, ctx->src_ray_dt, 4 * (size_t)nsrc);
  return hip_check(ctx, e, "sources");
}

// Internal cell index of every row-major cell (the C ABI's numbering).
static void cell_permutation(const frt_ctx* ctx, std::vector<int32_t>& perm) {
  const int64_t n3 = ctx->g.n3;
  perm.resize((size_t)n3);
  for (int64_t r = 0; r < n3; ++r) perm[(size_t)r] = frt_cell_from_rowmajor(ctx->g.lay, (int)r);
}

int frt_upload_field(frt_ctx* ctx, int32_t field, const float* host) {
  if (!ctx || !host) return FRT_E_ARG;
  float* d = field_ptr(ctx, field);
  if (!d) return set_error(ctx, FRT_E_ARG, "unknown field");
  const size_t n3 = (size_t)ctx->g.n3;
  std::vector<int32_t> perm;
  cell_permutation(ctx, perm);
  const int F = per_bin(field) ? ctx->g.nfreq : 1;
  // per-bin fields: [cell][nfreq] row-major in, [nfreq][cell] internal on the device
  std::vector<float> t((size_t)F * n3);
  for (size_t r = 0; r < n3; ++r)
    for (int nu = 0; nu < F; ++nu) t[(size_t)nu * n3 + (size_t)perm[r]] = host[r * F + nu];
  const hipError_t e = stream_copy(ctx, d, t.data(), 4 * t.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && field == FRT_F_RHO) ctx->have_rho = true;
  return hip_check(ctx, e, "upload field");
}

int frt_download_field(frt_ctx* ctx, int32_t field, float* host) {
  if (!ctx || !host) return FRT_E_ARG;
  float* d = field_ptr(ctx, field);
  if (!d) return set_error(ctx, FRT_E_ARG, "unknown field");
  const size_t n3 = (size_t)ctx->g.n3;
  int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (rc) return rc;
  std::vector<int32_t> perm;
  cell_permutation(ctx, perm);
  const int F = per_bin(field) ? ctx->g.nfreq : 1;
  std::vector<float> t((size_t)F * n3);
  const hipError_t e = stream_copy(ctx, t.data(), d, 4 * t.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (size_t r = 0; r < n3; ++r)
      for (int nu = 0; nu < F; ++nu) host[r * F + nu] = t[(size_t)nu * n3 + (size_t)perm[r]];
  return hip_check(ctx, e, "download field");
}

int frt_set_time(frt_ctx* ctx, float t) {
  if (!ctx) return FRT_E_ARG;
  ctx->t = t;
  return FRT_OK;
}

int frt_set_step(frt_ctx* ctx, int64_t step) {
  if (!ctx || step < 0) return FRT_E_ARG;
  ctx->step = step;
  ctx->cast_calls = (uint64_t)step + 1;
  return FRT_OK;
}

int frt_set_opacity_dither(frt_ctx* ctx, uint64_t seed) {
  if (!ctx) return FRT_E_ARG;
  ctx->g.dither_seed = seed;
  return FRT_OK;
}

int frt_set_dt(frt_ctx* ctx, float dt) {
  if (!ctx || !(dt > 0.f)) return FRT_E_ARG;
  ctx->g.dt = dt;
  return FRT_OK;
}

int frt_get_cell_size(const frt_ctx* ctx, float* dx) {
  if (!ctx || !dx) return FRT_E_ARG;
  *dx = ctx->g.dx;
  return FRT_OK;
}

int frt_update_cosmo_expansion(frt_ctx* ctx, float a_current, float a_prev) {
  if (!ctx) return FRT_E_ARG;
  if (!(a_current > 0.f) || !(a_prev > 0.f)) return set_error(ctx, FRT_E_ARG, "scale factors must be positive");
  if (!ctx->have_rho) return set_error(ctx, FRT_E_ARG, "update_cosmo_expansion needs rho: upload FRT_F_RHO first");
  // a compaction left to the next march runs first: split children take the
  // cell size of the step that split them
  if (int rc0 = materialize_rays(ctx)) return rc0;
  // dx[i] *= a_current/a_prev (float); rho *= pow(a_prev/a_current, 3.), the
  // scalar computed on the host with the C library's pow, as the reference does
  const float ratio = a_current / a_prev;
  const double rho_scale = std::pow((double)(a_prev / a_current), 3.);
  ctx->g.dx = ctx->g.dx * ratio;
  ctx->g.sf = ctx->g.dx / ctx->g.dx0;
  ctx->g.vol = (double)ctx->g.dx * (double)ctx->g.dx * (double)ctx->g.dx;
  ctx->g.avg_rho = (float)((double)ctx->g.avg_rho * rho_scale);  // avg_rho *= pow(a_prev/a_current, 3.) (:515)
  int rc = launch_cosmo_expansion(ctx, ratio, rho_scale, true);
  if (rc) return rc;
  return hip_check(ctx, hipGetLastError(), "update_cosmo_expansion");
}

int frt_update_hydro_step(frt_ctx* ctx, float zz, float a_rt, float avg_rho, float dt_cosmic) {
  if (!ctx) return FRT_E_ARG;
  if (!(a_rt > 0.f) || !(zz > -1.f)) return set_error(ctx, FRT_E_ARG, "need a_rt > 0 and zz > -1");
  if (!(avg_rho > 0.f) || !std::isfinite(dt_cosmic) || dt_cosmic == 0.f)
    return set_error(ctx, FRT_E_ARG, "need avg_rho > 0 and a finite, nonzero dt_cosmic");
  if (!ctx->have_rho) return set_error(ctx, FRT_E_ARG, "update_hydro_step needs rho: upload FRT_F_RHO first");
  if (int rc0 = materialize_rays(ctx)) return rc0;  // children with the cell size of their split
  ctx->cfg.zz = zz;
  ctx->cfg.a_rt = a_rt;
  set_redshift(ctx->g, ctx->cfg, zz);
  set_cell_size(ctx->g, ctx->cfg, a_rt);  // Lx = Lx0*a_rt, dx = Lx/Nx*kpc_to_cm
  ctx->g.avg_rho = avg_rho;
  int rc = launch_hydro_reload(ctx, std::pow((double)a_rt, -3.), dt_cosmic);
  if (rc) return rc;
  return hip_check(ctx, hipGetLastError(), "update_hydro_step");
}

int frt_otf_averages(frt_ctx* ctx, float avg_rho, float temp_0, frt_otf* out) {
  if (!ctx || !out) return FRT_E_ARG;
  int rc = ensure_compact(ctx);  // count_photons_io counts the rays in the buffer
  if (rc) return rc;
  const size_t nd = (size_t)FRT_OTF_BLOCKS * (FRT_OTF_NV + 1) + 64;
  if (!ctx->otf_buf && hipMalloc((void**)&ctx->otf_buf, nd * sizeof(double)) != hipSuccess) {
    ctx->otf_buf = nullptr;
    return set_error(ctx, FRT_E_NOMEM, "otf buffer");
  }
  double* dout = ctx->otf_buf + (size_t)FRT_OTF_BLOCKS * (FRT_OTF_NV + 1);
  rc = launch_otf(ctx, avg_rho, temp_0, ctx->otf_buf, dout);
  if (rc) return rc;
  double S[FRT_OTF_NV + 1];
  rc = hip_check(ctx, stream_copy(ctx, S, dout, sizeof(S), hipMemcpyDeviceToHost), "otf");
  if (rc) return rc;
  otf_fill(ctx, S, out);
  return FRT_OK;
}

int frt_otf_averages_noalloc(frt_ctx* ctx, float avg_rho, float temp_0, frt_otf* out) {
  if (!ctx || !out) return FRT_E_ARG;
  // the ray list must hold live rays only: compacting it here would use the
  // ray buffers' scan workspace, so removals pending since the last
  // find_avg_unu are refused instead
  if (ctx->have_alive)
    return set_error(ctx, FRT_E_STATE,
                     "rays were removed since the last compaction: call frt_step or frt_split_rays first");
  if (!ctx->otf_scratch) return set_error(ctx, FRT_E_STATE, "no device scratch on this context");
  const int64_t nb = std::min<int64_t>((ctx->g.n3 + 255) / 256, FRT_OTF_BLOCKS);
  const int64_t nbr = std::max<int64_t>(1, std::min<int64_t>((ctx->nrays + 255) / 256, FRT_OTF_BLOCKS));
  if ((size_t)(nb * FRT_OTF_NV + nbr) > (size_t)FRT_OTF_BLOCKS * (FRT_OTF_NV + 1))
    return set_error(ctx, FRT_E_CAPACITY, "write_otf partial sums exceed the device scratch");
  // every partial the final combination reads is written by this call, so
  // nothing needs clearing; calls that overlap on one GPU share the array and race
  double* part = (double*)ctx->otf_scratch;
  double* dout = part + (size_t)FRT_OTF_BLOCKS * (FRT_OTF_NV + 1);
  int rc = launch_otf(ctx, avg_rho, temp_0, part, dout);
  if (rc) return rc;
  double* S = (double*)ctx->hscalar;  // pinned, 512 bytes
  static_assert((FRT_OTF_NV + 1) * sizeof(double) <= 512, "pinned readback buffer");
  rc = hip_check(ctx, hipMemcpyAsync(S, dout, (FRT_OTF_NV + 1) * sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream), "otf");
  if (rc) return rc;
  rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "otf");
  if (rc) return rc;
  otf_fill(ctx, S, out);
  return FRT_OK;
}

}  // extern "C"

// calc_vol_avg / calc_mass_avg forms of the reduced sums (io_funcs.cc:500-632).
static void otf_fill(const frt_ctx* ctx, const double* S, frt_otf* out) {
  // calc_vol_avg: (float)tot / pow(N, 3.); calc_mass_avg: (float)tot / m_tot
  const double N3 = (double)ctx->g.n3;
  auto vol = [&](int i) { return (float)((double)(float)S[i] / N3); };
  auto mass = [&](int i, int m) { return (float)((double)(float)S[i] / S[m]); };
  frt_otf& o = *out;
  memset(&o, 0, sizeof(o));
  o.fH1[0] = vol(0); o.fH1[1] = mass(17, 7);
  o.fH2[0] = vol(1); o.fH2[1] = mass(18, 7);
  o.fHe1[0] = vol(2); o.fHe1[1] = mass(19, 8);
  o.fHe2[0] = vol(3); o.fHe2[1] = mass(20, 8);
  o.fHe3[0] = vol(4); o.fHe3[1] = mass(21, 8);
  o.nH1 = vol(5);
  o.nHe1 = vol(6);
  // fion = 1: weighting by fion is the volume average (m_tot = N^3 exactly)
  o.nH[0] = o.nH[1] = vol(7); o.nH[2] = mass(23, 22);
  o.nHe[0] = o.nHe[1] = vol(8); o.nHe[2] = mass(24, 22);
  o.ne = vol(9);
  const float GH1[7] = {vol(10), vol(10), mass(25, 22), mass(32, 7), mass(34, 33), (float)S[46], (float)S[45]};
  const float RH2[5] = {vol(11), vol(11), mass(26, 22), mass(35, 7), mass(37, 36)};
  const float T[5] = {vol(12), vol(12), mass(27, 22), mass(38, 7), mass(40, 39)};
  const float MFP[6] = {vol(13), vol(13), mass(28, 22), mass(41, 7), (float)S[48], (float)S[47]};
  const float M912[4] = {vol(14), vol(14), mass(29, 22), mass(42, 7)};
  const float Heat[5] = {vol(15), vol(15), mass(30, 22), mass(43, 7), mass(43, 7)};
  const float Cool[5] = {vol(16), vol(16), mass(31, 22), mass(44, 7), mass(44, 7)};
  memcpy(o.GH1, GH1, sizeof(GH1));
  memcpy(o.RH2, RH2, sizeof(RH2));
  memcpy(o.T, T, sizeof(T));
  memcpy(o.MFP, MFP, sizeof(MFP));
  memcpy(o.MFP_912, M912, sizeof(M912));
  memcpy(o.Heat, Heat, sizeof(Heat));
  memcpy(o.Cool, Cool, sizeof(Cool));
  o.npht = S[FRT_OTF_NV] * FRT_PHOT_NORM;
  o.nsrc = ctx->nsrc;
  o.fion[0] = (float)((double)(float)N3 / N3);
  o.fion[1] = (float)((double)(float)S[7] / S[7]);
}

extern "C" {

int frt_write_otf(frt_ctx* ctx, const char* path, float avg_rho, float temp_0) {
  if (!ctx || !path) return FRT_E_ARG;
  frt_otf o;
  int rc = frt_otf_averages(ctx, avg_rho, temp_0, &o);
  if (rc) return rc;
  FILE* f = fopen(path, "a");
  if (!f) return set_error(ctx, FRT_E_IO, std::string("cannot open ") + path + " for appending");
  const double yr_to_s = 3.154e7;  // global_constants.h:43
  // the reference's lines (io_funcs.cc:557-562, 633-654)
  fprintf(f, "Step Number    : %d\n", (int)ctx->step);
  fprintf(f, "Redshift       : %le\n", (double)ctx->g.zz);
  fprintf(f, "Time [Myr]     : %le\n", ctx->t / yr_to_s / 1e6);
  fprintf(f, "Time Step [Myr]: %le\n", ctx->g.dt / yr_to_s / 1e6);
  fprintf(f, "\n");
  fprintf(f, "Global Averages\n");
  fprintf(f, "fH1     : %le    %le\n", o.fH1[0], o.fH1[1]);
  fprintf(f, "fH2     : %le    %le\n", o.fH2[0], o.fH2[1]);
  fprintf(f, "fHe1    : %le    %le\n", o.fHe1[0], o.fHe1[1]);
  fprintf(f, "fHe2    : %le    %le\n", o.fHe2[0], o.fHe2[1]);
  fprintf(f, "fHe3    : %le    %le\n", o.fHe3[0], o.fHe3[1]);
  fprintf(f, "nH1     : %le       \n", o.nH1);
  fprintf(f, "nHe1    : %le       \n", o.nHe1);
  fprintf(f, "nH      : %le    %le    %le\n", o.nH[0], o.nH[1], o.nH[2]);
  fprintf(f, "nHe     : %le    %le    %le\n", o.nHe[0], o.nHe[1], o.nHe[2]);
  fprintf(f, "ne      : %le       \n", o.ne);
  fprintf(f, "GH1     : %le    %le    %le    %le    %le    %le    %le\n", o.GH1[0], o.GH1[1], o.GH1[2], o.GH1[3],
          o.GH1[4], o.GH1[5], o.GH1[6]);
  fprintf(f, "RH2     : %le    %le    %le    %le    %le\n", o.RH2[0], o.RH2[1], o.RH2[2], o.RH2[3], o.RH2[4]);
  fprintf(f, "T       : %le    %le    %le    %le    %le\n", o.T[0], o.T[1], o.T[2], o.T[3], o.T[4]);
  fprintf(f, "MFP     : %le    %le    %le    %le    %le    %le\n", o.MFP[0], o.MFP[1], o.MFP[2], o.MFP[3], o.MFP[4],
          o.MFP[5]);
  fprintf(f, "MFP_912 : %le    %le    %le    %le\n", o.MFP_912[0], o.MFP_912[1], o.MFP_912[2], o.MFP_912[3]);
  fprintf(f, "Heat    : %le    %le    %le     %le    %le\n", o.Heat[0], o.Heat[1], o.Heat[2], o.Heat[3], o.Heat[4]);
  fprintf(f, "Cool    : %le    %le    %le     %le    %le\n", o.Cool[0], o.Cool[1], o.Cool[2], o.Cool[3], o.Cool[4]);
  fprintf(f, "npht    : %le       \n", o.npht);
  fprintf(f, "nsrc    : %d        \n", o.nsrc);
  fprintf(f, "fion    : %le    %le\n", o.fion[0], o.fion[1]);
  fprintf(f, "\n");
  if (fclose(f) != 0) return set_error(ctx, FRT_E_IO, std::string("write to ") + path + " failed");
  return FRT_OK;
}

// ---- device checkpoints -------------------------------------------------------
}  // extern "C"

// The whole state a later step depends on, copied device to device: every
// cell field, the packed opacities, the live rays, the sources' accumulated
// ray_dt, and the scalars (time, step and casting counters, photon bounds,
// the grid constants that expansion and hydro steps change).
struct frt_checkpoint {
  int device = 0;
  int64_t n3 = 0, nrays = 0;
  int nfreq = 0, groups = 0;
  float* cells = nullptr;
  float2* avg2 = nullptr;
  float4* pos = nullptr;
  float4* dir = nullptr;
  int4* meta = nullptr;
  float4* phot = nullptr;
  float* src_ray_dt = nullptr;
  std::vector<int32_t> src_cell;
  std::vector<double> src_lum;
  GridConst g;
  frt_config cfg;
  float t = 0.f, nphot_merge = 0.f, gadget_step_time = 0.f;
  double nphot_min = 0., nphot_max = 0.;
  int64_t step = 0;
  uint64_t cast_calls = 0;
  bool have_rho = false;
};

static void checkpoint_release(frt_checkpoint* ck) {
  hipFree(ck->cells); hipFree(ck->avg2); hipFree(ck->pos); hipFree(ck->dir); hipFree(ck->meta);
  hipFree(ck->phot); hipFree(ck->src_ray_dt);
  delete ck;
}

extern "C" {

int frt_checkpoint_save(frt_ctx* ctx, frt_checkpoint** out) {
  if (!ctx || !out) return FRT_E_ARG;
  *out = nullptr;
  if (ctx->subgrid) return set_error(ctx, FRT_E_STATE, "device checkpoints do not hold the subgrid state");
  int rc = ensure_compact(ctx);
  if (rc) return rc;
  frt_checkpoint* ck = new frt_checkpoint();
  ck->device = ctx->device;
  ck->n3 = ctx->g.n3;
  ck->nfreq = ctx->g.nfreq;
  ck->groups = frt_phot_planes(ctx->g.nfreq);
  ck->nrays = ctx->nrays;
  const size_t ncell = (size_t)(C_NSCALAR + 3 * ck->nfreq) * (size_t)ck->n3;
  const size_t nr = (size_t)std::max<int64_t>(ck->nrays, 1);
  hipError_t e = hipMalloc((void**)&ck->cells, ncell * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&ck->avg2, (size_t)ck->n3 * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc((void**)&ck->pos, nr * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&ck->dir, nr * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&ck->meta, nr * sizeof(int4));
  if (e == hipSuccess) e = hipMalloc((void**)&ck->phot, frt_phot_bytes(ck->nfreq, (int64_t)nr));
  if (e == hipSuccess) e = hipMalloc((void**)&ck->src_ray_dt, (size_t)std::max(ctx->nsrc, 1) * sizeof(float));
  if (e != hipSuccess) {
    checkpoint_release(ck);
    return set_error(ctx, FRT_E_NOMEM, "checkpoint allocation");
  }
  const RaySoA& b = ctx->buf[ctx->cur];
  const size_t n = (size_t)ck->nrays;
  e = hipMemcpyAsync(ck->cells, ctx->cell_block, ncell * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ck->avg2, ctx->c.avg2, (size_t)ck->n3 * sizeof(float2), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(ck->pos, b.pos, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(ck->dir, b.dir, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(ck->meta, b.meta, n * sizeof(int4), hipMemcpyDeviceToDevice, ctx->stream);
  // photons: the 4-bin vectors, then each further bin's plane (ray-buffer layout with cap = n)
  if (e == hipSuccess && n) e = hipMemcpyAsync(ck->phot, b.phot, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
  for (int k = 0; k < ck->groups && e == hipSuccess && n; ++k)
    e = hipMemcpyAsync((float*)(ck->phot + n) + (size_t)k * n, (const float*)(b.phot + ctx->g.cap) + (size_t)k * ctx->g.cap,
                       n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && ctx->nsrc > 0)
    e = hipMemcpyAsync(ck->src_ray_dt, ctx->src_ray_dt, (size_t)ctx->nsrc * sizeof(float), hipMemcpyDeviceToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    checkpoint_release(ck);
    return hip_check(ctx, e, "checkpoint");
  }
  ck->src_cell = ctx->h_src_cell;
  ck->src_lum = ctx->h_src_lum;
  ck->g = ctx->g;
  ck->cfg = ctx->cfg;
  ck->t = ctx->t;
  ck->nphot_merge = ctx->nphot_merge;
  ck->gadget_step_time = ctx->gadget_step_time;
  ck->nphot_min = ctx->nphot_min;
  ck->nphot_max = ctx->nphot_max;
  ck->step = ctx->step;
  ck->cast_calls = ctx->cast_calls;
  ck->have_rho = ctx->have_rho;
  *out = ck;
  return FRT_OK;
}

int frt_checkpoint_restore(frt_ctx* ctx, const frt_checkpoint* ck) {
  if (!ctx || !ck) return FRT_E_ARG;
  if (ck->n3 != ctx->g.n3 || ck->nfreq != ctx->g.nfreq || ck->device != ctx->device)
    return set_error(ctx, FRT_E_ARG, "checkpoint of another grid, bin count or device");
  if (ck->nrays > ctx->g.cap) return set_error(ctx, FRT_E_CAPACITY, "checkpoint holds more rays than the buffer");
  int rc = FRT_OK;
  if (ck->src_cell != ctx->h_src_cell || ck->src_lum != ctx->h_src_lum) {
    rc = frt_set_sources(ctx, (int32_t)ck->src_cell.size(), ck->src_cell.data(), ck->src_lum.data());
    if (rc) return rc;
  }
  const size_t ncell = (size_t)(C_NSCALAR + 3 * ck->nfreq) * (size_t)ck->n3;
  const size_t n = (size_t)ck->nrays;
  const RaySoA& b = ctx->buf[ctx->cur];
  hipError_t e = hipMemcpyAsync(ctx->cell_block, ck->cells, ncell * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->c.avg2, ck->avg2, (size_t)ck->n3 * sizeof(float2), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(b.pos, ck->pos, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(b.dir, ck->dir, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(b.meta, ck->meta, n * sizeof(int4), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(b.phot, ck->phot, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
  for (int k = 0; k < ck->groups && e == hipSuccess && n; ++k)
    e = hipMemcpyAsync((float*)(b.phot + ctx->g.cap) + (size_t)k * ctx->g.cap, (const float*)(ck->phot + n) + (size_t)k * n,
                       n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && ctx->nsrc > 0)
    e = hipMemcpyAsync(ctx->src_ray_dt, ck->src_ray_dt, (size_t)ctx->nsrc * sizeof(float), hipMemcpyDeviceToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  rc = hip_check(ctx, e, "checkpoint restore");
  if (rc) return rc;
  const int64_t cap = ctx->g.cap;
  ctx->g = ck->g;
  ctx->g.cap = cap;
  ctx->cfg = ck->cfg;
  ctx->nrays = ck->nrays;
  ctx->deferred = false;
  ctx->psum_state = FRT_PSUM_NONE;
  ctx->t = ck->t;
  ctx->nphot_merge = ck->nphot_merge;
  ctx->gadget_step_time = ck->gadget_step_time;
  ctx->nphot_min = ck->nphot_min;
  ctx->nphot_max = ck->nphot_max;
  ctx->step = ck->step;
  ctx->cast_calls = ck->cast_calls;
  ctx->have_rho = ck->have_rho;
  ctx->have_segments = false;
  ctx->have_alive = false;
  ctx->adj_valid = false;
  return FRT_OK;
}

int frt_checkpoint_free(frt_checkpoint* ck) {
  if (!ck) return FRT_E_ARG;
  checkpoint_release(ck);
  return FRT_OK;
}

int frt_get_time(const frt_ctx* ctx, float* t) {
  if (!ctx || !t) return FRT_E_ARG;
  *t = ctx->t;
  return FRT_OK;
}

int frt_init_rt(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  // fresh start: empty buffer, run clock and merge threshold at their initial values
  ctx->nrays = 0;
  ctx->cur = 0;
  ctx->deferred = false;
  ctx->psum_state = FRT_PSUM_NONE;
  ctx->t = 0.f;
  ctx->step = 0;
  ctx->cast_calls = 0;
  ctx->nphot_merge = 0.f;
  ctx->have_alive = false;
  ctx->have_segments = false;
  if (ctx->subgrid) {
    // rt_funcs.cc:50-68 for a fresh start with I-front tracking: fion,
    // fion_max, tcross and the crossing bins 0, flags FALSE
    SgPtrs& sp = ctx->sg;
    const size_t n3 = (size_t)ctx->g.n3;
    hipError_t e = hipMemsetAsync(sp.fion, 0, 4 * n3 * (6 + 2 * FRT_SG_NTCROSS), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->c.f[C_TREION_FLAG], 0, 4 * n3, ctx->stream);
    if (e != hipSuccess) return hip_check(ctx, e, "init_rt");
  }
  if (ctx->nsrc > 0) {
    std::vector<float> rdt((size_t)ctx->nsrc, ctx->g.dt);
    hipError_t e = stream_copy(ctx, ctx->src_ray_dt, rdt.data(), 4 * (size_t)ctx->nsrc, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_check(ctx, e, "init_rt");
  }
  return FRT_OK;
}

int frt_get_nphot_min_max(frt_ctx* ctx, double* nphot_min, double* nphot_max) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  int counter = 0;
  double avg = 0.;
  for (int s = 0; s < ctx->nsrc; ++s) {
    const double lum = ctx->h_src_lum[s];
    if (lum > 0.) {
      double tmp = 0.;
      for (int nu = 0; nu < ctx->g.nfreq; ++nu)
        tmp += (double)ctx->g.spectrum[nu] * lum * (double)ctx->g.dt / 12. / dm_pow(2., 2. * ctx->cfg.hpx_lvl) / FRT_PHOT_NORM;
      avg += dm_log10(tmp);
      counter += 1;
    }
  }
  avg /= counter;
  avg = dm_pow(10., avg);
  ctx->nphot_max = 1e3 * avg;
  ctx->nphot_min = 1e-10 * avg;
  ctx->g.nphot_min = ctx->nphot_min;
  if (nphot_min) *nphot_min = ctx->nphot_min;
  if (nphot_max) *nphot_max = ctx->nphot_max;
  return FRT_OK;
}

// Entry points on the step path write the host mirror of the step scalars to
// the device first (the host may have changed time, step or rays since) and
// pull them back at the end; nothing between reads the device.
int frt_new_rays(frt_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = push_step_scalars(ctx))) return rc;
  if ((rc = ensure_compact(ctx))) return rc;
  const int64_t before = ctx->nrays;
  if ((rc = launch_cast(ctx))) return rc;
  rc = finish_stats(ctx);
  ctx->stats.rays_new = ctx->nrays - before;
  return rc;
}

int frt_update_rays(frt_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = push_step_scalars(ctx))) return rc;
  if (ctx->have_alive && (rc = ensure_compact(ctx))) return rc;  // a deferred compaction stays for the march
  if ((rc = ensure_segment_capacity(ctx))) return rc;
  if ((rc = reset_counters(ctx))) return rc;
  if ((rc = do_update_rays(ctx))) return rc;
  return finish_stats(ctx, true);
}

int frt_find_avg_unu(frt_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = push_step_scalars(ctx))) return rc;
  if ((rc = do_find_avg_unu(ctx, false))) return rc;
  return finish_stats(ctx);
}

int frt_merge_rays(frt_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = push_step_scalars(ctx))) return rc;
  if ((rc = do_merge(ctx))) return rc;
  return finish_stats(ctx);
}

int frt_split_rays(frt_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = push_step_scalars(ctx))) return rc;
  if ((rc = do_split(ctx))) return rc;
  return finish_stats(ctx);
}

int frt_update_thermal(frt_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  // c2ray_iter = FALSE: update_gamma and update_chem come first (main.cc:186-191)
  if (ctx->c2ray_off && (rc = launch_gamma_chem(ctx))) return rc;
  if ((rc = launch_thermal(ctx))) return rc;
  return finish_stats(ctx);
}

int frt_step(frt_ctx* ctx, frt_step_stats* stats) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  ctx->g.nphot_min = ctx->nphot_min;
  if ((rc = ensure_segment_capacity(ctx))) return rc;
  if (graph_eligible(ctx)) {
    rc = graph_step(ctx);
  } else {
    rc = push_step_scalars(ctx);
    if (rc == FRT_OK && ctx->have_alive) rc = ensure_compact(ctx);  // a deferred compaction stays for the march
    if (rc == FRT_OK) rc = launch_step_begin(ctx, true);  // t += dt (main.cc:123)
    if (rc == FRT_OK) rc = step_body(ctx);
  }
  if (rc) return rc;
  if ((rc = finish_stats(ctx, true))) return rc;
  if (stats) *stats = ctx->stats;
  return FRT_OK;
}

int frt_device_arch(char* buf, int32_t len) {
  if (!buf || len <= 0) return FRT_E_ARG;
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return FRT_E_HIP;
  std::string a(p.gcnArchName);
  a = a.substr(0, a.find(':'));
  snprintf(buf, (size_t)len, "%s", a.c_str());
  return FRT_OK;
}

int frt_set_step_graph(frt_ctx* ctx, int32_t on) {
  if (!ctx) return FRT_E_ARG;
  ctx->graph_on = on ? 1 : 0;
  if (!on) graph_invalidate(ctx);
  return FRT_OK;
}

int frt_set_tuning(frt_ctx* ctx, const char* key, int64_t value) {
  if (!ctx || !key) return FRT_E_ARG;
  const std::string k(key);
  int32_t* slot = nullptr;
  if (k == "deposit_queue") slot = &ctx->tune_deposit_queue;
  else if (k == "march_queue") slot = &ctx->tune_march_queue;
  else if (k == "queue_chunk") slot = &ctx->tune_queue_chunk;
  else if (k == "blocks_per_cu") slot = &ctx->tune_blocks_per_cu;
  else if (k == "grid_cap") slot = &ctx->tune_grid_cap;
  else if (k == "fp64_chunk") slot = &ctx->tune_fp64_chunk;
  else return set_error(ctx, FRT_E_ARG, "unknown tuning key: " + k);
  if ((k == "deposit_queue" || k == "march_queue") && value != 0 && value != 1)
    return set_error(ctx, FRT_E_ARG, k + " is 0 or 1");
  if (k == "queue_chunk" && (value < 64 || value > 65536 || value % 64 != 0))
    return set_error(ctx, FRT_E_ARG, "queue_chunk is a multiple of 64 in [64, 65536]");
  if (k == "blocks_per_cu" && (value < 1 || value > 64))
    return set_error(ctx, FRT_E_ARG, "blocks_per_cu is in [1, 64]");
  if (k == "grid_cap" && (value < 64 || value > (1 << 20)))
    return set_error(ctx, FRT_E_ARG, "grid_cap is in [64, 2^20]");
  if (k == "fp64_chunk" && value != 2048 && value != 4096)
    return set_error(ctx, FRT_E_ARG, "fp64_chunk is 2048 or 4096");
  *slot = (int32_t)value;
  return FRT_OK;
}

int frt_get_tuning(const frt_ctx* ctx, const char* key, int64_t* value) {
  if (!ctx || !key || !value) return FRT_E_ARG;
  const std::string k(key);
  if (k == "deposit_queue") *value = ctx->tune_deposit_queue;
  else if (k == "march_queue") *value = ctx->tune_march_queue;
  else if (k == "queue_chunk") *value = ctx->tune_queue_chunk;
  else if (k == "blocks_per_cu") *value = ctx->tune_blocks_per_cu;
  else if (k == "grid_cap") *value = ctx->tune_grid_cap;
  else if (k == "fp64_chunk") *value = ctx->tune_fp64_chunk;
  else if (k == "num_cu") *value = ctx->num_cu;
  else return FRT_E_ARG;
  return FRT_OK;
}

int frt_get_stats(const frt_ctx* ctx, frt_step_stats* stats) {
  if (!ctx || !stats) return FRT_E_ARG;
  *stats = ctx->stats;
  return FRT_OK;
}

int frt_ray_count(const frt_ctx* ctx, int64_t* count) {
  if (!ctx || !count) return FRT_E_ARG;
  *count = ctx->nrays;
  return FRT_OK;
}

int frt_download_rays(frt_ctx* ctx, frt_ray_view* v) {
  if (!ctx || !v) return FRT_E_ARG;
  int rc = ensure_compact(ctx);
  if (rc) return rc;
  const int64_t n = ctx->nrays;
  if (v->count < n) {
    v->count = n;
    return set_error(ctx, FRT_E_ARG, "ray view too small");
  }
  v->count = n;
  if ((rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync"))) return rc;
  if (n == 0) return FRT_OK;
  RaySoA& b = ctx->buf[ctx->cur];
  const int F = ctx->g.nfreq;
  const int PL = frt_phot_planes(F);
  std::vector<float4> pos(n), dir(n), phot(n);
  std::vector<float> phot_hi((size_t)PL * n);
  std::vector<int4> meta(n);
  hipError_t e = stream_copy(ctx, pos.data(), b.pos, 16 * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = stream_copy(ctx, dir.data(), b.dir, 16 * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = stream_copy(ctx, meta.data(), b.meta, 16 * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = stream_copy(ctx, phot.data(), b.phot, 16 * (size_t)n, hipMemcpyDeviceToHost);
  for (int k = 0; k < PL && e == hipSuccess; ++k)
    e = stream_copy(ctx, phot_hi.data() + (size_t)k * n, (const float*)(b.phot + ctx->g.cap) + (size_t)k * ctx->g.cap,
                    4 * (size_t)n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_check(ctx, e, "download rays");
  for (int64_t i = 0; i < n; ++i) {
    if (v->x) v->x[i] = pos[i].x;
    if (v->y) v->y[i] = pos[i].y;
    if (v->z) v->z[i] = pos[i].z;
    if (v->time) v->time[i] = pos[i].w;
    if (v->xhat) v->xhat[i] = dir[i].x;
    if (v->yhat) v->yhat[i] = dir[i].y;
    if (v->zhat) v->zhat[i] = dir[i].z;
    if (v->r) v->r[i] = dir[i].w;
    float dr;
    memcpy(&dr, &meta[i].x, 4);
    if (v->dr) v->dr[i] = dr;
    if (v->cell) v->cell[i] = frt_cell_to_rowmajor(ctx->g.lay, meta[i].y);
    if (v->pixel) v->pixel[i] = meta[i].z;
    if (v->order) v->order[i] = (int16_t)(meta[i].w & 0xffff);
    if (v->itheta) v->itheta[i] = (uint16_t)((uint32_t)meta[i].w >> 16);
    if (v->nphot)
      for (int nu = 0; nu < F; ++nu) {
        const float4& q = phot[i];
        v->nphot[i * F + nu] = nu == 0 ? q.x : nu == 1 ? q.y : nu == 2 ? q.z : nu == 3 ? q.w
                                                                              : phot_hi[(size_t)(nu - 4) * n + i];
      }
  }
  return hip_check(ctx, e, "download rays");
}

int frt_upload_rays(frt_ctx* ctx, const frt_ray_view* v) {
  if (!ctx || !v || v->count < 0) return FRT_E_ARG;
  const int64_t n = v->count;
  if (n > ctx->g.cap) return set_error(ctx, FRT_E_CAPACITY, "more rays than capacity");
  if (n > 0 && (!v->order || !v->pixel || !v->itheta || !v->xhat || !v->yhat || !v->zhat || !v->nphot ||
                !v->r || !v->dr || !v->x || !v->y || !v->z || !v->time || !v->cell))
    return FRT_E_ARG;
  for (int64_t i = 0; i < n; ++i)
    if (v->cell[i] < 0 || v->cell[i] >= ctx->g.n3 || v->order[i] < 0 || v->order[i] > ctx->cfg.l_max ||
        v->itheta[i] >= FRT_NTHETA)
      return set_error(ctx, FRT_E_ARG, "ray field out of range");
  int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (rc) return rc;
  ctx->cur = 0;
  RaySoA& b = ctx->buf[0];
  hipError_t e = hipSuccess;
  if (n > 0) {
    const int F = ctx->g.nfreq;
    const int PL = frt_phot_planes(F);
    std::vector<float4> pos(n), dir(n), phot(n);
    std::vector<float> phot_hi((size_t)PL * n);
    std::vector<int4> meta(n);
    for (int64_t i = 0; i < n; ++i) {
      pos[i] = make_float4(v->x[i], v->y[i], v->z[i], v->time[i]);
      dir[i] = make_float4(v->xhat[i], v->yhat[i], v->zhat[i], v->r[i]);
      int dr;
      memcpy(&dr, &v->dr[i], 4);
      meta[i] = make_int4(dr, frt_cell_from_rowmajor(ctx->g.lay, v->cell[i]), v->pixel[i],
                          (v->order[i] & 0xffff) | ((int)v->itheta[i] << 16));
      float q[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      for (int nu = 0; nu < F; ++nu) q[nu] = v->nphot[i * F + nu];
      phot[i] = make_float4(q[0], q[1], q[2], q[3]);
      for (int k = 0; k < PL; ++k) phot_hi[(size_t)k * n + i] = q[4 + k];
    }
    e = stream_copy(ctx, b.pos, pos.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = stream_copy(ctx, b.dir, dir.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = stream_copy(ctx, b.meta, meta.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = stream_copy(ctx, b.phot, phot.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
    for (int k = 0; k < PL && e == hipSuccess; ++k)
      e = stream_copy(ctx, (float*)(b.phot + ctx->g.cap) + (size_t)k * ctx->g.cap, phot_hi.data() + (size_t)k * n,
                      4 * (size_t)n, hipMemcpyHostToDevice);
  }
  ctx->nrays = n;
  ctx->deferred = false;
  ctx->psum_state = FRT_PSUM_NONE;
  ctx->have_alive = false;
  ctx->have_segments = false;
  return hip_check(ctx, e, "upload rays");
}

int frt_download_paths(frt_ctx* ctx, int64_t n, int32_t* ray_end, int32_t* cells, float* dists) {
  if (!ctx || !ray_end || !cells || !dists) return FRT_E_ARG;
  if (!ctx->have_segments) return set_error(ctx, FRT_E_STATE, "no paths: call update_rays first");
  if (n != ctx->nrays) return set_error(ctx, FRT_E_ARG, "n must equal the ray count");
  int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (rc || n == 0) return rc;
  const int P = ctx->g.P;
  std::vector<int32_t> c((size_t)P * n);
  std::vector<float> d((size_t)P * n);
  hipError_t e = stream_copy(ctx, ray_end, ctx->seg_cnt, 4 * (size_t)n, hipMemcpyDeviceToHost);
  for (int p = 0; p < P && e == hipSuccess; ++p) {
    e = stream_copy(ctx, c.data() + (size_t)p * n, ctx->path_cell + (size_t)p * ctx->g.cap, 4 * (size_t)n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = stream_copy(ctx, d.data() + (size_t)p * n, ctx->path_dist + (size_t)p * ctx->g.cap, 4 * (size_t)n, hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) return hip_check(ctx, e, "download paths");
  for (int64_t i = 0; i < n; ++i)
    for (int p = 0; p < P; ++p) {
      const bool used = p < ray_end[i];
      cells[i * P + p] = used ? frt_cell_to_rowmajor(ctx->g.lay, c[(size_t)p * n + i]) : -1;
      dists[i * P + p] = used ? d[(size_t)p * n + i] : 0.f;
    }
  return FRT_OK;
}

int frt_set_profiling(frt_ctx* ctx, int32_t enable) {
  if (!ctx) return FRT_E_ARG;
  ctx->prof = enable ? 1 : 0;
  return FRT_OK;
}

int frt_kernel_time(const frt_ctx* ctx, const char* name, double* total_ms, int64_t* launches) {
  if (!ctx || !name) return FRT_E_ARG;
  auto it = ctx->prof_acc.find(name);
  if (total_ms) *total_ms = (it == ctx->prof_acc.end()) ? 0. : it->second.first;
  if (launches) *launches = (it == ctx->prof_acc.end()) ? 0 : it->second.second;
  return FRT_OK;
}

int frt_reset_kernel_times(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  prof_resolve(ctx);
  ctx->prof_acc.clear();
  return FRT_OK;
}

int frt_deposit_bytes(const frt_ctx* ctx, double* bytes_per_launch, int64_t* launches) {
  if (!ctx) return FRT_E_ARG;
  const int64_t l = ctx->deposit_launches_last;
  if (bytes_per_launch) *bytes_per_launch = l ? ctx->deposit_bytes_last / (double)l : 0.;
  if (launches) *launches = l;
  return FRT_OK;
}

int frt_copy_kernel(frt_ctx* ctx, int64_t nbytes, int32_t repeats) {
  if (!ctx || nbytes <= 0 || repeats < 1) return FRT_E_ARG;
  if (nbytes != ctx->copy_bytes) {
    hipStreamSynchronize(ctx->stream);
    hipFree(ctx->copy_src);
    hipFree(ctx->copy_dst);
    ctx->copy_src = ctx->copy_dst = nullptr;
    hipError_t e = hipMalloc(&ctx->copy_src, (size_t)nbytes);
    if (e == hipSuccess) e = hipMalloc(&ctx->copy_dst, (size_t)nbytes);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->copy_src, 0, (size_t)nbytes, ctx->stream);
    if (e != hipSuccess) return hip_check(ctx, e, "copy buffers");
    ctx->copy_bytes = nbytes;
  }
  int rc = launch_copy(ctx, nbytes, repeats);
  if (rc) return rc;
  rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "copy");
  if (ctx->prof) prof_resolve(ctx);
  return rc;
}

int frt_synchronize(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  const int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "synchronize");
  if (rc == FRT_OK && ctx->prof) prof_resolve(ctx);  // kernel times of calls outside frt_step
  return rc;
}

// ---- restart files ----------------------------------------------------------

int frt_write_ray_file(frt_ctx* ctx, const char* path) {
  if (!ctx || !path) return FRT_E_ARG;
  int rc = ensure_compact(ctx);
  if (rc) return rc;
  const int64_t n = ctx->nrays;
  const size_t rec = FRT_RAY_FILE_FIXED + 4 * (size_t)ctx->g.nfreq;
  std::vector<uint8_t> h((size_t)n * rec + 8);
  if (n > 0) {
    uint8_t* d = nullptr;
    if (hipMalloc((void**)&d, (size_t)n * rec) != hipSuccess)
      return set_error(ctx, FRT_E_NOMEM, "ray file staging buffer");
    rc = launch_pack_ray_file(ctx, d);
    hipError_t e = rc ? hipSuccess : stream_copy(ctx, h.data(), d, (size_t)n * rec, hipMemcpyDeviceToHost);
    hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (rc) return rc;
    rc = hip_check(ctx, e, "ray file download");
    if (rc) return rc;
  }
  const int64_t end = -1;  // end marker of read_rays_binary
  memcpy(h.data() + (size_t)n * rec, &end, 8);
  return write_file(ctx, path, h.data(), h.size());
}

int frt_read_ray_file(frt_ctx* ctx, const char* path) {
  if (!ctx || !path) return FRT_E_ARG;
  std::vector<uint8_t> h;
  int rc = read_file(ctx, path, h);
  if (rc) return rc;
  const int F = ctx->g.nfreq;
  const size_t rec = FRT_RAY_FILE_FIXED + 4 * (size_t)F;
  if (h.size() < 8 || (h.size() - 8) % rec != 0 || get_le<int64_t>(h.data() + h.size() - 8) != -1)
    return set_error(ctx, FRT_E_IO, std::string(path) + ": not a ray file of " + std::to_string(F) + " bins");
  const int64_t n = (int64_t)((h.size() - 8) / rec);
  std::vector<int16_t> order(n);
  std::vector<int32_t> pixel(n), cell(n);
  std::vector<uint16_t> itheta(n);
  std::vector<float> xh(n), yh(n), zh(n), nph((size_t)n * F), r(n), dr(n), x(n), y(n), z(n), tm(n);
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t* o = h.data() + (size_t)i * rec;
    if (get_le<int64_t>(o) < 0)
      return set_error(ctx, FRT_E_IO, std::string(path) + ": end marker before the end of the file");
    order[i] = get_le<int16_t>(o + 9);
    pixel[i] = get_le<int32_t>(o + 11);
    itheta[i] = get_le<uint16_t>(o + 15);
    xh[i] = get_le<float>(o + 17);
    yh[i] = get_le<float>(o + 21);
    zh[i] = get_le<float>(o + 25);
    for (int nu = 0; nu < F; ++nu) nph[(size_t)i * F + nu] = get_le<float>(o + 29 + 4 * nu);
    const uint8_t* q = o + 29 + 4 * F;
    r[i] = get_le<float>(q);
    dr[i] = get_le<float>(q + 4);
    x[i] = get_le<float>(q + 8);
    y[i] = get_le<float>(q + 12);
    z[i] = get_le<float>(q + 16);
    tm[i] = get_le<float>(q + 20);
    cell[i] = get_le<int32_t>(q + 24);
  }
  frt_ray_view v;
  memset(&v, 0, sizeof(v));
  v.count = n;
  v.order = order.data();
  v.pixel = pixel.data();
  v.itheta = itheta.data();
  v.xhat = xh.data();
  v.yhat = yh.data();
  v.zhat = zh.data();
  v.nphot = nph.data();
  v.r = r.data();
  v.dr = dr.data();
  v.x = x.data();
  v.y = y.data();
  v.z = z.data();
  v.time = tm.data();
  v.cell = cell.data();
  return frt_upload_rays(ctx, &v);
}

int frt_write_gas_file(frt_ctx* ctx, const char* path, const frt_file_extras* extras) {
  if (!ctx || !path) return FRT_E_ARG;
  const int64_t n3 = ctx->g.n3;
  const int F = ctx->g.nfreq;
  const size_t nval = 1 + (size_t)n3 * (FRT_GAS_FILE_FIXED + F);
  const float* src[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (extras) {
    src[0] = extras->rho;
    src[1] = extras->f_He3;
    src[2] = extras->mfp;
    src[3] = extras->mfp_912;
    src[4] = extras->fion;
    src[5] = extras->tcross;
    src[6] = extras->treion_out;
  }
  unsigned have = 0;
  for (int j = 0; j < 7; ++j)
    if (src[j]) have |= 1u << j;
  float *d = nullptr, *dext = nullptr;
  if (hipMalloc((void**)&d, nval * 4) != hipSuccess) return set_error(ctx, FRT_E_NOMEM, "gas file staging buffer");
  hipError_t e = hipSuccess;
  if (have && hipMalloc((void**)&dext, (size_t)7 * n3 * 4) != hipSuccess) e = hipErrorOutOfMemory;
  for (int j = 0; j < 7 && e == hipSuccess; ++j)
    if (src[j]) e = stream_copy(ctx, dext + (size_t)j * n3, src[j], (size_t)n3 * 4, hipMemcpyHostToDevice);
  std::vector<float> h(nval);
  int rc = (e == hipSuccess) ? launch_pack_gas_file(ctx, dext, have, ctx->have_rho ? 1 : 0, d) : FRT_OK;
  if (e == hipSuccess && rc == FRT_OK) e = stream_copy(ctx, h.data(), d, nval * 4, hipMemcpyDeviceToHost);
  hipStreamSynchronize(ctx->stream);
  hipFree(d);
  hipFree(dext);
  if (rc) return rc;
  rc = hip_check(ctx, e, "gas file");
  if (rc) return rc;
  h[0] = ctx->t;
  return write_file(ctx, path, h.data(), nval * 4);
}

int frt_read_gas_file(frt_ctx* ctx, const char* path, frt_file_extras* extras) {
  if (!ctx || !path) return FRT_E_ARG;
  std::vector<uint8_t> raw;
  int rc = read_file(ctx, path, raw);
  if (rc) return rc;
  const int64_t n3 = ctx->g.n3;
  const int F = ctx->g.nfreq;
  const int W = FRT_GAS_FILE_FIXED + F;
  if (raw.size() != 4 * (1 + (size_t)n3 * W))
    return set_error(ctx, FRT_E_IO, std::string(path) + ": not a gas file of this grid and bin count");
  const float* v = reinterpret_cast<const float*>(raw.data());
  std::vector<float> col((size_t)n3);
  auto column = [&](int j) {
    for (int64_t c = 0; c < n3; ++c) col[(size_t)c] = v[1 + (size_t)c * W + j];
  };
  // (file column, field) of the fields this library holds
  static const int held[][2] = {{0, FRT_F_RHO},           {1, FRT_F_TEMP},          {2, FRT_F_NH1},
                                {3, FRT_F_FH1},           {4, FRT_F_FHE1},          {5, FRT_F_FH2},
                                {6, FRT_F_FHE2},          {7, FRT_F_FHE3},          {8, FRT_F_GAMMA_H1_TOT},
                                {9, FRT_F_GAMMA_HE1_TOT}, {10, FRT_F_MFP},          {11, FRT_F_MFP_912},
                                {12, FRT_F_HEAT_RATE},    {13, FRT_F_COOL_RATE},    {14, FRT_F_AVG_NH1}};
  for (const auto& hf : held) {
    column(hf[0]);
    rc = frt_upload_field(ctx, hf[1], col.data());
    if (rc) return rc;
  }
  // restart: avg_gamma = gamma_H1_tot, avg_gamma_prev = avg_gamma (init_funcs.cc:115-118)
  column(8);
  rc = frt_upload_field(ctx, FRT_F_GAMMA_PREV, col.data());
  if (rc) return rc;
  std::vector<float> ng((size_t)n3 * F);
  for (int64_t c = 0; c < n3; ++c)
    for (int nu = 0; nu < F; ++nu) ng[(size_t)c * F + nu] = v[1 + (size_t)c * W + FRT_GAS_FILE_FIXED + nu];
  rc = frt_upload_field(ctx, FRT_F_NGAM_NU, ng.data());
  if (rc) return rc;
  if (extras) {
    float* dst[7] = {extras->rho, extras->f_He3, extras->mfp, extras->mfp_912, extras->fion, extras->tcross,
                     extras->treion_out};
    static const int cols[7] = {0, 7, 10, 11, 15, 16, 17};
    for (int j = 0; j < 7; ++j)
      if (dst[j])
        for (int64_t c = 0; c < n3; ++c) dst[j][c] = v[1 + (size_t)c * W + cols[j]];
  }
  return frt_set_time(ctx, v[0]);
}

int frt_write_treion_file(frt_ctx* ctx, const char* path, const frt_file_extras* extras) {
  if (!ctx || !path) return FRT_E_ARG;
  const int64_t n3 = ctx->g.n3;
  std::vector<float> tre((size_t)n3), flag((size_t)n3);
  int rc = frt_download_field(ctx, FRT_F_TREION, tre.data());
  if (!rc) rc = frt_download_field(ctx, FRT_F_TREION_FLAG, flag.data());
  if (rc) return rc;
  std::vector<uint8_t> h((size_t)n3 * 7);
  for (int64_t c = 0; c < n3; ++c) {
    uint8_t* o = h.data() + (size_t)c * 7;
    memcpy(o, &tre[(size_t)c], 4);
    o[4] = flag[(size_t)c] != 0.f ? 1 : 0;
    o[5] = (extras && extras->tcross_flag) ? (extras->tcross_flag[c] ? 1 : 0) : 0;
    // neutral_flag starts TRUE (rt_funcs.cc:69) and only the I-front mode clears it
    o[6] = (extras && extras->neutral_flag) ? (extras->neutral_flag[c] ? 1 : 0) : 1;
  }
  return write_file(ctx, path, h.data(), h.size());
}

int frt_write_vif_file(frt_ctx* ctx, const char* path) {
  if (!ctx || !path) return FRT_E_ARG;
  const size_t n3 = (size_t)ctx->g.n3;
  float4* d = nullptr;
  if (hipMalloc((void**)&d, n3 * sizeof(float4)) != hipSuccess) return set_error(ctx, FRT_E_NOMEM, "vIF staging buffer");
  std::vector<float> h(4 * n3);
  int rc = launch_pack_vif_file(ctx, d);
  hipError_t e = (rc == FRT_OK) ? stream_copy(ctx, h.data(), d, n3 * sizeof(float4), hipMemcpyDeviceToHost) : hipSuccess;
  hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (rc) return rc;
  rc = hip_check(ctx, e, "vIF file");
  if (rc) return rc;
  return write_file(ctx, path, h.data(), h.size() * 4);
}

int frt_update_gadget_var(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  if (!ctx->have_rho) return set_error(ctx, FRT_E_ARG, "update_gadget_var needs rho: upload FRT_F_RHO first");
  ctx->gadget_step_time += ctx->g.dt;  // io_funcs.cc:436
  return launch_gadget(ctx, 0, nullptr);
}

int frt_write_gadget_file(frt_ctx* ctx, const char* path) {
  if (!ctx || !path) return FRT_E_ARG;
  const size_t n3 = (size_t)ctx->g.n3;
  float* d = nullptr;
  if (hipMalloc((void**)&d, 2 * n3 * sizeof(float)) != hipSuccess)
    return set_error(ctx, FRT_E_NOMEM, "gadget file staging buffer");
  std::vector<float> h(2 * n3);
  int rc = launch_gadget(ctx, 1, d);
  hipError_t e = (rc == FRT_OK) ? stream_copy(ctx, h.data(), d, 2 * n3 * sizeof(float), hipMemcpyDeviceToHost)
                                : hipSuccess;
  hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (rc) return rc;
  rc = hip_check(ctx, e, "gadget file");
  if (rc) return rc;
  return write_file(ctx, path, h.data(), h.size() * 4);
}

int frt_reset_gadget_var(frt_ctx* ctx) {
  if (!ctx) return FRT_E_ARG;
  ctx->gadget_step_time = 0.f;
  return launch_gadget(ctx, 2, nullptr);
}

int frt_get_gadget_step_time(const frt_ctx* ctx, float* step_time) {
  if (!ctx || !step_time) return FRT_E_ARG;
  *step_time = ctx->gadget_step_time;
  return FRT_OK;
}

int frt_read_treion_file(frt_ctx* ctx, const char* path, frt_file_extras* extras) {
  if (!ctx || !path) return FRT_E_ARG;
  std::vector<uint8_t> h;
  int rc = read_file(ctx, path, h);
  if (rc) return rc;
  const int64_t n3 = ctx->g.n3;
  if (h.size() != (size_t)n3 * 7) return set_error(ctx, FRT_E_IO, std::string(path) + ": not a treion file of this grid");
  std::vector<float> tre((size_t)n3), flag((size_t)n3);
  for (int64_t c = 0; c < n3; ++c) {
    const uint8_t* o = h.data() + (size_t)c * 7;
    memcpy(&tre[(size_t)c], o, 4);
    flag[(size_t)c] = o[4] ? 1.f : 0.f;
    if (extras && extras->tcross_flag) extras->tcross_flag[c] = o[5];
    if (extras && extras->neutral_flag) extras->neutral_flag[c] = o[6];
  }
  rc = frt_upload_field(ctx, FRT_F_TREION, tre.data());
  if (!rc) rc = frt_upload_field(ctx, FRT_F_TREION_FLAG, flag.data());
  return rc;
}

}  // extern "C"

// ---- sub-grid I-front model ------------------------------------------------------
extern "C" {
int frt_set_subgrid(frt_ctx* ctx, const frt_subgrid* sg) {
  if (!ctx || !sg) return FRT_E_ARG;
  if (!ctx->subgrid) return set_error(ctx, FRT_E_STATE, "frt_set_subgrid: the context was created with subgrid = 0");
  if (sg->nt < 2 || !sg->time_rh || !sg->fh1_rh) return set_error(ctx, FRT_E_ARG, "frt_set_subgrid: tables");
  const size_t n = (size_t)27 * sg->nt;
  hipStreamSynchronize(ctx->stream);
  hipFree(ctx->sg_tables);
  ctx->sg_tables = nullptr;
  hipError_t e = hipMalloc((void**)&ctx->sg_tables, 8 * n);
  if (e == hipSuccess) e = stream_copy(ctx, ctx->sg_tables, sg->time_rh, 4 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = stream_copy(ctx, ctx->sg_tables + n, sg->fh1_rh, 4 * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_check(ctx, e, "frt_set_subgrid");
  ctx->sg.time_rh = ctx->sg_tables;
  ctx->sg.fh1_rh = ctx->sg_tables + n;
  SgConst& c = ctx->sgc;
  c.nt = sg->nt;
  for (int i = 0; i < 3; ++i) {
    c.tre_dom[i] = sg->tre_dom[i];
    c.delta_nl[i] = sg->delta_nl[i];
  }
  c.t_z12 = sg->t_z12;
  c.t_z8 = sg->t_z8;
  c.tinit = sg->tinit;
  c.avg_rho = sg->avg_rho;
  c.temp_0 = sg->temp_0;
  c.fh1_0 = sg->fh1_0;
  // time_RH[Nt-1][0][0][0], read by get_tradhydro_index (subgrid.cc:188): with
  // the one-element shift of the gamma index it holds row Nt-2 of the
  // (zre 6, delta +sqrt3, gamma 3e-12) table
  c.t_norm = sg->time_rh[(size_t)((2 * 3 + 2) * 3 + 2) * sg->nt + (sg->nt - 2)];
  c.iter = 0;
  c.last = 0;
  ctx->g.avg_rho = sg->avg_rho;  // also the thermal update's density term
  ctx->subgrid_set = true;
  return FRT_OK;
}

int frt_set_subgrid_delta_nl(frt_ctx* ctx, const float delta_nl[3]) {
  if (!ctx || !delta_nl) return FRT_E_ARG;
  if (!ctx->subgrid) return set_error(ctx, FRT_E_STATE, "subgrid = 0");
  for (int i = 0; i < 3; ++i) ctx->sgc.delta_nl[i] = delta_nl[i];
  return FRT_OK;
}
}  // extern "C"

namespace frt {
// Host <-> device copy of the I-front state, row-major cells on the host,
// [member][cell] planes in the internal order on the device.
int subgrid_state_copy(frt_ctx* ctx, const frt_subgrid_state* st, bool down) {
  if (!ctx || !st) return FRT_E_ARG;
  if (!ctx->subgrid) return set_error(ctx, FRT_E_STATE, "subgrid = 0");
  const size_t n3 = (size_t)ctx->g.n3;
  std::vector<int32_t> perm;
  cell_permutation(ctx, perm);
  SgPtrs& sp = ctx->sg;
  float* dev[9] = {sp.fion, sp.fion_max, sp.tcross, sp.tcross_flag, sp.treion_out,
                   sp.delta_fion, sp.tcross_new, sp.tcross_check, sp.fh1_nodes};
  float* host[9] = {st->fion, st->fion_max, st->tcross, st->tcross_flag, st->treion_out,
                    st->delta_fion, st->tcross_new, st->tcross_check, st->fh1_nodes};
  const int w[9] = {1, 1, 1, 1, 1, 1, FRT_SG_NTCROSS, FRT_SG_NTCROSS, 5};
  if (down) {
    int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (rc) return rc;
  }
  for (int m = 0; m < 9; ++m) {
    if (!host[m]) continue;
    std::vector<float> t((size_t)w[m] * n3);
    hipError_t e;
    if (down) {
      e = stream_copy(ctx, t.data(), dev[m], 4 * t.size(), hipMemcpyDeviceToHost);
      if (e == hipSuccess)
        for (size_t r = 0; r < n3; ++r)
          for (int q = 0; q < w[m]; ++q) host[m][r * w[m] + q] = t[(size_t)q * n3 + (size_t)perm[r]];
    } else {
      for (size_t r = 0; r < n3; ++r)
        for (int q = 0; q < w[m]; ++q) t[(size_t)q * n3 + (size_t)perm[r]] = host[m][r * w[m] + q];
      e = stream_copy(ctx, dev[m], t.data(), 4 * t.size(), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    if (e != hipSuccess) return hip_check(ctx, e, "subgrid state");
  }
  return FRT_OK;
}
}  // namespace frt

extern "C" {
int frt_download_subgrid_state(frt_ctx* ctx, const frt_subgrid_state* out) {
  return subgrid_state_copy(ctx, out, true);
}
int frt_upload_subgrid_state(frt_ctx* ctx, const frt_subgrid_state* in) {
  return subgrid_state_copy(ctx, in, false);
}
}  // extern "C"

// frt_internal.h -- device state and kernel launchers of the RT library.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/flexrt_gpu.h"
#include "frt_detmath.h"
#include "frt_healpix.h"
#include "frt_rates.h"
#include "frt_order.h"

#define FRT_NALPHA 35
#define FRT_PHOT_NORM 1e60
#define FRT_PDF_BINS 10000
#define FRT_MAX_LEVEL 16

// Ray buffer: each ray is four 16-byte vectors, each vector in its own array
// indexed by ray, so a kernel that needs a ray's geometry reads whole 16-byte
// words and a gather touches one sector per vector instead of one per field.
//   pos  = {x, y, z, time}
//   dir  = {xhat, yhat, zhat, r}
//   meta = {dr (float bits), cell, pixel, order | itheta << 16}
//   phot = photon counts: bins 0..3 of ray i in phot[i] (unused lanes zero),
//          bins 4..nfreq-1 in float planes after the cap vectors:
//          ((float*)(phot + cap))[(nu - 4) * cap + i], so a ray's photons
//          are 16 + 4 (nfreq - 4) bytes with no padding.
struct RaySoA {
  float4* pos;
  float4* dir;
  int4* meta;
  float4* phot;
};

// Float planes of the photon bins beyond the first four.
static inline __host__ __device__ int frt_phot_planes(int nfreq) { return nfreq > 4 ? nfreq - 4 : 0; }
// Bytes of the photon arrays for cap rays.
static inline __host__ __device__ size_t frt_phot_bytes(int nfreq, int64_t cap) {
  return (size_t)cap * (16 + 4 * (size_t)frt_phot_planes(nfreq));
}

// Scalars every kernel may need, passed by value.
struct GridConst {
  int n;            // cells per side
  FrtCellLayout lay;  // internal cell numbering (frt_order.h)
  int n3;           // n^3
  int nfreq;
  int P;            // path_length
  int64_t cap;      // ray capacity (stride of per-ray arrays)
  float dx;         // proper cell size, cm (dx[i], uniform)
  float dx0;        // comoving cell size, cm
  float sf;         // dx[0] / dx0
  float clight;     // cl_factor * cl (set_clight)
  float nc;         // nl_const
  float dt;         // time step, s
  float min_dist;
  float temp_min;
  double vol;       // dx*dy*dz as double
  double nphot_min;
  int l_max;
  int hpx_lvl;
  int frac_limit;
  int rec_case_b;
  int temp_ev, recomb_cool, free_free, compton;
  float hubble;         // H(1/(1+zz)) as float
  double compton_coef;  // compton_rate(zz)
  double t_cmb;         // 2.726*(1 + zz)
  float zz;
  float avg_rho;        // the driver's volume-average rho (thermal ddelta_dt term)
  uint64_t dither_seed; // opacity dither (frt_rng.h), 0 = off
  int pass;             // C2Ray pass of the current cell launch
  float nhe_to_nh;
  double sig912;        // sigmapi_H1(13.6/h_eV)
  double log4;          // dm_log(4.) (l_from_r's divisor)
  double sigH[FRT_MAX_FREQ];
  double sigHe[FRT_MAX_FREQ];
  double sigHe2[FRT_MAX_FREQ];  // sigmapi_He2(nu_phot) (update_gamma of the c2ray_off mode)
  float sigHf[FRT_MAX_FREQ];
  float sigHef[FRT_MAX_FREQ];
  double nu[FRT_MAX_FREQ];
  float spectrum[FRT_MAX_FREQ];
  float rmax_tab[FRT_MAX_LEVEL];
  float rmin_tab[FRT_MAX_LEVEL];
};

// Device-side counters of one step (integer reductions only).
struct DevCounters {
  unsigned long long moved;
  unsigned long long dumped;
  unsigned long long terminated;
  unsigned long long alive;
  unsigned long long merged_away;
  unsigned long long groups;
  unsigned long long splits;
  unsigned long long substeps;
  unsigned int path_overflow;
  unsigned int pad;
};
// The device holds FRT_CTR_COPIES copies of the counters; a block adds into
// copy blockIdx % FRT_CTR_COPIES, so a large grid does not serialize on one
// atomic address.  The host sums the copies (read_device_counters).
#define FRT_CTR_COPIES 16

// Counts and scalars that kernels of a step produce and later kernels of the
// same step consume, kept in device memory: no launch of a step needs a host
// read, so a step runs without stopping the stream and can be captured as
// one graph.  The host holds a mirror, written before and read after each
// entry point (push_step_scalars / pull_step_scalars).
struct StepDev {
  int64_t nrays;       // rays in the current list
  int64_t nseg;        // path segments of the current paths
  int64_t nmoved;      // rays with segments
  int64_t ncand;       // merge candidates
  int64_t nsplit;      // rays after the split (before new_rays)
  int64_t mcount;      // entries of the merge-candidate scan (0: no merge this step)
  int64_t ncompact;    // rays the compaction writes (0 after a capacity error)
  int64_t step;        // completed steps
  uint64_t cast_calls; // new_rays calls so far
  float t;             // simulation time
  float nphot_merge;   // merge threshold in photons / phot_norm
  int32_t pdf_on;      // count_rays exceeded the threshold: recompute nphot_merge
  int32_t merge_on;    // merge runs this step
  uint32_t itm;        // rotation index of this step's merge
  uint32_t err;        // FRT_SDERR_* bits
  int64_t queue[16];   // work-queue heads of the ray-queue kernels (FRT_Q_*)
};
#define FRT_Q_MARCH 0
#define FRT_Q_DEPOSIT 1  // + C2Ray pass, passes 0..14
#define FRT_SDERR_SPLIT_CAPACITY 1u  // rays after the split + new_rays exceed the buffer
#define FRT_SDERR_CAST_CAPACITY 2u   // new_rays would exceed the buffer

// Fixed grids of the device-count scan and sort (frt_scan_sort.hip).
#define FRT_SCAN_MAX_GRID 2048
#define FRT_SORT_MAX_GRID 4096

struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};

// Cell fields, each n^3 floats (per-bin fields bin-major [nfreq][n^3]).
enum CellField {
  C_NH = 0, C_NHE, C_NH1, C_NHE1, C_NH2, C_NHE2, C_NE, C_NTOT, C_FH1, C_FHE1, C_FH2, C_FHE2,
  C_TEMP, C_DNE_DT, C_CLUMP, C_NH1_PREV, C_AVG_NH1, C_AVG_NHE1, C_RECOMB_H2, C_RECOMB_HE2,
  C_GAMMA_H1_TOT, C_GAMMA_HE1_TOT, C_RECRATE_H2, C_RECRATE_HE2, C_TREION, C_TREION_FLAG,
  C_HEAT, C_COOL, C_IF_SPEED, C_INC_FLUX, C_RHO, C_FHE3, C_MFP, C_MFP912, C_GAMMA_PREV, C_GADGET_GAMMA,
  C_GADGET_HEAT, C_DDELTA_DT, C_RHO_PREV, C_NHE3, C_NH2_PREV, C_NHE1_PREV, C_NHE2_PREV, C_NHE3_PREV,
  C_RECOMB_HE3, C_NSCALAR
};

struct CellPtrs {
  float* f[C_NSCALAR];
  float* gamma_h1_nu;   // [nfreq][n3]
  float* gamma_he1_nu;  // [nfreq][n3]
  float* ngam_nu;       // [nfreq][n3]
  float* u_nu;          // [nfreq][n3] the c2ray_off mode's march deposit (reset every step)
  float2* avg2;         // [n3] {avg_nh1, avg_nhe1} of the next deposit pass, one 8-byte gather
};

// Deposits of path segments, one slot per segment in cell-sorted order (a
// cell's slots are contiguous and in ray order).
struct ContribPtrs {
  void* pair;          // [slot] {H, He} deposit summed over bins (2 values)
  void* rec;           // [slot] last pass: photons in per bin, then the signed length
  int rec_len;         // values per record (nfreq + 1 rounded up to a multiple of 4)
};

// Sub-grid I-front model (subgrid.cc, cfg.subgrid): constants by value and
// the per-cell state, in the internal cell order.
#define FRT_SG_NTCROSS 10  // Ntcross, global_variables.cc:224
struct SgConst {
  int nt;            // rows per RadHydro table
  float tre_dom[3];  // cosmic time at z_re = 12, 8, 6
  float t_z12, t_z8; // cosmic_time(1/13), cosmic_time(1/9)
  float tinit;
  float avg_rho;
  float temp_0;
  float fh1_0;
  float delta_nl[3]; // deltaNL_dom of the current step
  float t_norm;      // time_RH[Nt-1][0][0][0] as get_tradhydro_index reads it
  int iter;          // C2Ray pass (0 .. itercount-1)
  int last;          // the last pass
};
struct SgPtrs {
  float* fion;
  float* fion_max;
  float* tcross;
  float* tcross_flag;
  float* treion_out;
  float* delta_fion;
  float* tcross_new;    // [FRT_SG_NTCROSS][n3]
  float* tcross_check;  // [FRT_SG_NTCROSS][n3], 0 / 1
  float* fh1_nodes;     // [5][n3] fh1_gamma_interp
  double* ngam3;        // [3][n3] ngam of the current pass
  int32_t* rpos;        // [n3] first slot of the cell not yet resolved by the I-front count
  const float* time_rh; // [27][nt], [zre][delta][gamma] tables
  const float* fh1_rh;
};
// Ray state of the sub-grid attenuation walk between I-front rounds.
struct SgWalk {
  double* tau;          // [4][nfreq][cap]: tau_h1, tau_he1, dtau_h1, dtau_he1
  int32_t* p;           // [cap] next segment (99 after a dump or absorption)
  int32_t* np;          // [cap] segments processed
  int32_t* wait;        // [cap] slot of the pending I-front count, -1 if none
  uint8_t* flags;       // [cap] bit 0 dumped, bit 1 done
  double* ifr;          // [nfreq][seg_cap] photons into the neutral part, then left after it
  uint8_t* slot;        // [seg_cap] 0 not reached, 1 waiting, 2 absorbed, 3 no room, 4 skipped, 5 emptied
  int64_t* pending;     // device counter of rays still waiting
};

// Host values a captured step depends on (kernel arguments and launch
// shapes): a replay is valid while they are unchanged.
// psum_state: nothing known, the march's rays that did not move, every ray.
enum : int { FRT_PSUM_NONE = 0, FRT_PSUM_UNMOVED = 1, FRT_PSUM_VALID = 2 };

struct GraphKey {
  GridConst g;
  int32_t nsrc;
  const void* ptrs[8];
  int64_t seg_cap;
  int32_t tune[6];
  int32_t itercount;
  int64_t capture_rays;
  int32_t deferred;
};

struct frt_ctx {
  frt_config cfg;
  GridConst g;
  hipStream_t stream;
  int device;
  std::string err;

  // rays
  RaySoA buf[2];
  int cur;
  int64_t nrays;
  int32_t* path_cell;  // [P][cap]
  float* path_dist;    // [P][cap]
  int32_t* seg_cnt;    // [cap+1] segments per ray (0 if not moved)
  int32_t* seg_off;    // [cap+1]
  uint8_t* alive;      // [cap]
  int32_t* n_out;      // [cap+1]
  int32_t* out_off;    // [cap+1]
  int32_t* cand_flag;  // [cap+1]
  uint32_t* cand_key_tmp;  // [cap]
  int32_t* moved_idx;  // [cap] rays that moved this step (seg_cnt > 0), in list order
  int64_t nmoved;
  bool have_rho;       // rho was uploaded or read from a gas file
  double* otf_buf;     // write_otf partial sums (allocated at first use)
  void* otf_scratch;   // device address of this device's static scratch (frt_otf_averages_noalloc)
  float gadget_step_time;  // MP-Gadget coupling: dt summed since the last reset
  // adjoint of find_avg_unu (FP64 form): tape of the last recorded call
  bool adj_on;         // record a tape in find_avg_unu
  bool adj_valid;      // the tape and the paths belong to the last find_avg_unu
  int adj_iters;       // C2Ray passes on the tape
  int64_t adj_nrays;   // rays on the tape
  double* adj_tape;    // [itercount][4][n3] + [2][n3]
  float* adj_nph;      // [capacity][nfreq] photons at the start of the step
  uint8_t* adj_np;     // [itercount][capacity] per pass: segments processed, | 0x80 if dumped
  // multi-step adjoint: what the taped step read before overwriting it, and
  // the adjoint carried from step to step (frt_adj_run_*)
  float* adj_tstate;        // [5][n3] temp, heat, cool, nH1_prev, ne at the start of the taped step
  float* adj_raydt;         // [adj_raydt_cap] ray_dt of the taped step's cast
  int32_t adj_raydt_cap;
  int64_t split_gen;        // compactions so far (the taped step's split mapping is live until the next)
  int64_t adj_split_gen;
  int64_t adj_cast_base;    // first ray of that cast (-1: no cast taped)
  int64_t adj_split_n;      // rays before and after the taped step's split (-1: no split taped)
  int64_t adj_split_total;
  double* adjr_cell;        // [ADJR_NCELL][n3] carried state and parameter adjoints (internal cell order)
  double* adjr_ray;         // [capacity][nfreq] adjoint of the ray photons (stored units)
  double* adjr_ray2;        // [capacity][nfreq]
  double* adjr_lum;         // [nsrc] adjoint of the source photon rates
  double* adjr_work;        // per-step work arrays ([ADJR_NWORK][n3] + [2 nfreq][n3])
  int32_t* adjr_inv;        // [capacity] inverse of the compaction order
  int64_t adjr_nray;        // rays the ray adjoint refers to
  int64_t adjr_step;        // step count of the state the carried adjoint refers to
  bool adjr_on;
  bool have_segments;  // update_rays ran and find_avg_unu has not yet consumed the paths
  bool have_alive;     // find_avg_unu ran in this step (alive flags valid)
  int64_t nseg;
  StepDev* sd;          // device: this context's step scalars
  StepDev* hsd;         // pinned host mirror
  int32_t* split_sorted;  // compaction order of the last split: the source (ray << 2 | child) of each output
  bool deferred;          // the last split's compaction is left to the next march (or materialize_rays)
  // psum[i]: the float photon sum of ray i (sum_nphot, in bin order), the
  // value the merge histogram, the merge flag and the split flag compare.
  // The march writes it for the rays it does not move and for the rays it
  // gathers, the last deposit pass for the moved rays, merge_groups for the
  // merged rays; psum_state says which of them hold (FRT_PSUM_*).
  float* psum;            // [cap + 1]
  int psum_state;
  // sub-grid I-front model
  bool subgrid;
  bool c2ray_off;         // c2ray_iter = FALSE: deposit in the march, update_gamma + update_chem
  bool subgrid_set;       // frt_set_subgrid has been called
  SgConst sgc;
  SgPtrs sg;
  SgWalk sgw;
  void* sg_block;         // base of the per-cell state allocation
  float* sg_tables;       // time_rh then fh1_rh
  int64_t sg_slot_cap;    // segments the slot buffers hold
  int64_t* sg_pending_h;  // pinned host copy of sgw.pending
  int64_t sg_rounds;      // I-front rounds of the last find_avg_unu
  // scan/sort workspace (frt_scan_sort.hip)
  int64_t* ss_part;
  int32_t* ss_hist;
  int32_t* ss_hoff;
  int32_t* ss_grp;  // [2][ceil(n3 / 64)]: segment run of each group of 64 cells
  // step graph: one executable graph per ray-buffer parity, captured on first use
  int graph_on;         // frt_set_step_graph
  bool capturing;       // launches size their grids for capture_rays rays
  int64_t capture_rays; // ray-count bucket of the graph being captured
  bool graph_valid[2];
  hipGraphExec_t graph_exec[2];
  GraphKey graph_key[2];
  bool deferred_after_graph[2];   // host-side effects of a replay
  int32_t* sorted_after_graph[2];
  // launch tuning (frt_set_tuning)
  int32_t tune_deposit_queue;   // 1: deposit passes pull chunks of the moved list from a queue
  int32_t tune_march_queue;     // 1: the march pulls chunks of the ray list from a queue
  int32_t tune_queue_chunk;     // rays per queue pull (a multiple of 64)
  int32_t tune_blocks_per_cu;   // persistent grid of the queue kernels, blocks per CU
  int32_t tune_grid_cap;        // block cap of grid-stride launches
  int32_t tune_fp64_chunk;      // rays per block chunk of the FP64 deposit (1024, 2048, 4096, 8192)
  int32_t num_cu;

  // segment sort
  int64_t seg_cap;
  uint32_t* seg_key[2];
  int32_t* seg_val[2];
  int32_t* path_slot;  // [path_length][capacity]: cell-sorted slot of each path entry
  ContribPtrs contrib;
  int32_t* cell_start;
  int32_t* cell_end;
  int key_bits;   // bits of a cell index (segment sort)
  int tile_bits;  // bits of a tile index (ray-list sort)

  // merge
  int64_t mcap;
  uint32_t* m_key[2];
  int32_t* m_val[2];
  int32_t* pdf;  // [FRT_PDF_BINS]

  // cells
  CellPtrs c;
  float* cell_block;

  // sources
  int32_t nsrc;
  int32_t* src_cell;
  double* src_lum;
  float* src_ray_dt;
  std::vector<int32_t> h_src_cell;
  std::vector<double> h_src_lum;

  // tables
  float* rot_fwd;  // [NTHETA][9]
  float* rot_inv;  // [NTHETA][9]

  // scalars of the reference's global state
  float t;
  float nphot_merge;
  double nphot_min, nphot_max;
  int64_t step;
  uint64_t cast_calls;
  int spectrum_set;
  frt_step_stats stats;

  DevCounters* dcnt;       // device, FRT_CTR_COPIES copies
  DevCounters* hcnt;       // pinned host: FRT_CTR_COPIES raw copies, their sum in hcnt[0]
  int64_t* hscalar;        // pinned host scratch
  double* dsmall;          // device scratch for scalars computed on the device (64 doubles)
  void* cub_tmp;
  size_t cub_tmp_bytes;

  // profiling
  int prof;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, int64_t>> prof_acc;
  double deposit_bytes_last;
  int64_t deposit_launches_last;
  void* copy_src;
  void* copy_dst;
  int64_t copy_bytes;
};

// ---- kernel launchers (frt_kernels.hip) ---------------------------------------
namespace frt {
void build_rotation_tables(float* fwd, float* inv);
uint32_t itheta_hash(uint64_t seed, uint64_t stream, uint64_t counter);

int launch_march(frt_ctx* ctx);
int launch_segments(frt_ctx* ctx);
int launch_deposit_inline(frt_ctx* ctx);
int launch_gamma_chem(frt_ctx* ctx);
int reserve_workspace(frt_ctx* ctx);
int read_device_counters(frt_ctx* ctx);
int launch_find_avg_unu(frt_ctx* ctx, bool fuse_thermal);
int launch_thermal(frt_ctx* ctx);
int launch_merge(frt_ctx* ctx);
int launch_split_compact(frt_ctx* ctx, bool nosplit);
int launch_step_begin(frt_ctx* ctx, bool advance);
int launch_step_end(frt_ctx* ctx);
int materialize_rays(frt_ctx* ctx);
int ensure_psum(frt_ctx* ctx);
int launch_find_avg_unu_subgrid(frt_ctx* ctx, bool fuse_thermal);
int launch_thermal_subgrid(frt_ctx* ctx);
int alloc_subgrid(frt_ctx* ctx);
void free_subgrid(frt_ctx* ctx);
int subgrid_state_copy(frt_ctx* ctx, const frt_subgrid_state* st, bool down);
int launch_fill_alive(frt_ctx* ctx);
int ensure_segment_capacity(frt_ctx* ctx);
int launch_cast(frt_ctx* ctx);
int launch_add_ray_dt(frt_ctx* ctx);
int launch_copy(frt_ctx* ctx, int64_t nbytes, int repeats);
// Restart files (io_funcs.cc): fixed bytes of a ray record before the
// nfreq photon counts are added, and fixed floats of a gas-file cell record.
#define FRT_RAY_FILE_FIXED 57
#define FRT_GAS_FILE_FIXED 18
#define FRT_M_H 1.672622e-24  // global_constants.h:19
#define FRT_Y_HE 0.246        // global_constants.h:28
int launch_pack_ray_file(frt_ctx* ctx, uint8_t* out);
int launch_pack_gas_file(frt_ctx* ctx, const float* ext, unsigned have, int held_rho, float* out);
int launch_pack_vif_file(frt_ctx* ctx, float4* out);
int launch_gadget(frt_ctx* ctx, int mode, float* out);  // 0 accumulate, 1 average + pack, 2 clear
// Sums and extremes of write_otf: FRT_OTF_NSUM double sums, then max/min of
// gamma_H1_tot and of mfp; FRT_OTF_NV values in all, then the photon count.
#define FRT_OTF_NSUM 45
#define FRT_OTF_NV 49
#define FRT_OTF_BLOCKS 1024
int launch_otf(frt_ctx* ctx, float avg_rho, float temp_0, double* part, double* out);
int launch_cosmo_expansion(frt_ctx* ctx, float ratio, double rho_scale, bool rays);
// update_hydro_step's cell loop: densities from the new rho, then ddelta_dt
// and rho_prev (gas_funcs.cc:700-738).
int launch_hydro_reload(frt_ctx* ctx, double rho_scale, float dt_cosmic);

// device-count scan and stable radix sort (frt_scan_sort.hip): the count is
// *dcount + add (dcount may be NULL), bound a host-side upper bound that sizes
// the grid; dtotal (may be NULL) receives the scan's total; the sort orders
// by key bits [lo, bits) and leaves its result in key[*which], val[*which].
// dev_sort_cells: stable sort of *dcount (cell key, value) pairs by cell
// (keys < ncells <= 2^key_bits): inv[value] = sorted position and, for every
// cell, the range [cstart, cend) of its positions (empty: 0, 0); key and val
// are overwritten
int dev_scan(frt_ctx* ctx, const int32_t* in, int32_t* out, const int64_t* dcount, int64_t add, int64_t bound,
             int64_t* dtotal, const char* name);
// dev_scan_moved: dev_scan of the per-ray segment counts that also lists, in
// order, the first *dcount items with a positive count (moved[0 .. *dmoved))
int dev_scan_moved(frt_ctx* ctx, const int32_t* in, int32_t* out, int32_t* moved, const int64_t* dcount,
                   int64_t add, int64_t bound, int64_t* dtotal, int64_t* dmoved, const char* name);
// dev_compact: the items i < *dcount + add with flag[i] = 1, in order, to
// list[0 .. *dtotal) with kdst[r] = ksrc[list[r]]
int dev_compact(frt_ctx* ctx, const int32_t* flag, const uint32_t* ksrc, int32_t* list, uint32_t* kdst,
                const int64_t* dcount, int64_t add, int64_t bound, int64_t* dtotal, const char* name);
int dev_sort_pairs(frt_ctx* ctx, uint32_t* key[2], int32_t* val[2], const int64_t* dcount, int64_t add,
                   int64_t bound, int bits, int* which, const char* name, int lo = 0);
int dev_sort_cells(frt_ctx* ctx, uint32_t* key[2], int32_t* val[2], const int64_t* dcount, int64_t bound,
                   int key_bits, int64_t ncells, int32_t* inv, int32_t* cstart, int32_t* cend, const char* name);
int alloc_scan_sort(frt_ctx* ctx);
void free_scan_sort(frt_ctx* ctx);
// host mirror of the step scalars <-> device
int push_step_scalars(frt_ctx* ctx);
int pull_step_scalars(frt_ctx* ctx);
void graph_invalidate(frt_ctx* ctx);

// helpers shared with the C-ABI file
int ensure_cub_tmp(frt_ctx* ctx, size_t bytes);
hipError_t stream_copy(frt_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t stream_zero(frt_ctx* ctx, void* dst, size_t bytes);
int set_error(frt_ctx* ctx, int code, const std::string& msg);
int hip_check(frt_ctx* ctx, hipError_t e, const char* what);
void prof_begin(frt_ctx* ctx, const char* name, ProfRec* rec);
void prof_end(frt_ctx* ctx, ProfRec* rec);
void prof_resolve(frt_ctx* ctx);
}  // namespace frt

/* flexrt_gpu.h -- C ABI of the MI355X radiative-transfer core.
 *
 * The reference (bayu-wilson/FlexRT_imaging_reion) is one C++ translation unit
 * whose hot-path functions take no arguments and talk through ~120 global
 * arrays (global_variables.cc).  This library keeps that state resident in
 * HBM inside an frt_ctx and exposes one entry point per reference routine on
 * the per-timestep path, plus a fused step.  A driver that keeps main.cc
 * unchanged binds these from same-named replacement files for
 * ray_tracing.cc, rt_funcs.cc, healpix_funcs.cc, rates.cc and subgrid.cc
 * (see the binding stubs shipped with the library).
 *
 * Conventions
 *  - Every function returns an frt_status (0 = FRT_OK).  No C++ exception
 *    crosses the ABI; frt_last_error() holds a message for the last failure.
 *  - Host pointers are borrowed for the duration of the call only.
 *  - Calls on one frt_ctx are ordered on the context's HIP stream and must
 *    come from one host thread at a time.
 *  - Grid index of cell (i, j, k) is i*n*n + j*n + k (ray_tracing.cc:313).
 *  - Per-frequency fields are exchanged in the reference layout
 *    [cell][nfreq] (global_variables.cc:136-138, 210-212).
 */
#ifndef FLEXRT_GPU_H
#define FLEXRT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRT_MAX_FREQ 8
#define FRT_NTHETA 42875 /* Nalpha^3 rotation triples, healpix_funcs.cc:16-17 */

typedef struct frt_ctx frt_ctx;

typedef enum {
  FRT_OK = 0,
  FRT_E_ARG = 1,           /* invalid argument or configuration */
  FRT_E_HIP = 2,           /* HIP runtime or kernel launch failure */
  FRT_E_CAPACITY = 3,      /* ray buffer full ("processor full", ray_tracing.cc:275-287) */
  FRT_E_PATH_OVERFLOW = 4, /* a ray crossed more than path_length cells in a step (ray_tracing.cc:1323) */
  FRT_E_STATE = 5,         /* call out of order (e.g. find_avg_unu before update_rays) */
  FRT_E_NOMEM = 6,         /* device allocation failed */
  FRT_E_IO = 7             /* a restart file could not be written, read or parsed */
} frt_status;

typedef enum { FRT_PREC_FP64 = 0, FRT_PREC_MIXED = 1 } frt_precision;

/* Runtime form of the reference's compile-time configuration
 * (user_inputs.h, global_constants.h, global_variables.cc). */
typedef struct frt_config {
  int32_t n;                /* cells per side, Nx = Ny = Nz (user_inputs.h:75-77) */
  int32_t nfreq;            /* frequency bins (user_inputs.h:78), 1..FRT_MAX_FREQ */
  int32_t path_length;      /* segments stored per ray per step (user_inputs.h:107) */
  int32_t itercount;        /* C2Ray iterations per step (user_inputs.h:85) */
  int32_t precision;        /* frt_precision */
  int32_t hpx_lvl;          /* HEALPix level for casting and merging (user_inputs.h:81) */
  int32_t l_max;            /* maximum ray order (global_variables.cc:55) */
  int32_t ray_per_cell_max; /* target rays per cell (global_variables.cc:40) */
  int32_t rec_case_b;       /* 1: case B recombination, 0: case A (user_inputs.h:4) */
  int32_t temp_ev;          /* temperature evolution (user_inputs.h:22) */
  int32_t recomb_cool;      /* recombination cooling (user_inputs.h:17) */
  int32_t free_free;        /* free-free cooling (user_inputs.h:20) */
  int32_t compton;          /* Compton heating/cooling (user_inputs.h:21) */
  int32_t frac_limit;       /* 1: where both HI and HeI segment depths underflow exp()
                               the HI/HeI split of ray_tracing.cc:770-771 is 0/0; use
                               its limit 1/(1+exp(dtau_he - dtau_h)) instead of NaN */
  int32_t subgrid;          /* 1: the sub-grid I-front model in its production setting,
                               subgrid = ifront = equilibrium = TRUE, relaxed = FALSE
                               (user_inputs.h:11-15); FP64 form; needs frt_set_subgrid */
  int32_t c2ray_off;        /* 1: c2ray_iter = FALSE (user_inputs.h:10): photons deposited
                               during the march into u_nu (ray_tracing.cc:1329-1355,
                               1443-1470), then update_gamma and update_chem's backward-
                               difference solve (gas_funcs.cc:74-99,243-351) after the
                               split; FP64 form, subgrid 0, collisional ionization off */
  int64_t ray_capacity;     /* rays held per buffer; 0 selects 30 n^3 (global_variables.cc:40-43) */
  uint64_t seed;            /* seed of the rotation-index stream that replaces rand() */
  double box_ckpc;          /* comoving box side in kpc, Lx0 (user_inputs.h:92) */
  double a_rt;              /* scale factor of the grid: proper dx = a_rt * comoving dx */
  double zz;                /* redshift for H(a) and Compton terms (global_variables.cc:312) */
  double dt;                /* time step in s (global_variables.cc:38) */
  double cl_factor;         /* reduced speed of light / c (user_inputs.h:84) */
  double nl_const;          /* cell-crossing fraction that triggers a move (user_inputs.h:83) */
  double safety_factor;     /* rays-per-cell split parameter (user_inputs.h:86) */
  double min_dist;          /* in-cell snap distance in cell units (global_variables.cc:305) */
  double temp_min;          /* temperature floor in K (global_variables.cc:304) */
  double hh, omega_m, omega_l; /* cosmology used by H(a) (cosmo_funcs.cc:13-15) */
} frt_config;

/* Per-step counters (all reductions are deterministic). */
typedef struct frt_step_stats {
  int64_t step;             /* steps completed by this context */
  int64_t rays_start;       /* rays in the buffer when update_rays began */
  int64_t rays_moved;       /* rays that crossed the nc test (ray_tracing.cc:1252) */
  int64_t segments;         /* path segments stored by the march */
  int64_t rays_dumped;      /* rays whose last C2Ray pass ended in the dump rule (ray_tracing.cc:782) */
  int64_t rays_terminated;  /* rays removed by the photon minimum (ray_tracing.cc:931-933) */
  int64_t rays_counted;     /* count_rays() before merging (main.cc:158) */
  int64_t num_thresh;       /* merge threshold (ray_tracing.cc:157-170) */
  int64_t merge_candidates; /* rays with sum_nphot < nphot_merge (ray_tracing.cc:339) */
  int64_t merge_groups;     /* occupied (cell, direction) bins */
  int64_t rays_merged_away; /* rays absorbed into a base ray (ray_tracing.cc:401) */
  int64_t rays_split;       /* parent rays split into 4 (rt_funcs.cc:221) */
  int64_t rays_new;         /* rays cast by new_rays */
  int64_t rays_end;         /* rays in the buffer after new_rays */
  int64_t thermal_substeps; /* total update_thermal sub-steps over all cells */
  double nphot_min;         /* photon floor (ray_tracing.cc:153) */
  double nphot_max;
  float nphot_merge;        /* merge threshold in photons/phot_norm (ray_tracing.cc:217) */
  float t;                  /* simulation time after the step, s */
} frt_step_stats;

/* Cell fields (float32, n^3 values each, or n^3*nfreq for the _NU fields).
 * Names follow global_variables.cc.  RHO and F_HE3 are inputs of the cosmic
 * expansion (gas_funcs.cc:460-552) and of the gas file; MFP and MFP_912 are
 * written by the last C2Ray pass (ray_tracing.cc:1078-1081) for the gas
 * file. */
typedef enum {
  FRT_F_NH = 0,        /* nH */
  FRT_F_NHE,           /* nHe */
  FRT_F_NH1,           /* nH1 */
  FRT_F_NHE1,          /* nHe1 */
  FRT_F_NH2,           /* nH2 */
  FRT_F_NHE2,          /* nHe2 */
  FRT_F_NE,            /* ne */
  FRT_F_NTOT,          /* n_tot */
  FRT_F_FH1,           /* f_H1 */
  FRT_F_FHE1,          /* f_He1 */
  FRT_F_FH2,           /* f_H2 */
  FRT_F_FHE2,          /* f_He2 */
  FRT_F_TEMP,          /* temp */
  FRT_F_DNE_DT,        /* dne_dt */
  FRT_F_CLUMP,         /* clump (set_clump, subgrid.cc:799) */
  FRT_F_NH1_PREV,      /* nH1_prev */
  FRT_F_AVG_NH1,       /* avg_nh1 */
  FRT_F_AVG_NHE1,      /* avg_nhe1 */
  FRT_F_RECOMB_H2,     /* recomb_H2 */
  FRT_F_RECOMB_HE2,    /* recomb_He2 */
  FRT_F_GAMMA_H1_TOT,  /* gamma_H1_tot */
  FRT_F_GAMMA_HE1_TOT, /* gamma_He1_tot */
  FRT_F_RECRATE_H2,    /* recrate_H2 */
  FRT_F_RECRATE_HE2,   /* recrate_He2 */
  FRT_F_TREION,        /* treion */
  FRT_F_TREION_FLAG,   /* treion_flag as 0.0 / 1.0 */
  FRT_F_HEAT_RATE,     /* heat_rate */
  FRT_F_COOL_RATE,     /* cool_rate */
  FRT_F_IF_SPEED,      /* IF_speed */
  FRT_F_INC_FLUX,      /* Inc_Flux */
  FRT_F_GAMMA_H1_NU,   /* gamma_H1_nu[cell][nfreq] */
  FRT_F_GAMMA_HE1_NU,  /* gamma_He1_nu[cell][nfreq] */
  FRT_F_NGAM_NU,       /* ngam_nu[cell][nfreq] */
  FRT_F_RHO,           /* rho, g/cm^3 */
  FRT_F_FHE3,          /* f_He3 (0 after a C2Ray step, ray_tracing.cc:1128,1194) */
  FRT_F_MFP,           /* mfp = 1/sigma_HI(nu_phot[0])/avg_nh1 */
  FRT_F_MFP_912,       /* mfp_912 = 1/sigma_HI(13.6 eV)/avg_nh1 */
  FRT_F_GAMMA_PREV,    /* avg_gamma_prev (0 on a fresh start, Gamma_HI of the gas file on a restart) */
  FRT_F_GADGET_GAMMA,  /* gadget_gamma: sum of Gamma_HI dt since the last reset (then the average) */
  FRT_F_GADGET_HEAT,   /* gadget_heat: sum of heat_rate dt / rho since the last reset (then the average) */
  FRT_F_DDELTA_DT,     /* ddelta_dt: rate of change of rho/avg_rho between hydro steps (gas_funcs.cc:733), 1/s */
  FRT_F_RHO_PREV,      /* rho_prev: rho after the last hydro-step reload (gas_funcs.cc:737) */
  /* state of the c2ray_off mode's backward-difference solve (gas_funcs.cc:243-351): */
  FRT_F_NHE3,          /* nHe3 */
  FRT_F_NH2_PREV,      /* nH2_prev */
  FRT_F_NHE1_PREV,     /* nHe1_prev */
  FRT_F_NHE2_PREV,     /* nHe2_prev */
  FRT_F_NHE3_PREV,     /* nHe3_prev */
  FRT_F_RECOMB_HE3,    /* recomb_He3 */
  FRT_F_U_NU,          /* u_nu[cell][nfreq], the march's deposit in the c2ray_off mode
                          (ray_tracing.cc:1350,1465); 0 in the C2Ray mode */
  FRT_F_COUNT
} frt_field;

/* Host view of the ray buffer in structure-of-arrays form (ray_type,
 * global_variables.cc:20-36).  Arrays hold `count` entries; nphot holds
 * count*nfreq values in [ray][nfreq] order.  cell is path[0].ind. */
typedef struct frt_ray_view {
  int64_t count;
  int16_t* order;
  int32_t* pixel;
  uint16_t* itheta;
  float* xhat;
  float* yhat;
  float* zhat;
  float* nphot;
  float* r;
  float* dr;
  float* x;
  float* y;
  float* z;
  float* time;
  int32_t* cell;
} frt_ray_view;

/* ---- context ------------------------------------------------------------ */
const char* frt_version(void);
/* Fill cfg with the reference's compile-time values (user_inputs.h,
 * global_variables.cc) for an n^3 grid. */
int frt_config_default(frt_config* cfg, int32_t n);
int frt_create(const frt_config* cfg, frt_ctx** out);
int frt_destroy(frt_ctx* ctx);
const char* frt_last_error(const frt_ctx* ctx);

/* ---- inputs -------------------------------------------------------------- */
/* Spectrum weights per bin and the bin frequencies in Hz: spectrum[]
 * (gas_funcs.cc:44-70) and nu_phot[] (user_inputs.h:95). */
int frt_set_spectrum(frt_ctx* ctx, const double* nu_phot, const float* spectrum);
/* Source cells and photon rates (photons/s), as source_lum after
 * read_source_catalog (io_funcs.cc:157-233): one entry per source cell. */
int frt_set_sources(frt_ctx* ctx, int32_t nsrc, const int32_t* cell, const double* lum);
int frt_upload_field(frt_ctx* ctx, int32_t field, const float* host);
int frt_download_field(frt_ctx* ctx, int32_t field, float* host);
int frt_set_time(frt_ctx* ctx, float t);
int frt_get_time(const frt_ctx* ctx, float* t);
/* Restart: the number of completed steps, which keys the deterministic angle
 * streams that replace rand() in casting and merging.  Sets the state of a
 * context that ran `step` calls of frt_step after one frt_new_rays; the
 * step is reported in frt_step_stats.step. */
int frt_set_step(frt_ctx* ctx, int64_t step);
/* Rounding-sensitivity measurement, no counterpart in the reference (seed 0,
 * the default, turns it off): after every C2Ray pass each cell's stored
 * opacities avg_nh1 and avg_nhe1 move one float ulp up or down, chosen by a
 * hash of (seed, step, pass, row-major cell).  The reference stores these in
 * float, so the spread between runs with and without the dither is what
 * rounding at that storage does to the result; it bounds the agreement any
 * other precision form can be asked for.  The CPU checker applies the same
 * dither. */
int frt_set_opacity_dither(frt_ctx* ctx, uint64_t seed);
/* update_dt (rt_funcs.cc:9-12) computes dt on the host from dx; the driver
 * hands the new value over (dt > 0). */
int frt_set_dt(frt_ctx* ctx, float dt);
int frt_get_cell_size(const frt_ctx* ctx, float* dx);
/* update_cosmo_expansion (gas_funcs.cc:460-552) for the scale factors the
 * driver computed (scale_factor(t + cosmic_time(1/(1+zinit))) and the same
 * one dt earlier, cosmo_funcs.cc:42-78): the cell size grows by
 * a_current/a_prev; rho falls by (a_prev/a_current)^3 and nH, nHe, the ion
 * densities (from f_H2, f_He2, f_He3 with fion = 1), avg_nh1, avg_nhe1, ne,
 * n_tot and nH1_prev are recomputed from it; ray positions within their
 * cells are scaled and kept off the cell faces.  Requires RHO to have been
 * uploaded (or read from a gas file). */
int frt_update_cosmo_expansion(frt_ctx* ctx, float a_current, float a_prev);
/* The device part of update_hydro_step (gas_funcs.cc:554-752) once the driver
 * has read the next snapshot: upload its comoving density into FRT_F_RHO
 * (get_ugas io_funcs.cc:33-62) and its sources (frt_set_sources), then call
 * this with the snapshot's redshift, a_rt = scale_factor(t + tinit), the
 * driver's avg_rho after its a_rt^-3 scaling (gas_funcs.cc:697) and
 * dt_cosmic = cosmic_time(aa_next) - cosmic_time(aa) (float, :733): rho
 * becomes proper (times a_rt^-3), the cell size Lx0 a_rt / Nx, nH, nHe, the
 * ion densities, avg_nh1, avg_nhe1, ne, n_tot and nH1_prev are recomputed;
 * when the volume average of rho_prev (calc_vol_avg, data_funcs.cc:75-92) is
 * positive, ddelta_dt = (rho/avg_rho - rho_prev/avg_rho_prev)/dt_cosmic, and
 * rho_prev = rho (gas_funcs.cc:700-738); the redshift-dependent rates of
 * update_thermal follow zz, and its -2/3 ddelta_dt/(rho/avg_rho) dt term
 * (gas_funcs.cc:399) uses avg_rho, which frt_update_cosmo_expansion then
 * scales with rho (gas_funcs.cc:515). */
int frt_update_hydro_step(frt_ctx* ctx, float zz, float a_rt, float avg_rho, float dt_cosmic);

/* Global averages of write_otf (io_funcs.cc:500-657), reduced on the GPU,
 * each in the reference's form: volume average (float)(sum) / N^3 and mass
 * average (float)(sum f m) / sum m (calc_vol_avg / calc_mass_avg,
 * data_funcs.cc:75-112, sums in double), min / max as minarr / maxarr
 * (data_funcs.cc:22-51).  "fion"-weighted averages use fion = 1, "binary"
 * marks cells whose Gamma_HI and avg_gamma_prev (FRT_F_GAMMA_PREV) are both
 * >= 3e-14 /s and within a decade (io_funcs.cc:524-532); "mean_den" is the
 * temperature of cells within 10% of avg_rho with temp_0 for the neutral
 * share (:535-549). */
typedef struct frt_otf {
  float fH1[2], fH2[2], fHe1[2], fHe2[2], fHe3[2]; /* volume, mass */
  float nH1, nHe1;                                 /* volume averages of avg_nh1, avg_nhe1 */
  float nH[3], nHe[3];                             /* volume, fion, binary */
  float ne;
  float GH1[7];     /* volume, fion, binary, mass(nH), mass(nH1), min, max */
  float RH2[5];     /* recrate_H2: volume, fion, binary, mass(nH), mass(nH2) */
  float T[5];       /* volume, fion, binary, mass(nH), mean_den */
  float MFP[6];     /* volume, fion, binary, mass(nH), min, max */
  float MFP_912[4]; /* volume, fion, binary, mass(nH) */
  float Heat[5];    /* volume, fion, binary, mass(nH), mass(nH) */
  float Cool[5];
  double npht;      /* count_photons_io: photons in all rays */
  int32_t nsrc;
  float fion[2];
} frt_otf;
/* avg_rho: the driver's volume-average density (io_funcs.cc:61, scaled with
 * the expansion); temp_0: the initial gas temperature (user_inputs.h:96). */
int frt_otf_averages(frt_ctx* ctx, float avg_rho, float temp_0, frt_otf* out);
/* write_otf: appends the step's block of text to path, in the reference's
 * format (step, redshift, time and time step in Myr, then the averages). */
int frt_write_otf(frt_ctx* ctx, const char* path, float avg_rho, float temp_0);
/* No counterpart in the reference: frt_otf_averages for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and results; the partial sums live in a
 * static device array of the library (about 400 KiB per GPU, held until the
 * process ends) instead of a buffer allocated at the first call, and the
 * result is read back through pinned memory the context already holds.
 * Calls on one GPU share that array: calls that overlap on one GPU race.
 * Returns FRT_E_STATE when rays were removed since the last compaction (call
 * frt_step or frt_split_rays first) and FRT_E_CAPACITY if the partial sums
 * would not fit (not reachable with the fixed block counts). */
int frt_otf_averages_noalloc(frt_ctx* ctx, float avg_rho, float temp_0, frt_otf* out);

/* ---- reference routines on the per-step path ----------------------------- */
/* init_rt (rt_funcs.cc:36-126), fresh start: empty ray buffer, ray_dt = dt
 * for every source cell. */
int frt_init_rt(frt_ctx* ctx);
/* get_nphot_min_max (ray_tracing.cc:129-154). */
int frt_get_nphot_min_max(frt_ctx* ctx, double* nphot_min, double* nphot_max);
/* new_rays + init_rays (rt_funcs.cc:240-261, ray_tracing.cc:230-321). */
int frt_new_rays(frt_ctx* ctx);
/* update_rays + advance_ray (rt_funcs.cc:159-208, ray_tracing.cc:1222-1512),
 * C2Ray mode: stores each moved ray's path segments. */
int frt_update_rays(frt_ctx* ctx);
/* find_avg_unu (ray_tracing.cc:623-1216): itercount C2Ray passes of
 * attenuation, deposition and the time-averaged ionization update, then the
 * per-ray photon update and removal. */
int frt_find_avg_unu(frt_ctx* ctx);
/* main.cc:153-165: set_lmerge, get_num_thresh, count_rays,
 * calc_nphot_merge (when count > num_thresh) and merge. */
int frt_merge_rays(frt_ctx* ctx);
/* split_rays + split_ray + set_to_true (rt_funcs.cc:211-237,146-156,
 * ray_tracing.cc:482-619). */
int frt_split_rays(frt_ctx* ctx);
/* update_thermal (gas_funcs.cc:362-457). */
int frt_update_thermal(frt_ctx* ctx);
/* One timestep in main.cc:117-198 order (set_clump and set_clight are
 * uniform here): t += dt, update_rays, find_avg_unu, merge, split_rays,
 * update_thermal, new_rays.  Statistics are copied to *stats if non-NULL. */
int frt_step(frt_ctx* ctx, frt_step_stats* stats);
int frt_get_stats(const frt_ctx* ctx, frt_step_stats* stats);

/* ---- step graph and launch tuning ---------------------------------------------
 * No counterpart in the reference, whose timestep loop (main.cc:108-198) runs
 * the phases as OpenMP loops over per-thread ray partitions
 * (mem_funcs.cc:186-299).  Every launch of a step sizes itself from counts
 * in device memory, so a step never stops the stream to read one back.
 * frt_set_step_graph(ctx, 1): frt_step captures its launches into a HIP graph
 * on first use and replays it (one graph per ray-buffer parity; its grids
 * are shaped for the power-of-two bucket of the ray count at the start of
 * the step, and a step in another bucket, or a change of the grid, sources,
 * segment buffers or tuning, recaptures).  Results are the same bits as
 * without the graph.  Not used while profiling or recording an adjoint tape
 * (those steps run uncaptured). */
int frt_set_step_graph(frt_ctx* ctx, int32_t on);
/* Launch parameters, by name:
 *   "deposit_queue"  0/1: deposit passes pull chunks of the moved-ray list
 *                    from a device work queue with a persistent grid
 *   "march_queue"    0/1: the same for the march
 *   "queue_chunk"    rays per pull (multiple of 64)
 *   "blocks_per_cu"  persistent grid of the queue kernels, blocks per CU
 *   "grid_cap"       block cap of grid-stride launches
 *   "fp64_chunk"     2048 or 4096: rays per block chunk of the FP64 deposit
 * frt_get_tuning also reads "num_cu".  Results do not depend on them. */
int frt_set_tuning(frt_ctx* ctx, const char* key, int64_t value);
/* GPU architecture of the current device (e.g. "gfx950"), the key of the
 * tuning table's entries. */
int frt_device_arch(char* buf, int32_t len);
int frt_get_tuning(const frt_ctx* ctx, const char* key, int64_t* value);
/* Verification hook: the step's device-count exclusive scan of counts[0..n)
 * and stable radix sort of (keys[i], i) on the low `bits` key bits, results
 * to host arrays (any may be NULL). */
int frt_selftest_scan_sort(frt_ctx* ctx, int64_t n, int32_t bits, const uint32_t* keys, uint32_t* keys_out,
                           int32_t* perm_out, const int32_t* counts, int32_t* scan_out, int64_t* total);
/* Verification hook: the step's cell sort of the segments (the cell-sorted
 * slots that the deposit passes write and the cell kernel sums) on (cells[i],
 * i), cells[i] below the context's cell count: slot_out[i] = position of item
 * i in the stable sort by cell, cstart_out / cend_out = each cell's range of
 * positions (0, 0 when empty; arrays of N^3). */
int frt_selftest_sort_cells(frt_ctx* ctx, int64_t n, const uint32_t* cells, int32_t* slot_out, int32_t* cstart_out,
                            int32_t* cend_out);

/* ---- ray buffer ----------------------------------------------------------- */
int frt_ray_count(const frt_ctx* ctx, int64_t* count);
/* Copy the ray buffer to caller-owned arrays sized for view->count >= the
 * current count (set on return). */
int frt_download_rays(frt_ctx* ctx, frt_ray_view* view);
/* Replace the ray buffer with view->count rays (restart, tests). */
int frt_upload_rays(frt_ctx* ctx, const frt_ray_view* view);
/* Path segments stored by the last update_rays, before find_avg_unu
 * consumes them: ray_end[i] segments for ray i (0 if it did not move),
 * cells/dists laid out [ray][path_length] (ray[x].path, ray_end[x];
 * ray_tracing.cc:1323-1327,1439-1442,1471).  n must equal the ray count. */
int frt_download_paths(frt_ctx* ctx, int64_t n, int32_t* ray_end, int32_t* cells, float* dists);

/* ---- restart files (the reference's formats, io_funcs.cc) ----------------- */
/* Per-cell fields of the gas and treion files that this library does not
 * evolve (C2Ray mode, subgrid off), row-major [n^3]; any may be NULL.  A write
 * then uses the value noted; a read fills the non-NULL ones. */
typedef struct {
  float* rho;            /* g/cm^3; NULL: nH m_H / (1 - Y) */
  float* f_He3;          /* NULL: 0 */
  float* mfp;            /* NULL: 0 */
  float* mfp_912;        /* NULL: 0 */
  float* fion;           /* NULL: 1 (fion = 1 with subgrid off) */
  float* tcross;         /* NULL: 0 */
  float* treion_out;     /* NULL: 0 */
  uint8_t* tcross_flag;  /* treion file; NULL: 0 */
  uint8_t* neutral_flag; /* treion file; NULL: 0 */
} frt_file_extras;
/* write_rays_binary (io_funcs.cc:274-303): per ray its index (long), active
 * (bool, TRUE), order (short), pixel (int), itheta (unsigned short), xhat,
 * yhat, zhat, nphot[nfreq], r, dr, x, y, z, time (float) and its cell
 * path[0].ind (int, row-major), packed without padding; then the long -1.
 * The records are packed on the GPU and written with one host write. */
int frt_write_ray_file(frt_ctx* ctx, const char* path);
/* read_rays_binary (io_funcs.cc:119-145): the rays of a ray file, in file
 * order, replace the ray buffer (the indices are not used). */
int frt_read_ray_file(frt_ctx* ctx, const char* path);
/* write_gas_binary (io_funcs.cc:358-393): the time t (float), then per cell
 * in row-major order rho, temp, nH1, f_H1, f_He1, f_H2, f_He2, f_He3,
 * gamma_H1_tot, gamma_He1_tot, mfp, mfp_912, heat_rate, cool_rate, avg_nh1,
 * fion, tcross, treion_out and ngam_nu[nfreq] (float).  Packed on the GPU. */
int frt_write_gas_file(frt_ctx* ctx, const char* path, const frt_file_extras* extras);
/* read_grid_binary (io_funcs.cc:81-116): sets t and the fields of the file
 * that this library holds; the others go to extras (may be NULL). */
int frt_read_gas_file(frt_ctx* ctx, const char* path, frt_file_extras* extras);
/* write_treion / read_treion (io_funcs.cc:338-355,65-77): per row-major cell
 * treion (float), treion_flag, tcross_flag, neutral_flag (bool). */
int frt_write_treion_file(frt_ctx* ctx, const char* path, const frt_file_extras* extras);
int frt_read_treion_file(frt_ctx* ctx, const char* path, frt_file_extras* extras);
/* write_vIF_binary (io_funcs.cc:396-430): per row-major cell IF_speed,
 * Inc_Flux, fion (1 without I-front tracking) and Delta_fion (0), float. */
int frt_write_vif_file(frt_ctx* ctx, const char* path);

/* ---- sub-grid I-front model (cfg.subgrid = 1) ---------------------------------
 * The RadHydro tables of init_radhydro_data / import_subgrid_file
 * (subgrid.cc:20-44,132-184) as read: nt rows per (z_re, delta, Gamma) node,
 * in [zre][delta][gamma][row] order (zre 12, 8, 6; delta -sqrt3, 0, +sqrt3 in
 * linear delta/sigma; Gamma 3e-14, 3e-13, 3e-12), the time since ionization
 * already in s and the total HI fraction.  With equilibrium = TRUE the MFP
 * columns are not read (the MFP is the equilibrium value of
 * ray_tracing.cc:981-986) and with temp_ev = TRUE neither is the temperature
 * column (interp_temp, subgrid.cc:637-649, is not called).  The cosmic times
 * are the driver's cosmic_time() values (cosmo_funcs.cc:43-51). */
typedef struct frt_subgrid {
  int32_t nt;              /* rows per table (Nt, global_variables.cc:266), >= 2 */
  int32_t reserved;
  const float* time_rh;    /* [3][3][3][nt] time_RH, s */
  const float* fh1_rh;     /* [3][3][3][nt] fh1_RH */
  float tre_dom[3];        /* cosmic_time at z_re = 12, 8, 6 (subgrid.cc:180-183) */
  float t_z12;             /* cosmic_time(1/(1+12)) (subgrid.cc:262) */
  float t_z8;              /* cosmic_time(1/(1+8)) (subgrid.cc:265) */
  float tinit;             /* cosmic_time(1/(1+zinit)) (init_funcs.cc:11) */
  float avg_rho;           /* mean density avg_rho (global_variables.cc:84) */
  float temp_0;            /* neutral temperature temp_0 (user_inputs.h:96) */
  float fh1_0;             /* fH1_0, the cap of interp_nh1 (user_inputs.h:98) */
  float delta_nl[3];       /* deltaNL_dom of the first step (delta_dom_to_deltaNL_dom) */
} frt_subgrid;
/* Tables and constants of the sub-grid model (init_radhydro_data,
 * subgrid.cc:132-184, and main.cc:86-90); the I-front state starts as
 * init_rt leaves it (rt_funcs.cc:50-68: fion = fion_max = tcross = 0, no
 * crossing bins, flags FALSE). */
int frt_set_subgrid(frt_ctx* ctx, const frt_subgrid* sg);
/* deltaNL_dom for the next step, as main.cc:126-128 sets it each step from
 * delta_dom_to_deltaNL_dom (subgrid.cc:103-130; deltaNL_dom[1] = 0). */
int frt_set_subgrid_delta_nl(frt_ctx* ctx, const float delta_nl[3]);
/* Per-cell I-front state (global_variables.cc:206-233), row-major cells;
 * NULL members are skipped.  tcross_new and tcross_check hold Ntcross = 10
 * values per cell, fh1_nodes the 5 Gamma nodes of fh1_gamma_interp. */
typedef struct frt_subgrid_state {
  float* fion;
  float* fion_max;
  float* tcross;
  float* tcross_flag;      /* 0 / 1 */
  float* treion_out;
  float* delta_fion;
  float* tcross_new;       /* [cell][10] */
  float* tcross_check;     /* [cell][10], 0 / 1 */
  float* fh1_nodes;        /* [cell][5] */
} frt_subgrid_state;
int frt_download_subgrid_state(frt_ctx* ctx, const frt_subgrid_state* out);
int frt_upload_subgrid_state(frt_ctx* ctx, const frt_subgrid_state* in);

/* MP-Gadget coupling (io_funcs.cc:434-496), main.cc:229-240 with
 * MP_GADGET_HEAT: after each step update_gadget_var adds Gamma_HI dt and
 * heat_rate dt / rho (the I-front heat injection gadget_ini_heat is 0 without
 * the subgrid model, subgrid.cc:674-686) and the step's dt to the step time;
 * write_gadget_binary divides both by the step time, in place, and writes the
 * two row-major float cubes; reset_gadget_var clears them.  rho must have
 * been uploaded. */
int frt_update_gadget_var(frt_ctx* ctx);
int frt_write_gadget_file(frt_ctx* ctx, const char* path);
int frt_reset_gadget_var(frt_ctx* ctx);
/* gadget_step_time: the summed time steps since the last reset, s. */
int frt_get_gadget_step_time(const frt_ctx* ctx, float* step_time);

/* ---- device checkpoints ---------------------------------------------------------
 * No counterpart in the reference, whose checkpoints are the restart files
 * (io_funcs.cc:274-393, frt_write_*_file above): the whole state a later step
 * depends on (cell fields, live rays, the sources' ray_dt, time, step and
 * casting counters, photon bounds, grid constants), copied device to device
 * so that a run can return to it, e.g. every K steps of a gradient
 * computation.  A checkpoint lives on the context's device until freed, and
 * is restored into a context of the same grid and bin count. */
typedef struct frt_checkpoint frt_checkpoint;
int frt_checkpoint_save(frt_ctx* ctx, frt_checkpoint** out);
int frt_checkpoint_restore(frt_ctx* ctx, const frt_checkpoint* ck);
int frt_checkpoint_free(frt_checkpoint* ck);

/* ---- adjoint of find_avg_unu ---------------------------------------------------
 * No counterpart in the reference: the gradient of one find_avg_unu
 * (ray_tracing.cc:623-1216, C2Ray branch) with the ray paths, the dump
 * decisions and the recombination coefficients held at their forward values
 * (reverse sweep in FP64 for both precision forms).
 * frt_adj_enable(ctx, 1) makes every following find_avg_unu (also inside
 * frt_step) record a tape; frt_adj_find_avg_unu then takes the adjoint of the
 * n_HI field that call produced (a_nh1, row-major) and returns, for any
 * non-NULL output, the gradient with respect to the rays' starting photons
 * (g_nphot, [ray][nfreq] in the list order of that call, per unit of the
 * stored photons), the clumping factor (g_clump) and the inputs nH, nHe and
 * the starting n_HI, n_HeI (row-major cells).  Valid until the next
 * update_rays. */
int frt_adj_enable(frt_ctx* ctx, int32_t on);
/* Rays on the tape (the length of g_nphot in rays); 0 without a tape. */
int frt_adj_ray_count(const frt_ctx* ctx, int64_t* n);
int frt_adj_find_avg_unu(frt_ctx* ctx, const double* a_nh1, double* g_nphot, double* g_clump, double* g_nH,
                         double* g_nHe, double* g_nH1, double* g_nHe1);

/* ---- adjoint across steps ------------------------------------------------------
 * No counterpart in the reference: the gradient of a function of the state
 * after several frt_step calls (the timestep loop of main.cc:108-198 with the
 * C2Ray chemistry, thermal update and the split of every step) with respect to
 * the source rates, the clumping factor and the density of the initial state.
 * The split, dump, ray-survival and thermal sub-step decisions are held at
 * their forward values; a step that merged rays is refused (FRT_E_STATE).
 * Protocol, for a run of S steps from a saved initial state with the tape on
 * (frt_adj_enable(ctx, 1)):
 *   frt_adj_run_begin(ctx)            after the forward run, on its final state
 *   frt_adj_seed(ctx, a_nH1, a_nH)    adjoint of the final nH1 and nH (row-major)
 *   for s = S-1 .. 0: return to the state before step s (a checkpoint and
 *     recomputation), run frt_step once, then frt_adj_step_reverse(ctx)
 *   return to the initial state, then frt_adj_run_finish(ctx, ray_dt of the
 *     initial cast, g_lum[nsrc], g_clump[n3], g_log_rho[n3])
 * g_lum is per unit of each source's rate; g_log_rho scales every density of
 * the initial state together (nH, nHe, n_HI, n_HeI, n_e, nH1_prev).  The
 * reverse sweep runs in FP64 for both precision forms. */
int frt_adj_run_begin(frt_ctx* ctx);
int frt_adj_seed(frt_ctx* ctx, const double* a_nH1, const double* a_nH);
int frt_adj_step_reverse(frt_ctx* ctx);
int frt_adj_run_finish(frt_ctx* ctx, float init_ray_dt, double* g_lum, double* g_clump, double* g_log_rho);

/* ---- measurement ----------------------------------------------------------- */
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context stream and its duration accumulated per kernel name. */
int frt_set_profiling(frt_ctx* ctx, int32_t enable);
/* Accumulated time (ms) and launch count for the named kernel. */
int frt_kernel_time(const frt_ctx* ctx, const char* name, double* total_ms, int64_t* launches);
int frt_reset_kernel_times(frt_ctx* ctx);
/* Algorithmic bytes moved by the deposit kernel in the last step (model
 * used for the roofline figure). */
int frt_deposit_bytes(const frt_ctx* ctx, double* bytes_per_launch, int64_t* launches);
/* Device-to-device copy of nbytes through a plain copy kernel, used to
 * calibrate HBM counters. */
int frt_copy_kernel(frt_ctx* ctx, int64_t nbytes, int32_t repeats);
/* Block until all work queued on the context stream has finished; with
 * profiling on, the kernel times of the finished launches are then counted. */
int frt_synchronize(frt_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* FLEXRT_GPU_H */

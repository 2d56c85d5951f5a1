/*
 * letkf_engine.h -- C ABI of the MI355X LETKF local-analysis engine.
 *
 * Drop-in for the analysis step of sbPOM-LETKF v2:
 *   SUBROUTINE da_main(fcst3d,fcst2d,anal3d,anal2d)
 *   (letkf/run/src/common_da.f90:22-124, called from main.f90:92)
 * whose other inputs are the module globals of common_setting.f90:168-198
 * (local grid nij1/lon1/lat1/ri1/rj1/dlon1/dlat1/depth1/phi1 and the sorted
 * observation tables nobsgrd/obselm/obslon/obslat/obslev/obserr/obsdep/
 * obshdxf) and the PARAMETERs of common_setting.f90:20-131.
 *
 * Every real array has the caller's element type, chosen by
 * letkf_params_t.real_kind (4 = REAL(4), 8 = REAL(8): the reference's
 * r_size), and the reference's column-major layout:
 *   fcst3d/anal3d (nij1,nlev,nmem,nv3d)   fcst2d/anal2d (nij1,nmem,nv2d)
 *   depth1 (nij1,nlev)   nobsgrd (nlon,nlat) int32   obshdxf (nobs,nmem)
 * Functions return LETKF_OK (0) or a negative letkf_status_t; they never
 * abort the process.  letkf_last_error() describes the last failure of the
 * calling thread.
 */
#ifndef LETKF_ENGINE_H
#define LETKF_ENGINE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum letkf_status {
  LETKF_OK = 0,
  LETKF_ERR_ARG = -1,       /* bad argument / inconsistent sizes        */
  LETKF_ERR_HIP = -2,       /* HIP runtime error (allocation, launch)   */
  LETKF_ERR_EIG = -3,       /* eigensolve did not converge, or every    */
                            /* eigenvalue negative (common.f90:1019-1033) */
  LETKF_ERR_CAPACITY = -4,  /* local obs count beyond the engine limit  */
  LETKF_ERR_NONFINITE = -5  /* NaN/Inf met where a finite value is needed */
} letkf_status_t;

/* Arithmetic of the ensemble-space products. */
typedef enum letkf_precision {
  LETKF_PREC_FP64 = 0,      /* Gram, weights and transform in FP64       */
  LETKF_PREC_FP32 = 1,      /* Gram Y^T R^-1 Y on FP32 MFMA, FP64 eig    */
  LETKF_PREC_BF16X3 = 2,    /* Gram from hi/lo BF16 splits on MFMA with  */
                            /* FP32 accumulation, FP64 eig               */
  LETKF_PREC_FP32_PANEL = 3,/* as FP32, the accumulators summing one     */
                            /* panel of 32 obs at a time into FP32       */
                            /* running totals: 2-6x smaller Gram         */
                            /* rounding error, fewer waves per SIMD      */
  LETKF_PREC_FP32X3 = 4,    /* Gram from FP32 hi/lo splits of the FP64   */
                            /* operands, hi*hi + hi*lo + lo*hi on FP32   */
                            /* MFMA with FP32 accumulation, FP64 eig     */
  LETKF_PREC_FP32_DUAL64 = 5,/* FP32 MFMA Gram for the K x K form, FP64  */
                            /* Gram for the n x n dual form (batched)    */
  LETKF_PREC_BF16X3_TRI32 = 6 /* as BF16X3, the Householder reduction of */
                            /* the Gram (nmem <= 128) in FP32; the      */
                            /* matrix functions and the update in FP64  */
} letkf_precision_t;

/* How the per-point ensemble-space problem is solved. */
typedef enum letkf_form {
  LETKF_FORM_PLAIN = 0,     /* nmem x nmem eigensolve at every point with */
                            /* local obs, as letkf_core does             */
  LETKF_FORM_BATCHED = 1,   /* points binned by local-obs count; points  */
                            /* with nobsl < nmem solve the equivalent    */
                            /* nobsl x nobsl problem; (nmem <= 128) the  */
                            /* inverse and inverse square root of the    */
                            /* tridiagonalised matrix from shifted       */
                            /* solves (Zolotarev's rational function,    */
                            /* relative error < 1e-15) instead of its    */
                            /* eigen-decomposition                       */
  LETKF_FORM_BATCHED_QL = 2 /* as BATCHED, with the QL eigen-            */
                            /* decomposition of the tridiagonal matrix   */
} letkf_form_t;

typedef struct letkf_params {
  int32_t nmem;                 /* K                         (setting:36)  */
  int32_t nlon, nlat, nlev;     /* global grid               (setting:37-39) */
  int32_t nij1;                 /* horizontal points of this analysis      */
  int32_t nv3d, nv2d;           /*                           (setting:40-41) */
  int32_t iv3d_t, iv3d_s;       /* 1-based T and S slots for round_off    */
  int32_t id_z_obs;             /* SSH element id, 2567      (setting:59)  */
  int32_t real_kind;            /* 4 or 8: element type of all real arrays */
  int32_t roff;                 /* apply round_off           (setting:129) */
  int32_t precision;            /* letkf_precision_t                       */
  int32_t form;                 /* letkf_form_t                            */
  double pi, re;                /*                           (setting:27-28) */
  double cov_infl_mul;          /* rho                       (setting:71)  */
  double alpha_rtpp, alpha_rtps;/*                           (setting:74-75) */
  double sigma_obs, sigma_obsv; /* [m]                       (setting:94-95) */
  double dist_zero, dist_zerov; /* cut-offs, r_size values   (setting:97-98) */
  double tmin, tmax, smin, smax;/* round_off bounds          (setting:130-131) */
  int32_t stage_timing;         /* 1: time each stage with HIP events into */
                                /* letkf_stats_t (adds event records only); */
                                /* 2: also each eigensolve kernel launch    */
  int64_t workspace_bytes;      /* cap on the analysis workspaces a plan   */
                                /* sizes its point chunks by; 0: 45% of    */
                                /* the device memory free at the plan (plus */
                                /* what the handle holds), at most 96 GB    */
} letkf_params_t;

/* Local grid of the analysis (common_setting.f90:168-176). */
typedef struct letkf_grid {
  const void* lon1;   /* (nij1) */
  const void* lat1;
  const void* ri1;    /* 1-based global i of each point, as a real */
  const void* rj1;
  const void* dlon1;  /* grid spacing [m] */
  const void* dlat1;
  const void* phi1;   /* bottom depth; 1 marks land (round_off skips it) */
  const void* depth1; /* (nij1,nlev), negative [m], k = nlev the surface */
} letkf_grid_t;

/* Observation tables after set_obs (common_setting.f90:182-198): sorted by
 * (lat cell, lon cell), nobsgrd the cumulative count per cell. */
typedef struct letkf_obs {
  int32_t nobs;
  const int32_t* nobsgrd; /* (nlon,nlat) */
  const int32_t* elm;     /* (nobs) */
  const void* lon;        /* (nobs) */
  const void* lat;
  const void* lev;        /* depth [m], negative */
  const void* err;        /* obs error sd */
  const void* dep;        /* innovation y - H(xf mean) */
  const void* hdxf;       /* (nobs,nmem) dYf = H(dXf) */
} letkf_obs_t;

/* Per-call counters and stage times of the last analysis on a handle. */
typedef struct letkf_stats {
  int64_t points;          /* nij1*nlev analysed points                */
  int64_t active_points;   /* points with at least one local obs       */
  int64_t local_obs;       /* sum of nobsl over points                 */
  int64_t dual_points;     /* points solved in the nobsl x nobsl form  */
  int64_t max_nobsl;
  double ms_total, ms_search, ms_gram, ms_eig, ms_transform;
  double gram_flops;       /* 2*n*K^2 (or 2*n^2*K dual) summed          */
  double transform_bytes;  /* fcst read + anal written                 */
  int64_t chunks;          /* point chunks (one launch of each per-chunk */
                           /* kernel: fill, Gram, transform)            */
  double eig_flops;        /* (4/3) m^3 summed over active points, m the */
                           /* eigenproblem size (Householder reduction) */
  double ms_eigk[25];      /* stage_timing 2, K <= 128: summed time of   */
  int64_t n_eigk[25];      /* the eigensolve kernels and their launch     */
                           /* counts, index 5*kind + bucket; kind 0      */
                           /* tridiagonalisation, 1 QL, 2 Q^T on the     */
                           /* vectors, 3 rotation replay (form BATCHED: */
                           /* the rational functions' shifted solves),  */
                           /* 4 Q and the update; bucket 0..4 = size    */
                           /* <= 16, 32, 48, 64, 128                    */
  int64_t rotations;       /* stage_timing 2, K <= 128: QL rotations     */
                           /* logged over all slots (each replayed on   */
                           /* 1+nv3d+nv2d vectors, twice)               */
  int64_t fallback_slots;  /* K <= 128: points whose rotation log       */
                           /* overflowed (form BATCHED: whose matrix is */
                           /* not SPD or whose spectrum bound is past   */
                           /* the rational table or, with a reduced-    */
                           /* precision Gram, past LETKF_KAPPA_MAX --   */
                           /* their Gram then recomputed in FP64),      */
                           /* solved by the LDS eigensolver and the     */
                           /* full tile transform                       */
  int64_t bucket_slots[5]; /* K <= 128: points with local obs per size  */
                           /* bucket of the register eigensolve (size   */
                           /* <= 16, 32, 48, 64, 128)                   */
} letkf_stats_t;

typedef struct letkf_handle letkf_handle_t;

/* Fill p with the reference defaults of common_setting.f90 rounded to the
 * given real_kind (sigma_obs 300 km, sigma_obsv 100 m, rho 1, RTPP 0.9,
 * RTPS 0, ROFF bounds, cut-offs 2*sqrt(10/3)*sigma). Sizes are left 0. */
int letkf_params_default(letkf_params_t* p, int32_t real_kind);

/* Create / destroy an engine on the current HIP device.  The handle owns its
 * device workspaces and one HIP stream; it is not re-entrant, distinct
 * handles may be used from distinct threads. */
int letkf_create(const letkf_params_t* p, letkf_handle_t** out);
int letkf_destroy(letkf_handle_t* h);

/* da_main with HOST arrays (the Fortran drop-in): uploads, analyses,
 * downloads, synchronous.  Replaces common_da.f90:22-124. */
int letkf_da_main(letkf_handle_t* h, const letkf_grid_t* grid,
                  const letkf_obs_t* obs, const void* fcst3d,
                  const void* fcst2d, void* anal3d, void* anal2d);

/* da_main with DEVICE arrays (every pointer in grid/obs/fcst/anal is device
 * memory).  Work is queued on `stream` (NULL: the handle's stream) and the
 * call returns after it completes. */
int letkf_da_main_device(letkf_handle_t* h, const letkf_grid_t* grid,
                         const letkf_obs_t* obs, const void* fcst3d,
                         const void* fcst2d, void* anal3d, void* anal2d,
                         void* stream);

/* letkf_da_main_device in two parts, for callers that repeat an analysis
 * over one observation geometry (inflation or error tuning loops, the
 * adjoint's outer iterations) or capture it in a hipGraph.  No counterpart
 * in the reference, whose da_main (common_da.f90:22-124) recounts the local
 * obs on every call.
 *
 * letkf_da_main_plan: the geometry-dependent part -- local-obs counts and
 * offsets of every point, the chunk plan, the size buckets and the
 * workspaces -- from grid and from obs->nobsgrd/lon/lat/lev (and elm, err
 * for the search).  It reads results back to the host and may allocate;
 * it returns after the work on `stream` (NULL: the handle's) is done.
 *
 * letkf_da_main_enqueue: the rest of the analysis on the plan (transpose
 * of hdxf, local-obs fill, Gram, eigensolves, transform), as kernel
 * launches on `stream` only -- no allocation, no synchronisation, no host
 * read-back, no memset -- so it may be captured in a hipGraph.  Returns
 * after enqueueing.  grid, obs and the handle's sizes, form and
 * localisation (pi, re, sigma_obs(v), dist_zero(v)) must be those of the
 * plan (the array pointers are compared); the values of fcst, obs->dep,
 * obs->err, obs->hdxf and the inflation/precision parameters may differ
 * between enqueues.  A captured graph replays the kernel arguments of its
 * capture: new array contents are seen, new parameter values (rho, alpha,
 * precision) need a new capture.  The plan and the workspaces a captured
 * graph refers to stay valid until the next letkf_da_main_plan /
 * letkf_da_main_device / letkf_destroy on the handle.  Kernel status bits
 * stay on the device.
 *
 * letkf_da_main_status: waits for `stream` and returns the status of the
 * analyses enqueued since (each enqueue clears it first). */
int letkf_da_main_plan(letkf_handle_t* h, const letkf_grid_t* grid,
                       const letkf_obs_t* obs, void* stream);
int letkf_da_main_enqueue(letkf_handle_t* h, const letkf_grid_t* grid,
                          const letkf_obs_t* obs, const void* fcst3d,
                          const void* fcst2d, void* anal3d, void* anal2d,
                          void* stream);
int letkf_da_main_status(letkf_handle_t* h, void* stream);

/* In-project autotuner (no counterpart in the reference).  The register
 * eigensolve groups the active points of a chunk by local-problem size into
 * buckets of at most 16, 32, 48, 64 and 128; each bucket has up to three
 * vector paths: 1 = the orthogonal factor Q of the tridiagonalisation
 * accumulated and applied as a matrix (nmem <= 64), 2 = its Householder
 * reflectors applied to the vectors one wavefront sum per vector, 3 = the
 * reflectors applied to all the vectors at once in a (row group, column)
 * lane layout (nv3d + nv2d <= 7).  letkf_autotune times every path a bucket
 * can take on the first chunk of the given analysis (DEVICE arrays, as
 * letkf_da_main_device), keeps the fastest per bucket in the handle (the
 * fixed rule's path unless another is more than 3% faster), then runs the
 * whole analysis with them (anal3d/anal2d receive it).
 * ms (may be NULL) receives 15 times: ms[3q + v - 1] for bucket q (16, 32,
 * 48, 64, 128) under path v, -1 where not measured.  letkf_set_tuning /
 * letkf_get_tuning set or read the five choices (0 = the fixed rule: path
 * 3 where nv3d + nv2d <= 7, else 2 for buckets 64 and 128 or nmem > 64 and
 * 1 below; a path a bucket cannot take falls back to the rule). */
int letkf_autotune(letkf_handle_t* h, const letkf_grid_t* grid,
                   const letkf_obs_t* obs, const void* fcst3d,
                   const void* fcst2d, void* anal3d, void* anal2d, double* ms,
                   void* stream);
int letkf_set_tuning(letkf_handle_t* h, const int32_t* vpath);
int letkf_get_tuning(const letkf_handle_t* h, int32_t* vpath);

/* letkf_core (common_letkf.f90:37-210) batched over nb local problems given
 * in CSR form, DEVICE arrays: problem b owns rows offs[b]..offs[b+1]-1 of
 * hdxf (rows of nmem values, row-major), rdiag, rloc, dep; outputs
 * wvec (nb,nmem) and wmat (nb,nmem,nmem) with wmat[b][i][j] = Wmat(i,j).
 * A problem with no rows gets letkf_core_noobs (W = sqrt(rho) I, w = 0). */
int letkf_core_batched(letkf_handle_t* h, int32_t nb, const int64_t* offs,
                       const void* hdxf, const void* rdiag, const void* rloc,
                       const void* dep, void* wvec, void* wmat);

/* letkf_core_batched that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and results as letkf_core_batched, with
 * its scratch (Gram, eigenvectors, eigenvalues, innovation terms, row
 * counts, status) in a 64 MiB device array of the library, one per device,
 * held until the process ends.  Takes no handle: p supplies nmem (2..128),
 * real_kind (the element type of hdxf, rdiag, rloc, dep, wvec, wmat) and
 * cov_infl_mul; the other fields are not read.  At most 64 MiB / (16 nmem^2
 * + 16 nmem + 4) problems per call (1008 at nmem 64, 254 at nmem 128);
 * more return LETKF_ERR_CAPACITY.  Work is only enqueued on `stream` (NULL:
 * the null stream); the call returns at once.  *status (DEVICE int32)
 * receives the letkf_status_t of the problems (0, LETKF_ERR_EIG or
 * LETKF_ERR_NONFINITE) when the work completes.  Calls on one device share
 * the scratch: calls that overlap in time on one device race. */
int letkf_core_batched_noalloc(const letkf_params_t* p, int32_t nb,
                               const int64_t* offs, const void* hdxf,
                               const void* rdiag, const void* rloc,
                               const void* dep, void* wvec, void* wmat,
                               int32_t* status, void* stream);

/* obs_id_uniform (common.f90:112-168) and the cell sort of set_obs
 * (common_obs.f90:266-377) on DEVICE arrays: lon/lat the global (nlon,nlat)
 * grid, olon/olat the nobs unsorted positions.  Writes idx/idy (1-based
 * cells, 999 outside), perm (sorted position -> input index) and nobsgrd.
 * *nsorted receives the number of obs inside the grid. */
int letkf_bin_obs(letkf_handle_t* h, int32_t nlon, int32_t nlat,
                  const void* lon, const void* lat, int32_t nobs,
                  const void* olon, const void* olat, int32_t* idx,
                  int32_t* idy, int32_t* perm, int32_t* nobsgrd,
                  int32_t* nsorted);

/* AOEI (common_obs.f90:389-400) in place on DEVICE arrays:
 * err <- sqrt(max(err^2, dep^2 - hxfsprd^2)). */
int letkf_aoei(letkf_handle_t* h, int32_t nobs, void* err, const void* dep,
               const void* hxfsprd);

/* Grid of the observation operator: the global model grid and the element
 * constants Trans_XtoY reads from common_setting (common_setting.f90:30,
 * 42-60, 155-161).  Fields are column-major in the reference's layouts:
 * lon, lat, fsm (nlon,nlat); depth (nlon,nlat,nlev), negative metres,
 * k = nlev the surface, k = 1 a dummy layer.  Variable slots are 1-based.
 * All real arrays have the element type of real_kind (4 or 8). */
typedef struct letkf_xtoy_grid {
  int32_t nlon, nlat, nlev, nv3d, nv2d;
  int32_t real_kind;
  int32_t id_u_obs, id_v_obs, id_t_obs, id_s_obs, id_z_obs; /* 2819 2820 3073 3332 2567 */
  int32_t iv3d_u, iv3d_v, iv3d_t, iv3d_s, iv2d_z;
  double hnosea, undef;                                     /* land value of fsm; 999 */
  const void* lon;
  const void* lat;
  const void* depth;
  const void* fsm;
} letkf_xtoy_grid_t;

/* H(x) of every member on DEVICE arrays: Trans_XtoY (common.f90:391-500)
 * called by set_obs once per member (common_io.f90:829-831), batched.
 * elm/idx/idy/lono/lato/levo describe nobs obs (idx, idy 1-based cells);
 * v3d and v2d hold the members: element (i,j,k,v,m) of v3d at
 * (i-1) + nlon (j-1) + nlon nlat (k-1) + s[0] (v-1) + s[1] (m-1), of v2d
 * at (i-1) + nlon (j-1) + s[2] (v-1) + s[3] (m-1), with s = strides
 * (elements); strides NULL or zero entries give one packed
 * (nlon,nlat,nlev,nv3d) / (nlon,nlat,nv2d) block per member, and the
 * analysis layout fcst3d(nij1,nlev,nmem,nv3d) with nij1 = nlon nlat is
 * s = {nij1 nlev nmem, nij1 nlev, nij1 nmem, nij1}.  hx (nobs,nmem)
 * receives H(x_m), undef where the reference gives
 * undef; qc (nobs, may be NULL) receives 0 where any member is undef (the
 * tmpqc rule of common_io.f90:833-839), 1 elsewhere.  Requires nlev >= 6
 * (the Akima stencil).  Runs on `stream` (NULL: the default stream) and
 * returns when done.  No counterpart entry point in the reference, which
 * loops over members on the host. */
int letkf_trans_xtoy_device(const letkf_xtoy_grid_t* g, int32_t nobs,
                            const int32_t* elm, const int32_t* idx,
                            const int32_t* idy, const void* lono,
                            const void* lato, const void* levo, int32_t nmem,
                            const void* v3d, const void* v2d,
                            const int64_t* strides, void* hx, int32_t* qc,
                            void* stream);

/* Settings of set_obs (common_setting.f90:56-63, 96, 101-118 and the
 * gross_error parameter of common_obs.f90:35).  Bound classes of the
 * gross-error check, in the reference's test order: 0 SSH, 1 SST (T at
 * lev 0), 2 T, 3 SSS, 4 S, 5 SSU, 6 U, 7 SSV, 8 V; other elements are
 * rejected when |dep| > gross_error * err. */
typedef struct letkf_setobs_cfg {
  int32_t id_u_obs, id_v_obs, id_t_obs, id_s_obs, id_z_obs;
  int32_t lge_io; /* gross-error check on (applied only when aoei == 0) */
  int32_t aoei;   /* adaptive obs-error inflation after the sort */
  int32_t nslots; /* time slots; obs come ordered by slot */
  int32_t nbslot; /* base slot (1-based) */
  double sigma_obst, undef, gross_error;
  double qc_min[9], qc_max[9];
} letkf_setobs_cfg_t;

/* Fills the reference's settings: LGE_IO off, AOEI on, one slot. */
void letkf_setobs_default(letkf_setobs_cfg_t* c);

/* Observation tables written by letkf_set_obs_device: DEVICE buffers of
 * capacity nobs (hdxf nobs*nmem, nobsgrd nlon*nlat), the sorted obs first.
 * hdxf is (nobs_out,nmem) column-major, element (o,m) at o + nobs_out*m:
 * with nobsgrd, elm, lon, lat, lev, err, dep it is the letkf_obs_t that
 * letkf_da_main_device takes.  ins, idx, idy, dat, hxfmean, hxfsprd may be
 * NULL (not written).  nobs_out, ndel_proj and ndel_gross are written on
 * the host: obs kept, obs dropped by the projection (qc 0 from H(x) or a
 * cell outside the grid: numdelqc1) and by the gross-error check
 * (numdelqc2). */
typedef struct letkf_obs_set {
  int32_t* nobsgrd;
  int32_t* elm;
  int32_t* ins;
  int32_t* idx;
  int32_t* idy;
  void* lon;
  void* lat;
  void* lev;
  void* dat;
  void* err;
  void* dep;
  void* hxfmean;
  void* hxfsprd;
  void* hdxf;
  int32_t nobs_out, ndel_proj, ndel_gross;
} letkf_obs_set_t;

/* The body of set_obs after H(x) (common_obs.f90:139-402) on DEVICE
 * arrays of nobs unsorted obs (elm, lon, lat, lev, dat, err, idx, idy;
 * ins may be NULL), hx (nobs,nmem) column-major and qc from
 * letkf_trans_xtoy_device: ensemble mean, spread, perturbations and
 * innovation per obs, gross-error check, temporal localisation of err,
 * removal of qc 0 obs, stable sort by (idy, idx) with nobsgrd, AOEI.
 * nobslots (HOST, nslots counts, may be NULL when nslots == 1) splits the
 * obs into time slots.  Real arrays have the handle's real_kind; nmem is
 * the handle's.  Returns when done. */
int letkf_set_obs_device(letkf_handle_t* h, const letkf_setobs_cfg_t* cfg,
                         int32_t nlon, int32_t nlat, int32_t nobs,
                         const int32_t* nobslots, const int32_t* elm,
                         const int32_t* ins, const void* lon, const void* lat,
                         const void* lev, const void* dat, const void* err,
                         const int32_t* idx, const int32_t* idy,
                         const void* hx, const int32_t* qc,
                         letkf_obs_set_t* out);

/* lpfgm_core (common_lpfgm.f90:444-494) batched over nb local problems in
 * the CSR form of letkf_core_batched (DEVICE offs, hdxf, rdiag, rloc, dep;
 * outputs wvec (nb,nmem) = 0 and wmat (nb,nmem,nmem), wmat[b][i][j] =
 * Wmat(i,j)): mode 2 = lpf_core (213-263), the average of nmonte resampling
 * matrices (resampling_matrix, 89-197); mode 3 = gm_core (279-405) times
 * lpf_core, gm_core's alpha = (nmem-1)/cov_infl_gm (its Pa = C^-1 from a
 * Cholesky factorisation rather than mtx_eigen).  The reference draws
 * each resampling's nmem uniforms from MT19937 seeded by the wall clock
 * (com_rand_seed(nmem,0,.), common.f90:1066-1095); here the caller passes
 * them: rand (HOST, nmonte x nmem doubles in [0,1), resampling m in row m),
 * used by every problem.  rm_sr 0 = "MR" (draws sorted ascending), 1 =
 * "SR" (rand(i) = (rand(1) + i - 1)/nmem in place, as the reference
 * writes it); dp = diagonal priority.  nmem <= 128, nmonte <= 255.  A
 * problem with no rows gets lpfgm_core_noobs (W = I). */
int letkf_lpfgm_core_batched(letkf_handle_t* h, int32_t mode,
                             double cov_infl_gm, int32_t nmonte, int32_t rm_sr,
                             int32_t dp, const double* rand, int32_t nb,
                             const int64_t* offs, const void* hdxf,
                             const void* rdiag, const void* rloc,
                             const void* dep, void* wvec, void* wmat);

/* Settings of the LPF / LPFGM analysis (iswitch_da, nmonte, RM, DP and
 * cov_infl_gm of common_setting.f90:65, 79-89). */
typedef struct letkf_lpfgm_cfg {
  int32_t mode;        /* iswitch_da: 2 = LPF, 3 = LPFGM                     */
  int32_t nmonte;      /* resamplings averaged per point (100), <= 255       */
  int32_t rm_sr;       /* 0 = "MR" (multinomial), 1 = "SR" (systematic)      */
  int32_t dp;          /* diagonal priority (.true.)                         */
  double cov_infl_gm;  /* gm_core's multiplicative inflation (1.0)           */
  const double* rand;  /* HOST nmonte x nmem uniforms in [0,1): the draws    */
} letkf_lpfgm_cfg_t;

/* da_main (common_da.f90:22-124) with iswitch_da = 2 or 3: at every point
 * lpfgm_core (or lpfgm_core_noobs without local obs) replaces letkf_core,
 * then Analysis_Ensemble, RTPP, RTPS and round_off as for the LETKF.
 * DEVICE arrays as letkf_da_main_device; the handle's real_kind, sizes,
 * localisation, inflation and round_off settings apply (its precision and
 * form do not: the LPF core runs in FP64).  The draws are shared by every
 * point, as the reference's clock-seeded draws are within one clock tick.
 * nmem <= 128.  Runs on `stream` (NULL: the handle's) and returns when
 * done. */
/* The same on HOST arrays (as letkf_da_main): what the Fortran driver calls
 * in place of da_main when iswitch_da is 2 or 3. */
int letkf_da_main_lpfgm(letkf_handle_t* h, const letkf_grid_t* grid,
                        const letkf_obs_t* obs, const letkf_lpfgm_cfg_t* cfg,
                        const void* fcst3d, const void* fcst2d, void* anal3d,
                        void* anal2d);
int letkf_da_main_lpfgm_device(letkf_handle_t* h, const letkf_grid_t* grid,
                               const letkf_obs_t* obs,
                               const letkf_lpfgm_cfg_t* cfg, const void* fcst3d,
                               const void* fcst2d, void* anal3d, void* anal2d,
                               void* stream);

/* ensemble_mesp (common.f90:505-596) on DEVICE arrays, as write_ens_mpi
 * calls it for the analysis and the forecast (common_io.f90:922-924):
 * v3d(nij1,nlev,nmem,nv3d), v2d(nij1,nmem,nv2d) in, v3dm/v3ds
 * (nij1,nlev,nv3d) and v2dm/v2ds (nij1,nv2d) out, Fortran order, element
 * type of real_kind (4 or 8).  Mean = member sum in member order / nmem;
 * spread = sqrt(sum of squared deviations from that mean / (nmem - 1)),
 * undef when nmem = 1; bit-identical to the reference's loops.  Runs on
 * `stream` (NULL: the default stream) and returns when done. */
int letkf_ensemble_mesp_device(int32_t real_kind, int32_t nij1, int32_t nlev,
                               int32_t nmem, int32_t nv3d, int32_t nv2d,
                               double undef, const void* v3d, const void* v2d,
                               void* v3dm, void* v3ds, void* v2dm, void* v2ds,
                               void* stream);

/* The REAL(4) global grids of nfield fields as the output step forms them:
 * gather_grd_mpi (common_mpi.f90:133-206) with one rank, where local value
 * i is grid point (ilon, ilat), i - 1 = (ilon - 1) + nlon (ilat - 1); then,
 * when wrap, the quasi-global copy of columns 1:2 into nlon-1:nlon
 * (lqglobal, common_io.f90:960-963; needs nlon >= 4); when file_order, the
 * vertical order of the written file (level k to nlev - k + 1,
 * write_state_vector_all, common_io.f90:655-661).  Field f, 3-D variable v,
 * level k, local value i is read at v3d[s3[0] f + s3[1] v + s3[2] k + i],
 * 2-D variable v at v2d[s2[0] f + s2[1] v + i] (strides in elements; the
 * analysis arrays anal3d(nij1,nlev,nmem,nv3d), the members being the
 * fields, have s3 = {nij1 nlev, nij1 nlev nmem, nij1}, s2 = {nij1, nij1
 * nmem}).  g3 receives (nlon,nlat,nlev,nv3d) and g2 (nlon,nlat,nv2d) per
 * field, packed.  Without file_order the grids are the v3dg/v2dg that
 * Trans_XtoY reads for H(x^a) (common_io.f90:966-989).  Runs on `stream`
 * and returns when done.  The reference has no single counterpart: it
 * gathers one field at a time over MPI. */
int letkf_grid_pack_device(int32_t real_kind, int32_t nlon, int32_t nlat,
                           int32_t nlev, int32_t nv3d, int32_t nv2d,
                           int32_t nfield, int32_t wrap, int32_t file_order,
                           const void* v3d, const int64_t* s3, const void* v2d,
                           const int64_t* s2, float* g3, float* g2,
                           void* stream);

/* Adjoint of letkf_da_main_device (no counterpart in the reference; for
 * gradient-based uses of the analysis: observation impact, tuning of the
 * inflation, of the observation errors and of the localisation,
 * torch.autograd).  DEVICE arrays; grid, obs, fcst and the handle's
 * parameters must be those of the forward call.  Given ga3/ga2 =
 * dL/d(anal3d, anal2d), writes
 *   g_dep[nobs]    dL/d obs->dep          (innovations, i.e. obs values)
 *   g_logvar[nobs] dL/d log(obs->err^2)   (per-obs log error variance)
 *   g_scalars[4]   dL/d alpha_rtpp, dL/d cov_infl_mul, dL/d sigma_obs,
 *                  dL/d sigma_obsv (the localisation scales of rloc,
 *                  common_da.f90:282-287, with the cut-offs dist_zero and
 *                  dist_zerov and so the local obs sets held fixed; 0 for
 *                  sigma_obsv <= 0).
 * The per-point eigen-decompositions are recomputed (FP64, K x K form);
 * round_off clamps pass no gradient.  Every sum is formed in a fixed
 * order (no atomics): the gradients are the same bits on every run.  Needs
 * real_kind 8, nmem <= 64 and alpha_rtps = 0.  Work is queued on `stream`
 * (NULL: the handle's stream) and the call returns after it completes. */
int letkf_da_adjoint_device(letkf_handle_t* h, const letkf_grid_t* grid,
                            const letkf_obs_t* obs, const void* fcst3d,
                            const void* fcst2d, const void* ga3,
                            const void* ga2, double* g_dep, double* g_logvar,
                            double* g_scalars, void* stream);

/* Replace the parameters of a handle (e.g. cov_infl_mul, alpha_rtpp,
 * precision, form) keeping its workspaces; sizes and real_kind must not
 * change.  No counterpart in the reference (its parameters are PARAMETERs
 * of common_setting.f90:71-131); for callers that vary them between calls. */
int letkf_set_params(letkf_handle_t* h, const letkf_params_t* p);

int letkf_get_stats(const letkf_handle_t* h, letkf_stats_t* out);

/* nobsl of every point -- count_nobsl (common_da.f90:133-205) as the last
 * analysis, plan or adjoint on the handle counted it -- into the DEVICE
 * buffer nobsl (nij1,nlev) int32, Fortran order.  The reference keeps the
 * count per point only inside da_main's loop (common_da.f90:70); this
 * exposes it for diagnostics and for checking the local-obs selection.
 * Runs on `stream` (NULL: the handle's) and returns when done. */
int letkf_get_nobsl(letkf_handle_t* h, int32_t* nobsl, void* stream);

/* One entry of the coefficient table of the batched form's matrix
 * functions (LETKF_FORM_BATCHED): Zolotarev's rational approximation
 * x^-1/2 ~ d + sum_l a[l] / (x + s[l]) on [1, 2^(cls+1)], kind 0 (the
 * nmem x nmem form, max relative error < 1e-15), 1 (the nobsl x nobsl
 * form, < 1e-16), 2 and 3 (the same forms after the FP32 reduction of
 * precision LETKF_PREC_BF16X3_TRI32, < 1e-10 and < 1e-11).  Writes d and
 * the poles' a, s (up to 40 each) and
 * returns their count (0: no entry; the class's matrices go to the
 * fallback), or a negative status.  No counterpart in the reference (it
 * uses the eigen-decomposition, common_letkf.f90:132-208); exposed for
 * tests and for callers that check the approximation. */
int letkf_rational_table_entry(int32_t kind, int32_t cls, double* d, double* a, double* s);
const char* letkf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

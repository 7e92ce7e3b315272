/*
 * sht.h -- C ABI of libsht_hip.so, the MI355X (gfx950) spherical-harmonic
 * transform on a Gaussian grid with triangular truncation.
 *
 * It replaces SPML's wa_module (linked by Dennou-CCM as -lspml-omp -lisp-fort,
 * Mkinclude:15-21 and :83-84), which DCPAM's dynamical core calls inside the
 * 'Dyn' region of atm/dcpam_main_mod.f90:2167-2203 (DynamicsHSplVAS83 at
 * :2173-2180, initialised at :5036).  SPML and ISPACK are not in the reference
 * tree; each entry point below names the SPML routine it stands in for, and the
 * in-tree call site or data layout that fixes its arguments.
 *
 * Conventions
 *   grid     xya[f][j][i]   FP64, i = 0..im-1 longitude (2 pi i/im), j = 0..jm-1
 *            Gauss latitudes south to north, f = level/field: the Fortran array
 *            xya(0:im-1, 1:jm, 1:k) of atm/dcpam_main_mod.f90:4107-4125.
 *   spectra, SPML packing (host drop-in):  wa[f][l], l = n(n+1)+m, m = -n..n,
 *            (N+1)^2 reals per field: Fortran wa((N+1)**2, k).  Coefficients of
 *            real harmonics with unit mean square over the sphere:
 *            Pt_n^0, sqrt2 Pt_n^m cos(m lon), sqrt2 Pt_n^m sin(m lon) (m<0 slot).
 *   spectra, canonical (device entry points): complex s_n^m,
 *            spec[((off(m) + n - m) * nf + f) * 2 + {0,1}], off(m) = m(N+1) - m(m-1)/2,
 *            f(lon,mu) = sum_n sum_{|m|<=n} s_n^m Pt_n^|m|(mu) e^{i m lon},
 *            Pt normalised to (1/2) int Pt^2 dmu = 1 (the zonal normalisation of
 *            exp/APEI07Couple/ebm/EBM_main.f90:170-174, :242-246).
 *            wa(n,0) = Re s_n^0, wa(n,m) = sqrt2 Re s_n^m, wa(n,-m) = -sqrt2 Im s_n^m.
 *   gradients  GradLon = (1/cos lat) df/dlon, GradLat = df/dlat (unit sphere).
 *
 * Errors: every int-returning call returns SHT_OK or a negative SHT_ERR_*; the
 * text is in sht_last_error() (thread-local).  SPML aborts through
 * MessageNotify instead (style at atm/dccm_atm_mod.f90:144-147); the Fortran
 * module in dennou-ccm_amd/fortran turns a nonzero code into an error stop.
 *
 * Threading: a plan is immutable after creation apart from its scratch; calls
 * on one plan are stream ordered and must not overlap on different streams.
 * Different plans may run concurrently on different streams.
 */
#ifndef SHT_H
#define SHT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHT_ABI_VERSION 1

enum {
  SHT_OK = 0,
  SHT_ERR_INVALID = -1,     /* bad argument or plan descriptor */
  SHT_ERR_NOMEM = -2,       /* device or host allocation failed */
  SHT_ERR_DEVICE = -3,      /* a HIP call failed */
  SHT_ERR_UNSUPPORTED = -4, /* valid but not implemented (e.g. im not a power of 2) */
  SHT_ERR_BUSY = -5         /* request server: no memory or slot free for admission */
};

/* Arithmetic of the Legendre step.  The FFT, the table generation and the
 * accumulation across 32-wide K steps are FP64 in every form; the mixed forms
 * keep the Fourier intermediate in FP32. */
enum {
  SHT_PREC_FP64 = 0,         /* FP64 MFMA (v_mfma_f64_16x16x4) */
  SHT_PREC_SPLIT_BF16 = 1,   /* BF16 hi/lo pieces, three products, FP32 step sums */
  SHT_PREC_SCALED_FP16 = 2,  /* FP16 hi/lo pieces scaled per m and per column, three products */
  SHT_PREC_SPLIT_BF16X6 = 3  /* BF16 hi/mid/lo pieces, six products */
};

enum { SHT_FLAG_GRAD = 1 }; /* also build the derivative table (GradLat) */

typedef struct sht_plan sht_plan;

typedef struct {
  int nmax;      /* triangular truncation N (gridset_nml nmax, dcpam_APEI07Couple_T21L16.conf:56-66) */
  int im;        /* longitudes (imax); a power of two, >= 3N+1 for alias-free products */
  int jm;        /* Gaussian latitudes (jmax); even, > N */
  int kmax;      /* largest number of fields per call (sizes the plan's scratch) */
  int precision; /* SHT_PREC_* */
  int flags;     /* SHT_FLAG_* */
} sht_desc;

/* ---- plan: SPML wa_Initial(n, i, j, k) / wa_Finalize (called inside
 *      DynamicsHSplVAS83Init, atm/dcpam_main_mod.f90:5036).  Gauss nodes and
 *      weights, Pt_n^m and derivative tables (FP64, computed on the GPU),
 *      FFT twiddles.  Uses the current HIP device. */
int sht_plan_create(const sht_desc *desc, sht_plan **plan);
/* (SHT_INJECT_PLAN_FAILURE=k in the environment makes the k-th device
 * allocation of every sht_plan_create fail with SHT_ERR_NOMEM, for tests of
 * the clean-up path; unset in production.) */
void sht_plan_destroy(sht_plan *plan);
const char *sht_last_error(void);
int sht_abi_version(void);

/* ---- grid description: SPML w_module xy_Lon, xy_Lat, x_Lon_Weight,
 *      y_Lat_Weight (read by tool/gmapgen/gmapgen_main.f90:296-303).
 *      Host arrays of im, jm, im, jm doubles; any pointer may be NULL.
 *      Latitudes in radians, S->N; latitude weights sum to 2. */
int sht_plan_grid(const sht_plan *plan, double *lon, double *lat,
                  double *lon_weight, double *lat_weight);

/* ---- spectral index helpers: SPML l_nm / nm_l (commented use at
 *      exp/APEI07Couple/ebm/EBM_main.f90:100, :122).  1-based l as in Fortran. */
int sht_l_nm(int n, int m);
int sht_nm_l(int l, int *n, int *m);

/* ---- device entry points, canonical spectra, stream ordered (no host sync).
 *      stream is a hipStream_t (NULL = legacy default stream).               */
/* xya_wa: spectral -> grid */
int sht_spec_to_grid_dev(sht_plan *plan, const double *spec, double *grid,
                         int nfield, void *stream);
/* xya_wa like sht_spec_to_grid_dev, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory at call time
 * (HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The Fourier intermediate goes to a static 32 MiB device array per GPU
 * (looked up at the first sht_plan_create on that device) instead of the
 * plan's scratch; calls whose intermediate (jm/2 * ceil(nfield/G) * (nmax+1)
 * * 4G values, G = max(2048/im, 8), 8 bytes each in FP64 plans, 4 in the mixed
 * forms) exceeds it fail with SHT_ERR_INVALID.  All calls on one GPU share the
 * array: calls that overlap on one GPU race.  Results equal
 * sht_spec_to_grid_dev's bitwise. */
int sht_spec_to_grid_dev_noalloc(sht_plan *plan, const double *spec, double *grid,
                                 int nfield, void *stream);
/* wa_xya: grid -> spectral */
int sht_grid_to_spec_dev(sht_plan *plan, const double *grid, double *spec,
                         int nfield, void *stream);
/* xya_wa, xya_GradLon_wa and xya_GradLat_wa of one spectral input in one pass;
 * any of the three outputs may be NULL.  GradLat needs SHT_FLAG_GRAD. */
int sht_spec_to_grad_dev(sht_plan *plan, const double *spec, double *grid,
                         double *gradlon, double *gradlat, int nfield, void *stream);
/* wa_DivLon_xya, wa_DivLat_xya and their sum (the flux divergence of DCPAM's
 * dynamics, atm/dcpam_main_mod.f90:2173-2180): spectra of
 *   (1/cos lat) dA/dlon + (1/cos lat) d(B cos lat)/dlat
 * for grid fields A and B, either of which may be NULL.  The lat term is the
 * quadrature of -B (1-mu^2) dPt/dmu / cos lat (integration by parts, exact
 * for band-limited input); it needs SHT_FLAG_GRAD.  FP64 in every form. */
int sht_grid_to_div_dev(sht_plan *plan, const double *a, const double *b, double *spec,
                        int nfield, void *stream);
/* wa_Lapla_wa (power 1), wa_LaplaInv_wa (power -1) and their powers (the
 * spectral hyperdiffusion, HDOrder in dcpam_APEI07Couple_T21L16.conf:205-230):
 * out = (-n(n+1))^power in on canonical spectra, power in [-4, 4]; negative
 * powers leave n = 0 at zero.  in == out is allowed. */
int sht_spec_lapla_dev(const sht_plan *plan, const double *in, double *out, int power,
                       int nfield, void *stream);
/* SPML real packing <-> canonical complex layout */
int sht_spml_to_canon_dev(const sht_plan *plan, const double *wa, double *spec,
                          int nfield, void *stream);
int sht_canon_to_spml_dev(const sht_plan *plan, const double *spec, double *wa,
                          int nfield, void *stream);
/* AvrLonLat_xy: global mean of each field (atm/dccm_atm_mod.f90:668-669, :774) */
int sht_avr_lonlat_dev(const sht_plan *plan, const double *grid, double *avg,
                       int nfield, void *stream);

/* ---- host drop-in, SPML packing, synchronous: wa_module's array-valued
 *      functions.  nlev <= kmax.  Arrays are host memory.                    */
int sht_xya_wa(sht_plan *plan, const double *wa, double *xya, int nlev);
int sht_wa_xya(sht_plan *plan, const double *xya, double *wa, int nlev);
int sht_xya_gradlon_wa(sht_plan *plan, const double *wa, double *xya, int nlev);
int sht_xya_gradlat_wa(sht_plan *plan, const double *wa, double *xya, int nlev);
int sht_wa_divlon_xya(sht_plan *plan, const double *xya, double *wa, int nlev);
int sht_wa_divlat_xya(sht_plan *plan, const double *xya, double *wa, int nlev);  /* SHT_FLAG_GRAD */
int sht_wa_lapla_wa(sht_plan *plan, const double *wa, double *wa_out, int nlev);     /* on the host */
int sht_wa_laplainv_wa(sht_plan *plan, const double *wa, double *wa_out, int nlev);  /* on the host */

/* ---- zonal transform: SPML l_module (l_Initial(nm, jm), y_Lat, y_Lat_Weight,
 *      l_y, y_l) as the energy-balance model uses it
 *      (exp/APEI07Couple/ebm/EBM_main.f90:64 l_initial(1024, 1024); l_y at
 *      :170 and :242).  m = 0 only:
 *        l_y:  l_n = (1/2) sum_j w_j Pt_n(mu_j) y_j,   y_l:  y_j = sum_n l_n Pt_n(mu_j),
 *      Pt_n = sqrt(2n+1) P_n.  nmax may equal jm (degree jm vanishes at every
 *      Gauss node).  Layouts: y[f][j] (jm per field, S->N), l[f][n] (nmax+1
 *      per field).  FP64 on v_mfma_f64_16x16x4f64.                          */
typedef struct sht_zonal sht_zonal;
int sht_zonal_create(int nmax, int jm, int kmax, sht_zonal **zonal);
void sht_zonal_destroy(sht_zonal *zonal);
/* y_Lat (radians, S->N) and y_Lat_Weight (sum 2); host arrays of jm, either may be NULL */
int sht_zonal_grid(const sht_zonal *zonal, double *lat, double *lat_weight);
/* device, stream ordered */
int sht_l_y_dev(sht_zonal *zonal, const double *y, double *l, int nfield, void *stream);
int sht_y_l_dev(sht_zonal *zonal, const double *l, double *y, int nfield, void *stream);
/* host, synchronous (nlev <= kmax) */
int sht_l_y(sht_zonal *zonal, const double *y, double *l, int nlev);
int sht_y_l(sht_zonal *zonal, const double *l, double *y, int nlev);

/* ---- latitude cells and the surface exchange grid of the coupler's grid
 *      generator (tool/gmapgen/gmapgen_main.f90:336-405; the same cells at
 *      exp/APEI07Couple/ebm/EBM_main.f90:200-204).  Host arithmetic over the
 *      Gauss weights of a plan (sht_plan_grid / sht_zonal_grid), S->N.      */
/* edges[0] = -pi/2, edges[j] = asin(w[j-1] + sin(edges[j-1])), edges[jm] = pi/2 */
int sht_lat_cell_edges(const double *lat_weight, int jm, double *edges);
/* Merges the cell edges of grids A and B (A's edges 0..jm_a-1, B's 1..jm_b),
 * sorts them and keeps each band of area > 1e-12 as one latitude (band
 * midpoint) with its area as weight; jm_a == jm_b returns grid A.  lat and
 * lat_weight hold up to jm_a + jm_b - 1 values; *jm_out receives the count.
 * lat_b may be NULL. */
int sht_exchange_grid(const double *lat_a, const double *w_a, int jm_a, const double *lat_b,
                      const double *w_b, int jm_b, double *lat, double *lat_weight, int *jm_out);

/* ---- restart state: RestartFileOutput / RestartFileGet of the dynamical
 *      core (atm/dcpam_main_mod.f90:3063-3066 and :4134-4137), the grid
 *      fields of the two leapfrog time levels B and N, in a netCDF classic
 *      file (64-bit offset form) written and read without a netCDF library.
 *      Host arrays; lon fastest as in Fortran: xyz(0:im-1, 1:jm, 1:km) for
 *      U, V, Temp; xyzf(0:im-1, 1:jm, 1:km, 1:nq) for QMix; xy(0:im-1, 1:jm)
 *      for Ps.  Variables UB, VB, TempB, QMixB, PsB, UN, VN, TempN, QMixN,
 *      PsN over dimensions (nq, sig, lat, lon), with coordinate variables lon,
 *      lat, sig and a global attribute time.  (gtool5's own names are not in
 *      the reference tree.)                                                 */
typedef struct {
  int im, jm, km, nq;           /* longitudes, latitudes, levels, tracers   */
  double time;                  /* restart time in seconds (write)          */
  const double *lon, *lat, *sig;/* im, jm, km coordinates (write only)      */
} sht_restart_desc;
/* fields[10] in the order UB, VB, TempB, QMixB, PsB, UN, VN, TempN, QMixN, PsN */
int sht_restart_write(const char *path, const sht_restart_desc *desc, const double *const fields[10]);
/* Reads the fields of a file whose grid matches desc's im, jm, km, nq (any
 * fields[i] may be NULL to skip it) and the restart time (time may be NULL). */
int sht_restart_read(const char *path, const sht_restart_desc *desc, double *time, double *const fields[10]);

/* ---- ensemble batches: many members of uneven truncation and level count
 *      transformed by one hipGraph launch.  There is no SPML counterpart (SPML
 *      holds one global plan); this is the batched form of wa_xya/xya_wa for
 *      ensemble runs.  Members are grouped by grid (nmax, im, jm); each group's
 *      members are consecutive field ranges of one resident batch, so every
 *      kernel launch covers all members of a group, and the groups run as
 *      independent branches of the graph.                                    */
typedef struct sht_batch sht_batch;
typedef struct {
  int nmax, im, jm; /* truncation and grid of the member */
  int nfield;       /* 2-D fields of the member (levels x variables)        */
} sht_member_desc;
enum { SHT_DIR_INVERSE = 0, SHT_DIR_FORWARD = 1, SHT_DIR_ROUNDTRIP = 2 };
/* Allocates the member state in HBM (canonical spectra in, grid, spectra out)
 * and one plan per grid with the given precision. */
int sht_batch_create(const sht_member_desc *members, int nmember, int precision,
                     sht_batch **batch);
void sht_batch_destroy(sht_batch *batch);
/* Device views of member j.  Spectra: complex s_n^m of field f at
 * spec[(k * ld + 2 f) + {0,1}], k = off(m) + n - m (ld = row stride in doubles;
 * the member's fields are columns of its group's batch).  Grid: [f][j][i].
 * Any output pointer may be NULL. */
int sht_batch_member(const sht_batch *batch, int j, double **spec_in, double **grid,
                     double **spec_out, int *ld);
/* Runs the direction (INVERSE: spec_in -> grid; FORWARD: grid -> spec_out;
 * ROUNDTRIP: both) for every member with one graph launch on stream.  The
 * graph is captured on the first call of each direction. */
int sht_batch_run(sht_batch *batch, int direction, void *stream);
/* The same work launched kernel by kernel (no graph), for comparison. */
int sht_batch_run_eager(sht_batch *batch, int direction, void *stream);
/* Autotunes the Legendre kernel shapes of every group's plan at the group's
 * field count (sht_plan_autotune; cache_path may be NULL), then drops the
 * captured graphs so the next sht_batch_run captures the tuned kernels.
 * Synchronizes the device; call it while the batch is idle. */
int sht_batch_autotune(sht_batch *batch, int reps, const char *cache_path);
/* The same with sht_plan_autotune_flags' flags. */
int sht_batch_autotune_flags(sht_batch *batch, int reps, const char *cache_path, int flags);

/* ---- request server: continuous batching of transform requests whose data
 *      are resident in HBM.  No SPML counterpart (SPML is a library called by
 *      one model); this serves many callers from resident tables.  A worker
 *      thread owns `slots` HIP streams, each with a plan per grid; whenever a
 *      slot is free it takes the pending requests of one (grid, direction) in
 *      arrival order (up to max_fields fields), gathers them into its resident
 *      batch buffers, transforms them and scatters the results, so requests
 *      that arrive while others run are admitted into the next free slot and
 *      their work overlaps on the device.  Submission is admitted against a
 *      memory budget (bytes of request data pending or in flight) and returns
 *      SHT_ERR_BUSY when it would exceed it.                                  */
typedef struct sht_server sht_server;
typedef struct {
  int ngrid;             /* grids served (<= 8)                              */
  int nmax[8], im[8], jm[8];
  int slots;             /* concurrent batches (streams), 1..8               */
  int max_fields;        /* fields per batch and per request                 */
  int precision;         /* SHT_PREC_*                                       */
  size_t mem_budget;     /* admission budget in bytes                        */
} sht_server_desc;
typedef struct {
  int nmax;              /* selects the grid                                 */
  int nfield;
  int direction;         /* SHT_DIR_INVERSE or SHT_DIR_FORWARD               */
  const double *in;      /* device: canonical spectra [K][nfield] (inverse) or grid */
  double *out;           /* device: grid (inverse) or canonical spectra      */
} sht_request;
int sht_server_create(const sht_server_desc *desc, sht_server **server);
/* Stops the worker after the requests in flight complete. */
void sht_server_destroy(sht_server *server);
/* Non-blocking; *ticket identifies the request.  SHT_ERR_BUSY: not admitted. */
int sht_server_submit(sht_server *server, const sht_request *req, long long *ticket);
/* Completed requests since the last call: tickets and submit-to-completion
 * latencies in milliseconds; *n returns how many (at most max). */
int sht_server_poll(sht_server *server, long long *tickets, double *latency_ms, int max, int *n);
/* Blocks until nothing is pending or in flight. */
int sht_server_drain(sht_server *server);
/* Counters: completed requests, fields, batches launched, requests refused. */
int sht_server_stats(const sht_server *server, long long *requests, long long *fields,
                     long long *batches, long long *refused);
/* Failed requests and the reason of the last failure (reason may be NULL).  A
 * batch that fails before it reaches the device (a launch error) fails only
 * its own requests -- sht_server_poll reports them with latency -1 -- and the
 * server goes on; a device fault seen on a batch's completion is sticky and
 * fails every later call with SHT_ERR_DEVICE.  Environment variable
 * SHT_INJECT_BATCH_FAILURE=k (read at sht_server_create) makes the k-th batch
 * fail before launch, for tests of this path. */
int sht_server_failures(const sht_server *server, long long *failed, char *reason, int reason_len);

/* ---- coupler remap: the sparse interpolation Jcup applies to exchanged
 *      fields (common/interpolation_data_latlon_mod.f90:274-306,
 *      interpolate_data_latlon):  recv(:, d) = 0;
 *      recv(r_i, d) += send(s_i, d) * coef_i for every triplet i, in triplet
 *      order.  Indices are 0-based here (the reference's are 1-based).  The
 *      GPU sums each receiving point's triplets in their original order with
 *      unfused multiply and add, so results are bitwise those of the loop
 *      evaluated without fused multiply-adds. */
typedef struct sht_remap sht_remap;
int sht_remap_create(const int *send_index, const int *recv_index, const double *coef, int nnz,
                     int nsend, int nrecv, sht_remap **remap);
void sht_remap_destroy(sht_remap *remap);
/* send [ndata][nsend], recv [ndata][nrecv] (Fortran send_data(:, d)), device. */
int sht_remap_apply_dev(const sht_remap *remap, const double *send, double *recv, int ndata,
                        void *stream);

/* ---- stage timers: HIP events around every kernel launch of a plan, summed
 *      per stage -- the counterpart of DCPAM's ProfUtil_RapStart/RapEnd
 *      regions around the dynamical core (atm/dcpam_main_mod.f90:2167, :2203).
 *      Only for eager launches (not under graph capture).                   */
enum {
  SHT_STAGE_LEG_INV = 0, /* Legendre synthesis GEMMs   */
  SHT_STAGE_FFT_INV = 1, /* ring synthesis FFT        */
  SHT_STAGE_FFT_FWD = 2, /* ring analysis FFT         */
  SHT_STAGE_LEG_FWD = 3, /* Legendre analysis GEMMs   */
  SHT_NSTAGE = 4
};
int sht_plan_set_timing(sht_plan *plan, int enable);
/* Waits for the recorded launches, returns per-stage total milliseconds and
 * launch counts (arrays of SHT_NSTAGE; either may be NULL), then resets. */
int sht_plan_stage_times(sht_plan *plan, double *total_ms, long long *launches);

/* ---- Legendre kernel shapes and their autotuner.  No SPML counterpart: the
 *      best tiling of the per-m GEMMs depends on the truncation and on the
 *      number of fields per call, which the ensemble and serving paths vary.
 *      Shapes differ in the order of the K sums, so their results agree to
 *      rounding; for one shape a result depends neither on nf nor on the
 *      field's position in the batch.                                       */
enum {
  SHT_SHAPE_AUTO = -1,        /* the tuned or set entry for nf, else the built-in choice */
  SHT_SHAPE_KSPLIT = 0,       /* workgroup per (m, 16 columns); its 4 waves split K     */
  SHT_SHAPE_WAVE = 1,         /* wavefront per 16x16 tile, whole K, pipelined loads     */
  SHT_SHAPE_TILE = 2,         /* wavefront per 16 columns x all row tiles (registers)   */
  SHT_SHAPE_TILE_SPLIT = 3,   /* TILE with the K loop halved for the heaviest m         */
  SHT_SHAPE_STAGED = 4,       /* FP64: workgroup per (m, 16 columns), data block in LDS,
                                 each wave a (parity, row tile) over the whole K; mixed
                                 inverse on grids of <= 128 latitude pairs: workgroup per
                                 (m, 8 column tiles), the table held in registers while
                                 the spectral tiles stream through LDS                   */
  SHT_SHAPE_FUSED = 5,        /* inverse of mixed plans with 64 <= im <= 512 and
                                 nmax < 192: KSPLIT's Legendre items and the ring FFTs
                                 in one persistent launch (grid values; gradients and
                                 Legendre-only launches run KSPLIT)                      */
  SHT_NSHAPE = 6
};
/* Fixes the shapes for calls with nf fields (nf = 0: the plan default for every
 * count without an entry).  FP64 plans take every shape but SHT_SHAPE_FUSED.  Mixed-precision
 * plans take SHT_SHAPE_AUTO, SHT_SHAPE_KSPLIT (workgroup per (m, 16 columns),
 * data tile shared in LDS), SHT_SHAPE_WAVE (wavefront per 16 columns, data
 * tile in registers) and SHT_SHAPE_TILE (workgroup per (m, 32 columns), each
 * wave's table fragments used for both column tiles), and for the inverse
 * SHT_SHAPE_STAGED (jm <= 256) and SHT_SHAPE_FUSED where the plan has them
 * (SHT_ERR_UNSUPPORTED otherwise); their results are bitwise equal. */
int sht_plan_set_shapes(sht_plan *plan, int nf, int inv_shape, int fwd_shape);
/* The shapes a call with nf fields would run (resolved, never SHT_SHAPE_AUTO). */
int sht_plan_get_shapes(const sht_plan *plan, int nf, int *inv_shape, int *fwd_shape);
/* Waits of SHT_SHAPE_FUSED calls on the plan that gave up: a fused call's FFT
 * tile waits at most one second for its column tile's Legendre items (a
 * broken device or counter state), then sets its grid rows to NaN and counts
 * here.  Sticky for the plan's life; reads the device synchronously.  No
 * SPML counterpart.  SHT_INJECT_FUSED_TIMEOUT=1 in the environment at plan
 * creation makes every such wait give up after 10 us (failure testing). */
int sht_plan_fused_failures(const sht_plan *plan, long long *count);
/* Wavenumber classes of SHT_SHAPE_TILE_SPLIT for calls with nf fields in one
 * direction (SHT_DIR_INVERSE or SHT_DIR_FORWARD): wavenumbers are numbered
 * heaviest first (m = 0 has the longest K loop), the loop runs split four
 * ways for m < m4, two ways for m4 <= m < m2 and unsplit above; m4 = m2 = -1
 * restores the built-in classes (two-way for the wavenumbers with at least
 * 70 % of the heaviest one's work).  The split changes the order of the K
 * sums, so results agree with the other shapes to rounding. */
int sht_plan_set_classes(sht_plan *plan, int nf, int dir, int m4, int m2);
int sht_plan_get_classes(const sht_plan *plan, int nf, int dir, int *m4, int *m2);
/* Times every shape of both Legendre kernels at nf fields on the plan's device
 * -- for FP64 plans also the split shape over a range of wavenumber classes
 * (two-way class ending after 20 % to 70 % of the wavenumbers, with no, a
 * quarter or a half four-way class) -- (reps runs of the whole direction,
 * Legendre with FFT, per candidate in three interleaved rounds, on a private
 * stream; synchronizes it) and keeps the fastest shape and classes for calls
 * with nf fields.  With cache_path not NULL, a JSON file keyed by truncation,
 * grid, nf and precision (value: the two shapes and the two class pairs) is
 * consulted first and updated after timing.  best_us (2 doubles, may be NULL) receives the winners' mean
 * direction times in microseconds (0 when taken from the cache).  It uses the plan's
 * scratch: call it while no other call on the plan is in flight. */
int sht_plan_autotune(sht_plan *plan, int nf, int reps, const char *cache_path, double *best_us);
/* The same with flags: SHT_TUNE_SHAPES_ONLY times the shapes with their
 * built-in wavenumber classes only (one choice per direction, cached under
 * the key with the suffix "_shapes"); 0 = sht_plan_autotune. */
enum { SHT_TUNE_SHAPES_ONLY = 1 };
int sht_plan_autotune_flags(sht_plan *plan, int nf, int reps, const char *cache_path, int flags,
                            double *best_us);

#ifdef __cplusplus
}
#endif
#endif /* SHT_H */

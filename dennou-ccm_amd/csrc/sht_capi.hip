// sht_capi.hip -- the extern "C" entry points of libsht_hip.so (include/sht.h).
#include <cmath>
#include <cstring>
#include <new>

#include "sht_internal.h"

namespace sht {

void set_error(const std::string &msg) { host::set_error(msg); }
int fail(int code, const std::string &msg) {
  host::set_error(msg);
  return code;
}

static int validate_desc(const sht_desc *d) {
  std::string msg;
  const int rc = host::check_plan_desc(d, msg);
  return rc ? fail(rc, msg) : SHT_OK;
}

static void release(Plan &p) {
  for (auto &t : p.timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t e : p.event_pool) (void)hipEventDestroy(e);
  void *ptrs[] = {p.mu, p.w, p.coslat, p.twid, p.Pinv, p.Hinv, p.Pfwd, p.poff_inv, p.poff_fwd,
                  p.Linv, p.Lfwd, p.loff_inv, p.loff_fwd, p.lexp_inv, p.lexp_fwd, p.four, p.four2,
                  p.stage_a, p.stage_b, p.fwd_units, p.zero, p.xcd_m[0], p.xcd_m[1],
                  p.Hfwd, p.fsync};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (p.pinned) (void)hipHostFree(p.pinned);
  for (hipEvent_t e : p.host_events)
    if (e) (void)hipEventDestroy(e);
  if (p.host_stream) (void)hipStreamDestroy(p.host_stream);
  p = Plan{};
}

static int check_call(const sht_plan *h, const void *a, const void *b, int nf, const char *who) {
  if (!h) return fail(SHT_ERR_INVALID, std::string(who) + ": plan is NULL");
  if (!a || !b) return fail(SHT_ERR_INVALID, std::string(who) + ": NULL array");
  if (nf < 1 || nf > h->p.kmax)
    return fail(SHT_ERR_INVALID, std::string(who) + ": nfield must be in [1, kmax]");
  return SHT_OK;
}

// Created on the first host call; each part only while missing, so a call
// after a failed set-up completes it rather than running without it.
static int ensure_host_staging(Plan &p) {
  const size_t grid = (size_t)p.kmax * p.jm * p.im;
  const size_t wa = (size_t)p.kmax * (p.N + 1) * (p.N + 1);
  const size_t spec = (size_t)p.kmax * p.K * 2;
  const size_t big = grid > wa ? grid : wa;
  if (!p.host_stream) SHT_HIP(hipStreamCreateWithFlags(&p.host_stream, hipStreamNonBlocking));
  if (!p.stage_a) SHT_HIP(hipMalloc(&p.stage_a, sizeof(double) * big));
  if (!p.stage_b) SHT_HIP(hipMalloc(&p.stage_b, sizeof(double) * spec));
  for (hipEvent_t &e : p.host_events)
    if (!e) SHT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!p.pinned) {
    SHT_HIP(hipHostMalloc(&p.pinned, sizeof(double) * big, hipHostMallocDefault));
    p.pinned_bytes = sizeof(double) * big;
  }
  return SHT_OK;
}

// The host copies of a drop-in call, in chunks of at least kHostChunkBytes
// (at most kHostMaxChunks) so that the PCIe copy of one chunk overlaps the host copy of
// another: in, each chunk's host -> pinned copy is followed by its async
// upload; out, the downloads are queued behind the transform and each
// chunk's pinned -> host copy starts when its download has landed.  Both
// host copies are split over worker threads (host::parallel_copy): with one
// thread the host did not keep up with the PCIe transfers on every machine
// (T85L32, 97 levels: xya_wa 1.22 -> 0.81 ms, wa_xya 0.92-1.19 -> 0.79-0.81
// ms).
constexpr size_t kHostChunkBytes = (size_t)2 << 20;
// with the copies split over threads, fewer and larger chunks pay less
// per-transfer overhead (97 levels, 25 MB grid: at most 8 / 6 / 4 / 3
// chunks, wa_xya 0.79-0.81 / 0.75 / 0.74-0.75 / 0.76-0.77 ms)
constexpr size_t kHostMaxChunks = 4;

static int host_chunks(size_t bytes) {
  const size_t n = bytes / kHostChunkBytes;
  return n < 1 ? 1 : n > kHostMaxChunks ? kHostMaxChunks : (int)n;
}

static int upload_chunked(Plan &p, const double *in, size_t n, hipStream_t s) {
  const int nc = host_chunks(sizeof(double) * n);
  for (int c = 0; c < nc; ++c) {
    const size_t a = n * c / nc, b = n * (c + 1) / nc;
    host::parallel_copy(p.pinned + a, in + a, sizeof(double) * (b - a));
    SHT_HIP(hipMemcpyAsync(p.stage_a + a, p.pinned + a, sizeof(double) * (b - a), hipMemcpyHostToDevice, s));
  }
  return SHT_OK;
}

static int download_chunked(Plan &p, double *out, size_t n, hipStream_t s) {
  const int nc = host_chunks(sizeof(double) * n);
  for (int c = 0; c < nc; ++c) {
    const size_t a = n * c / nc, b = n * (c + 1) / nc;
    SHT_HIP(hipMemcpyAsync(p.pinned + a, p.stage_a + a, sizeof(double) * (b - a), hipMemcpyDeviceToHost, s));
    SHT_HIP(hipEventRecord(p.host_events[c], s));
  }
  for (int c = 0; c < nc; ++c) {
    const size_t a = n * c / nc, b = n * (c + 1) / nc;
    SHT_HIP(hipEventSynchronize(p.host_events[c]));
    host::parallel_copy(out + a, p.pinned + a, sizeof(double) * (b - a));
  }
  return SHT_OK;
}

static hipEvent_t pooled_event(Plan &p) {
  if (!p.event_pool.empty()) {
    hipEvent_t e = p.event_pool.back();
    p.event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

thread_local StageTimer *tl_stage_timer = nullptr;

// Runs one launcher; with timing on, its kernels record the stage's events
// themselves (launch_kernel), from the first kernel's start to the last's end.
template <class F>
static int timed(Plan &p, int stage, hipStream_t s, F &&launch) {
  if (!p.timing) return launch();
  StageTimer t;
  t.start = pooled_event(p);
  t.stop = pooled_event(p);
  if (!t.start || !t.stop) return fail(SHT_ERR_DEVICE, "stage timer: hipEventCreate failed");
  tl_stage_timer = &t;
  int rc = launch();
  tl_stage_timer = nullptr;
  (void)s;
  if (t.used) {
    p.timed.push_back({stage, t.start, t.stop});
  } else {
    p.event_pool.push_back(t.start);
    p.event_pool.push_back(t.stop);
  }
  return rc;
}

static int spec_to_grid(Plan &p, const double *spec, double *grid, int nf, hipStream_t s) {
  // the fused shape: one launch, timed as the Legendre stage
  if (p.prec != SHT_PREC_FP64 && resolve_shape(p, nf, 0) == SHT_SHAPE_FUSED)
    return timed(p, SHT_STAGE_LEG_INV, s, [&] {
      return launch_inv_fused(p, spec, reinterpret_cast<float *>(p.four), grid, nf, s);
    });
  int rc = timed(p, SHT_STAGE_LEG_INV, s, [&] {
    return p.prec == SHT_PREC_FP64
               ? launch_leg_inv_f64(p, spec, p.four, p.four2, true, false, nf, s)
               : launch_leg_inv_mixed(p, spec, reinterpret_cast<float *>(p.four), nf, s);
  });
  if (rc) return rc;
  return timed(p, SHT_STAGE_FFT_INV, s, [&] {
    return launch_fft_inv(p, p.four, nullptr, p.prec != SHT_PREC_FP64, grid, nullptr, nullptr, nf, s);
  });
}

static int grid_to_spec(Plan &p, const double *grid, double *spec, int nf, hipStream_t s) {
  int rc = timed(p, SHT_STAGE_FFT_FWD, s, [&] {
    return launch_fft_fwd(p, grid, p.four, p.prec != SHT_PREC_FP64, nf, s);
  });
  if (rc) return rc;
  return timed(p, SHT_STAGE_LEG_FWD, s, [&] {
    return p.prec == SHT_PREC_FP64
               ? launch_leg_fwd_f64(p, p.four, spec, nf, s)
               : launch_leg_fwd_mixed(p, reinterpret_cast<const float *>(p.four), spec, nf, s);
  });
}

// spectral -> grid with the Fourier intermediate at `four` (the static
// scratch of sht_spec_to_grid_dev_noalloc) instead of the plan's; the same
// kernels as spec_to_grid, without stage timers (they create events).
int spec_to_grid_into(Plan &p, const double *spec, double *grid, int nf, void *four, hipStream_t s) {
  if (p.prec != SHT_PREC_FP64 && resolve_shape(p, nf, 0) == SHT_SHAPE_FUSED)
    return launch_inv_fused(p, spec, static_cast<float *>(four), grid, nf, s);
  int rc = p.prec == SHT_PREC_FP64
               ? launch_leg_inv_f64(p, spec, static_cast<double *>(four), nullptr, true, false, nf, s)
               : launch_leg_inv_mixed(p, spec, static_cast<float *>(four), nf, s);
  if (rc) return rc;
  return launch_fft_inv(p, four, nullptr, p.prec != SHT_PREC_FP64, grid, nullptr, nullptr, nf, s);
}

// Gradients always run the FP64 Legendre step (the derivative table is FP64 only).
static int spec_to_grad(Plan &p, const double *spec, double *grid, double *glon, double *glat,
                        int nf, hipStream_t s) {
  const bool value_tab = grid || glon;
  const bool deriv_tab = glat != nullptr;
  int rc = timed(p, SHT_STAGE_LEG_INV, s, [&] {
    return launch_leg_inv_f64(p, spec, p.four, p.four2, value_tab, deriv_tab, nf, s);
  });
  if (rc) return rc;
  return timed(p, SHT_STAGE_FFT_INV, s, [&] {
    return launch_fft_inv(p, p.four, p.four2, false, grid, glon, glat, nf, s);
  });
}

// Flux divergence (1/cos) dA/dlon + (1/cos) d(B cos)/dlat to spectra, FP64 in
// every precision form: A's Fourier coefficients times i m / cos meet the
// value table, B's times 1 / cos meet -(w/2)(1-mu^2) dPt/dmu (the lat term
// integrated by parts) with the parity slots swapped; the second GEMM adds
// into the first's output.
static int grid_to_div(Plan &p, const double *a, const double *b, double *spec, int nf,
                       hipStream_t s) {
  double *fa = a ? p.four : nullptr;
  double *fb = b ? (a ? p.four2 : p.four) : nullptr;
  int rc = timed(p, SHT_STAGE_FFT_FWD, s, [&] { return launch_fft_fwd_div(p, a, fa, b, fb, nf, s); });
  if (rc) return rc;
  return timed(p, SHT_STAGE_LEG_FWD, s, [&] {
    int r = SHT_OK;
    if (a) r = launch_leg_fwd_f64_table(p, p.Pfwd, fa, spec, false, false, nf, s);
    if (!r && b) r = launch_leg_fwd_f64_table(p, p.Hfwd, fb, spec, true, a != nullptr, nf, s);
    return r;
  });
}

// One direction of the transform on a plan (used by the ensemble batches).
int run_dir_on(sht_plan *plan, const double *spec_in, double *grid, double *spec_out, int nf,
               int dir, hipStream_t s) {
  Plan &p = plan->p;
  int rc = SHT_OK;
  if (dir == SHT_DIR_INVERSE || dir == SHT_DIR_ROUNDTRIP) rc = spec_to_grid(p, spec_in, grid, nf, s);
  if (!rc && (dir == SHT_DIR_FORWARD || dir == SHT_DIR_ROUNDTRIP))
    rc = grid_to_spec(p, grid, spec_out, nf, s);
  return rc;
}

}  // namespace sht

using namespace sht;

extern "C" {

const char *sht_last_error(void) { return host::last_error(); }
int sht_abi_version(void) { return SHT_ABI_VERSION; }

int sht_plan_create(const sht_desc *desc, sht_plan **out) {
  if (!out) return fail(SHT_ERR_INVALID, "sht_plan_create: out is NULL");
  *out = nullptr;
  int rc = validate_desc(desc);
  if (rc) return rc;
  sht_plan *h = new (std::nothrow) sht_plan;
  if (!h) return fail(SHT_ERR_NOMEM, "sht_plan_create: host allocation failed");
  Plan &p = h->p;
  p.N = desc->nmax;
  p.im = desc->im;
  p.jm = desc->jm;
  p.P2 = desc->jm / 2;
  p.P2p = roundup(p.P2, 16);
  p.K = (p.N + 1) * (p.N + 2) / 2;
  p.kmax = desc->kmax;
  p.prec = desc->precision;
  p.flags = desc->flags;
  p.maxke = ceil_half(p.N + 1);
  p.alloc_fail = host::injected_count("SHT_INJECT_PLAN_FAILURE");
  p.fused_inject = host::injected_count("SHT_INJECT_FUSED_TIMEOUT") > 0;
  hipError_t e = hipGetDevice(&p.device);
  if (e != hipSuccess) {
    delete h;
    return fail(SHT_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  }
  // the static scratch of sht_spec_to_grid_dev_noalloc on this device; a
  // failed lookup leaves that entry point reporting it, not plan creation
  (void)noalloc_init_device(p.device);
  rc = build_tables(p);
  if (!rc) fft_fwd_capacity(p, p.fft_fwd_resident);
  if (!rc && p.prec != SHT_PREC_FP64) rc = build_mixed_tables(p);
  if (!rc) {
    // [P2][nfg][N+1][2][2G]: field groups (G = max(2048/im, 8)) are allocated whole
    const size_t four = (size_t)p.P2 * (p.N + 1) * 4 * (size_t)roundup(p.kmax, layout_group(fft_group(p.im)));
    rc = plan_alloc(p, reinterpret_cast<void **>(&p.four), sizeof(double) * four);
    if (!rc && (p.flags & SHT_FLAG_GRAD)) rc = plan_alloc(p, reinterpret_cast<void **>(&p.four2), sizeof(double) * four);
  }
  if (rc) {
    std::string keep = host::last_error();
    release(p);
    delete h;
    host::set_error(keep);
    return rc;
  }
  *out = h;
  return SHT_OK;
}

int sht_plan_fused_failures(const sht_plan *h, long long *count) {
  if (!h || !count) return fail(SHT_ERR_INVALID, "sht_plan_fused_failures: NULL argument");
  return fused_wait_failures(h->p, count);
}

void sht_plan_destroy(sht_plan *h) {
  if (!h) return;
  release(h->p);
  delete h;
}

int sht_plan_grid(const sht_plan *h, double *lon, double *lat, double *lon_weight,
                  double *lat_weight) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_grid: plan is NULL");
  const Plan &p = h->p;
  for (int i = 0; i < p.im; ++i) {
    if (lon) lon[i] = 2.0 * M_PI * (double)i / (double)p.im;
    if (lon_weight) lon_weight[i] = 2.0 * M_PI / (double)p.im;
  }
  for (int j = 0; j < p.jm; ++j) {
    if (lat) lat[j] = atan2(p.h_mu[j], p.h_coslat[j]);
    if (lat_weight) lat_weight[j] = p.h_w[j];
  }
  return SHT_OK;
}

int sht_plan_set_timing(sht_plan *h, int enable) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_set_timing: plan is NULL");
  h->p.timing = enable != 0;
  return SHT_OK;
}

int sht_plan_stage_times(sht_plan *h, double *total_ms, long long *launches) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_stage_times: plan is NULL");
  Plan &p = h->p;
  double ms[SHT_NSTAGE] = {0, 0, 0, 0};
  long long cnt[SHT_NSTAGE] = {0, 0, 0, 0};
  int rc = SHT_OK;
  for (auto &t : p.timed) {
    float e = 0.f;
    if (rc == SHT_OK) {
      hipError_t st = hipEventSynchronize(t.b);
      if (st == hipSuccess) st = hipEventElapsedTime(&e, t.a, t.b);
      if (st != hipSuccess)
        rc = fail(SHT_ERR_DEVICE, std::string("sht_plan_stage_times: ") + hipGetErrorString(st));
    }
    ms[t.stage] += e;
    cnt[t.stage] += 1;
    p.event_pool.push_back(t.a);
    p.event_pool.push_back(t.b);
  }
  p.timed.clear();
  for (int i = 0; i < SHT_NSTAGE; ++i) {
    if (total_ms) total_ms[i] = ms[i];
    if (launches) launches[i] = cnt[i];
  }
  return rc;
}

int sht_l_nm(int n, int m) {
  const int l = host::l_of_nm(n, m);
  if (l < 1) return fail(SHT_ERR_INVALID, "sht_l_nm: need 0 <= |m| <= n <= 46339");
  return l;
}

int sht_nm_l(int l, int *n, int *m) {
  if (!n || !m || !host::nm_of_l(l, *n, *m)) return fail(SHT_ERR_INVALID, "sht_nm_l: need l >= 1");
  return SHT_OK;
}

int sht_spec_to_grid_dev(sht_plan *h, const double *spec, double *grid, int nf, void *stream) {
  int rc = check_call(h, spec, grid, nf, "sht_spec_to_grid_dev");
  if (rc) return rc;
  return spec_to_grid(h->p, spec, grid, nf, (hipStream_t)stream);
}

int sht_grid_to_spec_dev(sht_plan *h, const double *grid, double *spec, int nf, void *stream) {
  int rc = check_call(h, grid, spec, nf, "sht_grid_to_spec_dev");
  if (rc) return rc;
  return grid_to_spec(h->p, grid, spec, nf, (hipStream_t)stream);
}

int sht_spec_to_grad_dev(sht_plan *h, const double *spec, double *grid, double *gradlon,
                         double *gradlat, int nf, void *stream) {
  double *any = grid ? grid : (gradlon ? gradlon : gradlat);
  int rc = check_call(h, spec, any, nf, "sht_spec_to_grad_dev");
  if (rc) return rc;
  if (gradlat && !(h->p.flags & SHT_FLAG_GRAD))
    return fail(SHT_ERR_INVALID, "sht_spec_to_grad_dev: GradLat needs a plan built with SHT_FLAG_GRAD");
  return spec_to_grad(h->p, spec, grid, gradlon, gradlat, nf, (hipStream_t)stream);
}

int sht_grid_to_div_dev(sht_plan *h, const double *a, const double *b, double *spec, int nf,
                        void *stream) {
  int rc = check_call(h, a ? a : b, spec, nf, "sht_grid_to_div_dev");
  if (rc) return rc;
  if (b && !(h->p.flags & SHT_FLAG_GRAD))
    return fail(SHT_ERR_INVALID, "sht_grid_to_div_dev: the lat term needs a plan built with SHT_FLAG_GRAD");
  return grid_to_div(h->p, a, b, spec, nf, (hipStream_t)stream);
}

int sht_spec_lapla_dev(const sht_plan *h, const double *in, double *out, int power, int nf,
                       void *stream) {
  int rc = check_call(h, in, out, nf, "sht_spec_lapla_dev");
  if (rc) return rc;
  if (power < -4 || power > 4) return fail(SHT_ERR_INVALID, "sht_spec_lapla_dev: power must be in [-4, 4]");
  return launch_spec_lapla(h->p, in, out, power, nf, (hipStream_t)stream);
}

int sht_spml_to_canon_dev(const sht_plan *h, const double *wa, double *spec, int nf, void *stream) {
  int rc = check_call(h, wa, spec, nf, "sht_spml_to_canon_dev");
  if (rc) return rc;
  return launch_spml_to_canon(h->p, wa, spec, nf, (hipStream_t)stream);
}

int sht_canon_to_spml_dev(const sht_plan *h, const double *spec, double *wa, int nf, void *stream) {
  int rc = check_call(h, spec, wa, nf, "sht_canon_to_spml_dev");
  if (rc) return rc;
  return launch_canon_to_spml(h->p, spec, wa, nf, (hipStream_t)stream);
}

int sht_avr_lonlat_dev(const sht_plan *h, const double *grid, double *avg, int nf, void *stream) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_avr_lonlat_dev: plan is NULL");
  if (!grid || !avg || nf < 1) return fail(SHT_ERR_INVALID, "sht_avr_lonlat_dev: bad arguments");
  return launch_avr_lonlat(h->p, grid, avg, nf, (hipStream_t)stream);
}

// ---- host drop-in (SPML packing, synchronous)
enum HostOp { kXyaWa, kWaXya, kGradLon, kGradLat, kDivLon, kDivLat };

static int host_call(sht_plan *h, const double *in, double *out, int nlev, HostOp op,
                     const char *who) {
  int rc = check_call(h, in, out, nlev, who);
  if (rc) return rc;
  Plan &p = h->p;
  if ((op == kGradLat || op == kDivLat) && !(p.flags & SHT_FLAG_GRAD))
    return fail(SHT_ERR_INVALID, std::string(who) + ": needs a plan built with SHT_FLAG_GRAD");
  rc = ensure_host_staging(p);
  if (rc) return rc;
  hipStream_t s = p.host_stream;
  const size_t ngrid = (size_t)nlev * p.jm * p.im;
  const size_t nwa = (size_t)nlev * (p.N + 1) * (p.N + 1);
  const bool to_spec = op == kWaXya || op == kDivLon || op == kDivLat;
  const size_t nin = to_spec ? ngrid : nwa;
  const size_t nout = to_spec ? nwa : ngrid;
  rc = upload_chunked(p, in, nin, s);
  if (!rc) switch (op) {
    case kXyaWa:
      if (!(rc = launch_spml_to_canon(p, p.stage_a, p.stage_b, nlev, s)))
        rc = spec_to_grid(p, p.stage_b, p.stage_a, nlev, s);
      break;
    case kGradLon:
      if (!(rc = launch_spml_to_canon(p, p.stage_a, p.stage_b, nlev, s)))
        rc = spec_to_grad(p, p.stage_b, nullptr, p.stage_a, nullptr, nlev, s);
      break;
    case kGradLat:
      if (!(rc = launch_spml_to_canon(p, p.stage_a, p.stage_b, nlev, s)))
        rc = spec_to_grad(p, p.stage_b, nullptr, nullptr, p.stage_a, nlev, s);
      break;
    case kWaXya:
      if (!(rc = grid_to_spec(p, p.stage_a, p.stage_b, nlev, s)))
        rc = launch_canon_to_spml(p, p.stage_b, p.stage_a, nlev, s);
      break;
    case kDivLon:
    case kDivLat: {
      const double *a = op == kDivLon ? p.stage_a : nullptr;
      const double *b = op == kDivLat ? p.stage_a : nullptr;
      if (!(rc = grid_to_div(p, a, b, p.stage_b, nlev, s)))
        rc = launch_canon_to_spml(p, p.stage_b, p.stage_a, nlev, s);
      break;
    }
  }
  if (rc) {
    // uploads may still read the pinned buffer: let them finish before the
    // next call refills it
    (void)hipStreamSynchronize(s);
    return rc;
  }
  return download_chunked(p, out, nout, s);
}

int sht_xya_wa(sht_plan *h, const double *wa, double *xya, int nlev) {
  return host_call(h, wa, xya, nlev, kXyaWa, "sht_xya_wa");
}
int sht_wa_xya(sht_plan *h, const double *xya, double *wa, int nlev) {
  return host_call(h, xya, wa, nlev, kWaXya, "sht_wa_xya");
}
int sht_xya_gradlon_wa(sht_plan *h, const double *wa, double *xya, int nlev) {
  return host_call(h, wa, xya, nlev, kGradLon, "sht_xya_gradlon_wa");
}
int sht_xya_gradlat_wa(sht_plan *h, const double *wa, double *xya, int nlev) {
  return host_call(h, wa, xya, nlev, kGradLat, "sht_xya_gradlat_wa");
}
int sht_wa_divlon_xya(sht_plan *h, const double *xya, double *wa, int nlev) {
  return host_call(h, xya, wa, nlev, kDivLon, "sht_wa_divlon_xya");
}
int sht_wa_divlat_xya(sht_plan *h, const double *xya, double *wa, int nlev) {
  return host_call(h, xya, wa, nlev, kDivLat, "sht_wa_divlat_xya");
}

// Diagonal operators on the SPML packing, applied on the host (no transfer):
// slot l = n(n+1)+m+1 scales by (-n(n+1))^power, n = 0 to zero for power < 0.
static int host_lapla(sht_plan *h, const double *in, double *out, int nlev, int power,
                      const char *who) {
  int rc = check_call(h, in, out, nlev, who);
  if (rc) return rc;
  const int N = h->p.N, L = (N + 1) * (N + 1);
  for (int n = 0; n <= N; ++n) {
    const double e = -(double)n * (n + 1);
    const double fac = power < 0 ? (n == 0 ? 0.0 : 1.0 / e) : e;
    for (int f = 0; f < nlev; ++f)
      for (int l = n * n; l < (n + 1) * (n + 1); ++l) out[(size_t)f * L + l] = fac * in[(size_t)f * L + l];
  }
  return SHT_OK;
}
int sht_wa_lapla_wa(sht_plan *h, const double *wa, double *out, int nlev) {
  return host_lapla(h, wa, out, nlev, 1, "sht_wa_lapla_wa");
}
int sht_wa_laplainv_wa(sht_plan *h, const double *wa, double *out, int nlev) {
  return host_lapla(h, wa, out, nlev, -1, "sht_wa_laplainv_wa");
}

}  // extern "C"

// sht_internal.h -- plan layout, error plumbing and kernel launchers shared by
// the translation units of libsht_hip.so (gfx950 only).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

// The public entry points are the only exported symbols (-fvisibility=hidden).
#pragma GCC visibility push(default)
#include "../../include/sht.h"
#pragma GCC visibility pop
#include "sht_host.h"

namespace sht {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SHT_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::sht::fail(SHT_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ layout
// Device tables of one plan.  P2 = jm/2 latitude pairs; pair p couples the
// north node jn = P2 + p and the south node js = P2 - 1 - p (mu_js = -mu_jn).
// Table rows over pairs are padded to P2p = roundup(P2, 16) with zeros.
// Within one m the degrees split by parity of n - m: k-th even row n = m + 2k,
// k-th odd row n = m + 2k + 1.
//
//   Pinv  per m: [par 2][KP(m)][P2p]  Pt_n^m(mu_p)          (inverse, p fastest)
//   Hinv  per m: [par 2][KP(m)][P2p]  (1-mu^2) dPt/dmu      (inverse GradLat)
//   Hfwd  per m: [par 2][P2p][KR(m)]  -(w_p/2) (1-mu^2) dPt/dmu (forward DivLat)
//   Pfwd  per m: [par 2][P2p][KR(m)]  (w_p/2) Pt_n^m(mu_p)  (forward, k fastest)
//   KP(m) = roundup(ceil((N-m+1)/2), 4), KR(m) = roundup(ceil((N-m+1)/2), 16),
//   padding entries are zero.
//
//   four  [P2][nfg][N+1][par 2][2G]  the Fourier intermediate: per latitude pair
//         and wavenumber the equatorially even/odd halves, complex with the
//         fields of one group innermost (G = max(2048/im, 8) fields, nfg =
//         ceil(nf/G) groups), so each FFT block reads or writes one contiguous,
//         aligned slice (or, on grids above 256 longitudes, its part of each
//         run).  Forward: (F(mu)+F(-mu), F(mu)-F(-mu)); inverse: the even-n and
//         odd-n Legendre sums (north = even + odd, south = even - odd).
//
// Mixed-precision plans also hold the tables split into 16-bit pieces in the
// matrix-core operand layout (sht_legendre_mixed.hip) and use an FP32 Fourier
// intermediate.

struct Plan {
  int N = 0, im = 0, jm = 0, P2 = 0, P2p = 0, K = 0, kmax = 0, prec = 0, flags = 0;
  int device = 0;
  int fft_fwd_resident[2] = {0, 0};  // forward FFT blocks the device holds at once, big / small (0: unknown)
  int maxke = 0;            // ceil((N+1)/2): most rows of one parity
  // device tables
  double *mu = nullptr, *w = nullptr, *coslat = nullptr; // jm each, S->N
  double *twid = nullptr;                                // im complex: e^{2 pi i t/im}
  double *zero = nullptr;   // 64 zeros: the source of masked-off operand loads
  double *Pinv = nullptr, *Hinv = nullptr, *Pfwd = nullptr;
  double *Hfwd = nullptr;  // -(w/2) (1-mu^2) dPt/dmu, Pfwd layout (divergence; GRAD plans)
  int *poff_inv = nullptr, *poff_fwd = nullptr;          // N+1 offsets (doubles)
  std::vector<int> h_poff_inv, h_poff_fwd;
  size_t n_inv = 0, n_fwd = 0;
  // mixed-precision tables
  void *Linv = nullptr, *Lfwd = nullptr;
  int *loff_inv = nullptr, *loff_fwd = nullptr;
  int *lexp_inv = nullptr, *lexp_fwd = nullptr; // FP16 power-of-two per (m, parity)
  size_t n_linv = 0, n_lfwd = 0;
  // wave-tile forward Legendre: (m << 8 | 16-degree row tile), heaviest m first
  int *fwd_units = nullptr;
  int n_fwd_units = 0;
  // Wavenumbers dealt to the 8 XCDs, [xcd][mlen] with -1 padding, balanced by
  // the work of each m (greedy, heaviest first): one list for the inverse
  // (work ~ K steps), one for the forward (work ~ 16-degree row tiles).
  int *xcd_m[2] = {nullptr, nullptr};
  int xcd_mlen[2] = {0, 0};
  // Fused inverse (mixed forms): its queue counters and per-column-tile
  // completion counts ((kFusedSyncHead + column tiles) slots of kSyncStride
  // ints, zero between calls), and the workgroups the device holds at once
  // (0: not available)
  int *fsync = nullptr;
  int fused_blocks = 0;
  int fused_inject = 0;  // SHT_INJECT_FUSED_TIMEOUT: every fused wait times out
  // Legendre kernel shapes (SHT_SHAPE_*): the default for every field count,
  // and entries fixed by sht_plan_set_shapes / sht_plan_autotune for one count
  // and the wavenumber classes of the split shape (host::Tuning: K loop split
  // four ways for m < m4, two ways for m4 <= m < m2; -1 = built-in classes)
  int shape_inv = SHT_SHAPE_AUTO, shape_fwd = SHT_SHAPE_AUTO;
  struct ShapeEntry {
    int nf;
    host::Tuning t;
  };
  std::vector<ShapeEntry> shapes;
  // scratch
  double *four = nullptr, *four2 = nullptr; // [P2][N+1][2][2*kmax]
  double *stage_a = nullptr, *stage_b = nullptr; // device staging for host calls
  double *pinned = nullptr;                      // pinned host staging
  size_t pinned_bytes = 0;
  hipStream_t host_stream = nullptr;
  hipEvent_t host_events[8] = {};                // chunks of the host result copy
  // host copies of the grid
  std::vector<double> h_mu, h_w, h_coslat;
  // stage timers (sht_plan_set_timing)
  bool timing = false;
  struct TimedLaunch { int stage; hipEvent_t a, b; };
  std::vector<TimedLaunch> timed;
  std::vector<hipEvent_t> event_pool;
  // SHT_INJECT_PLAN_FAILURE: the alloc_fail-th device allocation of this plan fails
  int alloc_fail = 0, alloc_count = 0;
};

// Device allocation of plan state: SHT_ERR_NOMEM on failure, or when the
// injected failure index is reached.
inline int plan_alloc(Plan &p, void **ptr, size_t bytes) {
  if (p.alloc_fail && ++p.alloc_count == p.alloc_fail)
    return fail(SHT_ERR_NOMEM, "sht_plan_create: injected allocation failure (SHT_INJECT_PLAN_FAILURE)");
  const hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail(SHT_ERR_NOMEM, std::string("sht_plan_create: hipMalloc: ") + hipGetErrorString(e));
  return SHT_OK;
}
#define SHT_PLAN_ALLOC(p, ptr, bytes)                                              \
  do {                                                                             \
    const int rc_ = ::sht::plan_alloc((p), reinterpret_cast<void **>(ptr), (bytes)); \
    if (rc_) return rc_;                                                           \
  } while (0)

__host__ __device__ inline int roundup(int a, int b) { return (a + b - 1) / b * b; }

// Fourier scratch geometry: G = fields per FFT workgroup, cgs = 2G columns per
// group (a power of two), column c = 2*field + {re, im}.
struct FourGeom {
  int M1, nfg, lg;  // lg = log2(cgs)
  int stream = 0;   // writers should use non-temporal stores (see four_geom_out)
  __host__ __device__ size_t at(int p, int m, int par, int c) const {
    return ((((size_t)p * nfg + (c >> lg)) * M1 + m) * 2 + par) * ((size_t)1 << lg) + (c & ((1 << lg) - 1));
  }
};
// Store of the Fourier intermediate; NT (non-temporal) is a template
// argument of the writing kernels because a run-time choice between the two
// stores is merged by the compiler into one plain store.
template <bool NT, typename T>
__device__ __forceinline__ void put_four(T *dst, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, dst);
  else
    *dst = v;
}
// Store of forward Legendre results (spectra, rows of 2 nf doubles):
// streaming.  A workgroup writes 16-column strips of its rows; the row pitch
// is no multiple of a 128-byte line, so a strip covers parts of two lines
// and HBM receives partial lines (T85, 97 fields: 7.7-8.0 MB written per
// dispatch for 5.81 MB of spectra).  Regular write-back stores let the XCD's
// L2 join the parts (all column strips of one m run on one XCD) and write
// exactly 5.81 MB, but the write-back at the kernel's end delays the next
// launch: the T85 round trip measured 40.15 against 39.40 us (split-BF16
// x6) and 46.45 against 46.17 us (FP64), so the stores stream.
__device__ __forceinline__ void put_spec(double *dst, double v) { __builtin_nontemporal_store(v, dst); }
// FFT blocks: a workgroup transforms C complex values, G = C / im ring pairs
// of one latitude pair, with C / 8 threads.  On grids of at most
// kFftSmallMaxIm longitudes, calls that give the big blocks fewer than
// kFftSmallBelow workgroups take the small ones (one wave), which hold the
// Fourier intermediate in field groups of a quarter the size (fft_block).
// Measured on MI355X: T21L16 (64 longitudes, 49 fields: 32 big blocks) round
// trip 15.2 -> 13.9 us; T42 (128) unchanged; T85 (256) with 1-8 fields
// 0.3-0.6 us slower per FFT launch (a one-wave block hides less latency and
// still runs whole radix-16 butterflies per thread).
constexpr int kFftBig = 2048, kFftSmall = 512;
constexpr int kFftSmallMaxIm = 64, kFftSmallBelow = 256;
inline int fft_group(int im) { return kFftBig / im; }  // the big blocks' field group
// Field groups of the Fourier intermediate: the FFT block's G fields, but at
// least kLayoutMinGroup, so that a (pair, m, parity) run holds at least 16
// values (64 bytes in FP32) and the Legendre kernels, which write and read
// across wavenumbers, move whole 64-byte pieces.  On grids of 512 to 2048
// longitudes (G = 4, 2, 1) an FFT block then covers 1 / KS of a group.
// Measured at T170 (512 longitudes), 776 fields, split-BF16 x6, groups of 8
// fields instead of 4: inverse Legendre 214 -> 185-189 us, forward Legendre
// 187 -> 173-175 us, forward FFT 231 -> 238-241 us (its blocks store half
// runs), round trip 896 -> 863-866 us.
constexpr int kLayoutMinGroup = 8;
__host__ __device__ constexpr int layout_group(int G) { return G >= kLayoutMinGroup ? G : kLayoutMinGroup; }
inline int fft_block(const Plan &p, int nf) {
  const int g = kFftBig / p.im;
  return p.im <= kFftSmallMaxIm && (long)p.P2 * ((nf + g - 1) / g) < kFftSmallBelow ? kFftSmall : kFftBig;
}
// the layout for FFT blocks of G fields (groups of layout_group(G) fields)
inline FourGeom four_geom_g(int N, int G, int nf) {
  const int GL = layout_group(G);
  int lg = 0;
  while ((1 << lg) < 2 * GL) ++lg;
  return FourGeom{N + 1, (nf + GL - 1) / GL, lg};
}
// the layout a call with nf fields uses (its FFT block's field groups)
inline FourGeom four_geom(const Plan &p, int nf) { return four_geom_g(p.N, fft_block(p, nf) / p.im, nf); }
// Geometry for the Legendre inverse, which writes the Fourier intermediate:
// non-temporal stores up to kStreamFourBytes of it, regular stores above.
// Measured on MI355X at T85 (97 to 388 fields): streaming saves the write-back
// of the dirty L2 lines when the kernel ends (2.6 us at 8.5 MB), while above
// ~25 MB the following FFT reads the regularly stored intermediate from the
// Infinity Cache 6 us faster than the streamed one.
constexpr size_t kStreamFourBytes = (size_t)24 << 20;
inline FourGeom four_geom_out(const Plan &p, int nf, size_t elem_bytes, int G = 0) {
  FourGeom g = G ? four_geom_g(p.N, G, nf) : four_geom(p, nf);
  const size_t bytes = (size_t)p.P2 * g.nfg * g.M1 * 2 * ((size_t)1 << g.lg) * elem_bytes;
  g.stream = bytes <= kStreamFourBytes;
  return g;
}
__host__ __device__ inline int ceil_half(int a) { return (a + 1) / 2; }
__host__ __device__ inline int64_t coef_off(int N, int m) {
  return (int64_t)m * (N + 1) - (int64_t)m * (m - 1) / 2;
}

// ------------------------------------------------------------------ kernel launch
// Stage timing: while a timer is open on the calling thread, transform kernels
// are launched through hipExtLaunchKernel with events that the runtime records
// at the kernel's own start (first launch of the stage) and end (every launch;
// the last one stands), so a stage's time is its kernels' execution, not the
// launch overhead that events recorded around the launch would add.
struct StageTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  bool used = false;
};
extern thread_local StageTimer *tl_stage_timer;

template <typename... KArgs, typename... Args>
inline void launch_kernel(void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem,
                          hipStream_t s, Args... args) {
  if (StageTimer *t = tl_stage_timer) {
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)shmem, s, t->used ? nullptr : t->start,
                          t->stop, 0u, args...);
    t->used = true;
  } else {
    kernel<<<grid, block, shmem, s>>>(args...);
  }
}

// ------------------------------------------------------------------ launchers
// Table construction (sht_tables.hip)
int build_tables(Plan &p);
// Gauss-Legendre nodes mu (S->N), weights and cos(lat) of jm (even) latitudes
int launch_gauss(int jm, double *mu, double *w, double *coslat, hipStream_t s);

// FFT stage (sht_fft.hip).  mode: 0 value, 1 GradLon (x i m / cos), 2 GradLat (x 1/cos).
// f32 = the Fourier intermediate is FP32 (mixed-precision plans); the FFT itself is FP64.
// Forward FFT blocks one device holds at once (occupancy x CUs; 0 if unknown).
void fft_fwd_capacity(const Plan &p, int (&out)[2]);
int launch_fft_fwd(const Plan &p, const double *grid, void *four, bool f32, int nf, hipStream_t s);
// Divergence analysis FFT: a -> fa with x i m / cos, b -> fb with x 1 / cos (FP64;
// either pair may be NULL), one launch.
int launch_fft_fwd_div(const Plan &p, const double *a, double *fa, const double *b, double *fb,
                       int nf, hipStream_t s);
int launch_fft_inv(const Plan &p, const void *four, const void *four2, bool f32, double *grid,
                   double *gradlon, double *gradlat, int nf, hipStream_t s);

// Legendre stage, FP64 MFMA (sht_legendre.hip).  Shapes resolve through
// resolve_shape: an entry for nf, else the plan default, else the built-in one.
int resolve_shape(const Plan &p, int nf, int dir);  // dir 0 inverse, 1 forward
// Wavenumber classes of the split shape (SHT_SHAPE_TILE_SPLIT), numbered
// heaviest first: the K loop runs split four ways for m < m4, two ways for
// m4 <= m < m2, unsplit for m >= m2.
struct KSplit {
  int m4, m2;
};
// The classes a call with nf fields uses: an entry's, else the built-in ones.
KSplit resolve_classes(const Plan &p, int nf, int dir);
KSplit default_classes(const Plan &p, int dir);
int launch_leg_inv_f64_shape(const Plan &p, int shape, const double *spec, double *four,
                             double *four2, bool value_table, bool deriv_table, int nf, hipStream_t s,
                             KSplit cls = KSplit{-1, -1});
int launch_leg_fwd_f64_shape(const Plan &p, int shape, const double *four, double *spec, int nf,
                             hipStream_t s, KSplit cls = KSplit{-1, -1});
int launch_leg_inv_f64(const Plan &p, const double *spec, double *four, double *four2,
                       bool value_table, bool deriv_table, int nf, hipStream_t s);
int launch_leg_fwd_f64(const Plan &p, const double *four, double *spec, int nf, hipStream_t s);
// Forward Legendre (K-split shape) with any forward-layout table; swap = take the
// opposite parity slot of the Fourier data (H tables), accumulate = add into spec.
int launch_leg_fwd_f64_table(const Plan &p, const double *table, const double *four, double *spec,
                             bool swap, bool accumulate, int nf, hipStream_t s);
// spec_out = (-n(n+1))^power spec_in (negative powers leave n = 0 at zero)
int launch_spec_lapla(const Plan &p, const double *in, double *out, int power, int nf, hipStream_t s);

// Legendre stage, mixed precision (sht_legendre_mixed.hip)
int build_mixed_tables(Plan &p);
int launch_leg_inv_mixed(const Plan &p, const double *spec, float *four, int nf, hipStream_t s);
// Streaming inverse of the mixed forms (SHT_SHAPE_STAGED): eight waves hold
// the table of up to 8 row tiles, so grids of at most 128 latitude pairs.
inline bool mixed_stream_available(const Plan &p) { return p.prec != SHT_PREC_FP64 && p.P2p <= 128; }
// Fused inverse of the mixed forms (SHT_SHAPE_FUSED): Legendre step and FFT
// in one persistent launch, grid values only; available for im in [64, 512]
// and nmax < 192.
constexpr int kFusedSyncHead = 18, kSyncStride = 64;
bool fused_inv_available(const Plan &p);
int launch_inv_fused(const Plan &p, const double *spec, float *four, double *grid, int nf, hipStream_t s);
// Waits of the plan's fused calls that gave up (synchronous read; sticky).
int fused_wait_failures(const Plan &p, long long *count);
int launch_leg_inv_mixed_shape(const Plan &p, int shape, const double *spec, float *four, int nf,
                               hipStream_t s);
int launch_leg_fwd_mixed_shape(const Plan &p, int shape, const float *four, double *spec, int nf,
                               hipStream_t s);
int launch_leg_fwd_mixed(const Plan &p, const float *four, double *spec, int nf, hipStream_t s);

// Allocation-free spectral -> grid (sht_noalloc.hip): the static scratch's
// device address is looked up once per device at plan creation.
int noalloc_init_device(int device);
int spec_to_grid_into(Plan &p, const double *spec, double *grid, int nf, void *four, hipStream_t s);

// Layout helpers (sht_pack.hip)
int launch_spml_to_canon(const Plan &p, const double *wa, double *spec, int nf, hipStream_t s);
int launch_canon_to_spml(const Plan &p, const double *spec, double *wa, int nf, hipStream_t s);
int launch_avr_lonlat(const Plan &p, const double *grid, double *avg, int nf, hipStream_t s);

// Index of the calling wavefront in its workgroup, in a scalar register: the
// compiler then knows that everything derived from it is uniform across the
// wave, so loops and branches on it stay scalar (no exec-mask divergence and
// no per-iteration memory waits).
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// XCD-aware block order: the hardware deals consecutive workgroup ids to the 8
// XCDs in turn; this maps them so that each XCD receives a contiguous range of
// logical tiles (neighbouring tiles share table rows in that XCD's L2).
// Balanced, L2-local dealing of whole wavenumbers: block b runs on XCD b % 8;
// that XCD's (b / 8)-th block is tile t of the j-th m of its list, with
// tiles_per_m blocks per m.  Returns false for the padding blocks.
__device__ __forceinline__ bool xcd_m_block(int b, const int *__restrict__ mlist, int mlen,
                                            int tiles_per_m, int &m, int &t) {
  const int x = b & 7, i = b >> 3, j = i / tiles_per_m;
  t = i - j * tiles_per_m;
  if (j >= mlen) return false;
  m = mlist[x * mlen + j];
  return m >= 0;
}
inline int xcd_m_blocks(const Plan &p, int dir, int tiles_per_m) {
  return 8 * p.xcd_mlen[dir] * tiles_per_m;
}

__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int nx = 8;
  int xcd = b % nx, idx = b / nx;
  int per = nblocks / nx, rem = nblocks % nx;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

}  // namespace sht

struct sht_plan {
  sht::Plan p;
};

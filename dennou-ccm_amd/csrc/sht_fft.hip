// sht_fft.hip -- the ring FFT stage of the transform (ISPACK's real FFT along
// each latitude ring, called by SPML wa_xya / xya_wa).
//
// One workgroup of 256 threads owns one latitude pair and G fields.  The north
// and south rings of a pair travel as the real and imaginary parts of ONE
// complex FFT of length L = im, so each ring costs half a complex transform and
// the equatorial split the Legendre step needs falls out of the pre/post
// processing for free:
//   forward:  z = f_N + i f_S  ->  Z = FFT(z);  X_N = (Z_m + conj Z_{L-m}) / 2L,
//             X_S = (Z_m - conj Z_{L-m}) / 2iL;   stored S = X_N + X_S, A = X_N - X_S
//   inverse:  from the Legendre sums E (even n-m) and O (odd): X_N = E + O,
//             X_S = E - O, Z_m = X_N + i X_S, Z_{L-m} = conj X_N + i conj X_S,
//             z = IFFT(Z): f_N = Re z, f_S = Im z.
// The Fourier coefficients are read/written in the [pair][m][parity][field]
// layout the Legendre GEMMs want, so this kernel is also the (lat,lon) <->
// (m,lat) transpose: global traffic is row-contiguous on the grid side and
// runs of 2G values on the Fourier side; the reordering happens in LDS.
// The complex FFT is a Stockham autosort with radix 8/16 butterflies held in
// registers (two or three passes per transform) over a single 32 KB LDS
// buffer holding G*L = 2048 complex doubles.  LDS bandwidth bounds the
// passes, so the ends skip what they can: the inverse places only the
// coefficients m < M1 and L - m, reads the band between them as zeros, and
// stores its last pass straight to the grid; the forward writes only those
// wavenumbers in its last pass and, when the launch fits the device at once,
// feeds its first pass straight from the grid.
#include "sht_fft_dev.h"

// No implicit fused multiply-adds in this file (see sht_fft_dev.h): the
// direct and LDS-staged forward kernels must round identically.
#pragma clang fp contract(off)

namespace sht {
namespace {

template <typename TF>
struct FwdIO {
  const double *src[2];
  TF *dst[2];
  int mode[2];  // 0 value, 1 divergence lon term (x i m / cos), 2 lat term (x 1 / cos)
};

// grid -> four (S, A halves).  Blocks: the FFT tiles (four_tiles) x inputs.
// DIRECT: the first FFT pass reads its butterflies' ring samples straight
// from global memory (runs of L/R samples), which saves an LDS round trip and
// wins while every block of the launch is resident at once; otherwise all
// threads first load whole rows into LDS (runs of 64 samples per wave), which
// streams better when blocks queue.  The last pass writes only the
// wavenumbers m < M1 and L - m the Legendre step reads.
// spread (both FFT kernels): the launch fits the device at once, so blocks
// take tiles in hardware order (block b runs on XCD b mod 8), which deals the
// workgroups of the last, partly empty field group evenly over the XCDs (T85,
// 97 fields: inverse 9.8 -> 9.5 us); otherwise each XCD takes a contiguous
// range of tiles, which measured ~1 % faster when blocks queue.
template <int L, typename TF, bool DIRECT, int C>
__global__ __launch_bounds__(kThreadsOf<C>) void fft_fwd_kernel(FwdIO<TF> io,
                                                                const double2 *__restrict__ tw,
                                                                const double *__restrict__ coslat,
                                                                int P2, int M1, int nf, int jm, int spread) {
  const double *__restrict__ grid = io.src[blockIdx.y];
  TF *__restrict__ four = io.dst[blockIdx.y];
  const int mode = io.mode[blockIdx.y];
  constexpr int G = C / L;
  constexpr int kThreads = kThreadsOf<C>;
  __shared__ double2 buf[kPaddedOf<C>];
  __shared__ double2 twl[L];
  const int tile = spread ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  const FourTile<G> ft(tile, P2);
  const int p = ft.p, f0 = ft.fgr * G;
  const int jn = P2 + p, js = P2 - 1 - p;
  const int tid = threadIdx.x;
  // the twiddles are loaded after the first ring loads are issued (one wait
  // covers both; loaded first, their wait held back every ring load)
  auto twiddles = [&] {
    for (int q = tid; q < L; q += kThreads) twl[q] = tw[q];
  };
  // z = f_N + i f_S; transforms past nf are zero
  auto ring = [&](int g, int i) {
    const int f = f0 + g;
    double2 z = make_double2(0.0, 0.0);
    if (f < nf) {
      z.x = __builtin_nontemporal_load(grid + ((size_t)f * jm + jn) * L + i);
      z.y = __builtin_nontemporal_load(grid + ((size_t)f * jm + js) * L + i);
    }
    return z;
  };
  auto keep = [&](int g, int i, double2 v) {
    if (i < M1 || i > L - M1) buf[px<L>(g * L + i)] = v;
  };
  if constexpr (DIRECT) {
    fft_lds<L, -1, C>(buf, twl, ring, keep, threadIdx.x, twiddles);
  } else {
    constexpr int NQ = C / kThreads;  // samples per thread
    double2 z[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int idx = tid + q * kThreads;
      z[q] = ring(idx / L, idx % L);
    }
    twiddles();
#pragma unroll
    for (int q = 0; q < NQ; ++q) buf[px<L>(tid + q * kThreads)] = z[q];
    __syncthreads();
    fft_lds<L, -1, C>(buf, twl, [&](int g, int i) { return buf[px<L>(g * L + i)]; }, keep, threadIdx.x);
  }
  const double sc = 0.5 / L * (mode == 0 ? 1.0 : 1.0 / coslat[jn]);
  for (int idx = tid; idx < M1 * G; idx += kThreads) {
    const int m = idx / G, g = idx % G, f = f0 + g;
    if (f >= nf) continue;
    const double2 zm = buf[px<L>(g * L + m)];
    const double2 zc = buf[px<L>(g * L + ((L - m) & (L - 1)))];
    // X_N = (Z_m + conj Z_{L-m}) sc ;  X_S = -i (Z_m - conj Z_{L-m}) sc
    const double2 xn = make_double2((zm.x + zc.x) * sc, (zm.y - zc.y) * sc);
    const double2 xs = make_double2((zm.y + zc.y) * sc, -(zm.x - zc.x) * sc);
    double2 S = cadd(xn, xs), A = csub(xn, xs);
    if (mode == 1) {  // x i m
      S = make_double2(-m * S.y, m * S.x);
      A = make_double2(-m * A.y, m * A.x);
    }
    if (m == 0) S.y = A.y = 0.0;
    TF *o = four + ft.off(m, M1, nf) + 2 * g;
    store_pair(o, S);
    store_pair(o + 2 * FourTile<G>::GL, A);
  }
}


template <typename TF>
struct InvOutputs {
  const TF *src[3];
  double *dst[3];
  int mode[3]; // 0 value, 1 GradLon, 2 GradLat
};

// four (even/odd Legendre sums) -> grid.  Blocks: the FFT tiles (four_tiles)
// x noutputs, one tile each (fft_inv_tile).
template <int L, typename TF, int C>
__global__ __launch_bounds__(kThreadsOf<C>) void fft_inv_kernel(InvOutputs<TF> io,
                                                                const double2 *__restrict__ tw,
                                                                const double *__restrict__ coslat,
                                                                int P2, int M1, int nf, int jm, int spread) {
  __shared__ double2 buf[kPaddedOf<C>];
  __shared__ double2 twl[L];
  const int tile = spread ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  const int sel = blockIdx.y;
  fft_inv_tile<L, TF, false, C>(buf, twl, tw, io.src[sel], io.dst[sel], io.mode[sel], coslat, P2, M1, nf, jm,
                                tile, threadIdx.x);
}

template <int L, int C, typename TF>
int fwd_launch(const Plan &p, const FwdIO<TF> &io, int nin, int nf, hipStream_t s) {
  const int nblk = four_tiles<C / L>(p.P2, nf);
  const bool resident = nblk * nin <= p.fft_fwd_resident[C == kFftSmall];
  auto kern = resident ? fft_fwd_kernel<L, TF, true, C> : fft_fwd_kernel<L, TF, false, C>;
  launch_kernel(kern, dim3(nblk, nin), kThreadsOf<C>, 0, s,
      io, reinterpret_cast<const double2 *>(p.twid), p.coslat, p.P2, p.N + 1, nf, p.jm, resident ? 1 : 0);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

// the small blocks exist for grids of at most kFftSmallMaxIm longitudes
template <int L, typename TF>
int fwd_dispatch(const Plan &p, const FwdIO<TF> &io, int nin, int nf, hipStream_t s) {
  if constexpr (L <= kFftSmallMaxIm)
    if (fft_block(p, nf) == kFftSmall) return fwd_launch<L, kFftSmall>(p, io, nin, nf, s);
  return fwd_launch<L, kFftBig>(p, io, nin, nf, s);
}

template <int L, int C>
int fwd_resident(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fft_fwd_kernel<L, double, true, C>, kThreadsOf<C>,
                                                   0) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * ncu;
}

template <int L, int C, typename TF>
int inv_launch(const Plan &p, const InvOutputs<TF> &io, int nout, int nf, hipStream_t s) {
  const int nblk = four_tiles<C / L>(p.P2, nf);
  // same LDS and block size as the forward kernel, hence the same residency
  const bool resident = nblk * nout <= p.fft_fwd_resident[C == kFftSmall];
  launch_kernel(fft_inv_kernel<L, TF, C>, dim3(nblk, nout), kThreadsOf<C>, 0, s,
      io, reinterpret_cast<const double2 *>(p.twid), p.coslat, p.P2, p.N + 1, nf, p.jm, resident ? 1 : 0);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int L, typename TF>
int inv_dispatch(const Plan &p, const InvOutputs<TF> &io, int nout, int nf, hipStream_t s) {
  if constexpr (L <= kFftSmallMaxIm)
    if (fft_block(p, nf) == kFftSmall) return inv_launch<L, kFftSmall>(p, io, nout, nf, s);
  return inv_launch<L, kFftBig>(p, io, nout, nf, s);
}

template <typename TF>
int fwd_typed(const Plan &p, const FwdIO<TF> &io, int nin, int nf, hipStream_t s) {
  if (nin == 0) return SHT_OK;
  switch (p.im) {
    case 16: return fwd_dispatch<16>(p, io, nin, nf, s);
    case 32: return fwd_dispatch<32>(p, io, nin, nf, s);
    case 64: return fwd_dispatch<64>(p, io, nin, nf, s);
    case 128: return fwd_dispatch<128>(p, io, nin, nf, s);
    case 256: return fwd_dispatch<256>(p, io, nin, nf, s);
    case 512: return fwd_dispatch<512>(p, io, nin, nf, s);
    case 1024: return fwd_dispatch<1024>(p, io, nin, nf, s);
    case 2048: return fwd_dispatch<2048>(p, io, nin, nf, s);
  }
  return fail(SHT_ERR_UNSUPPORTED, "im must be a power of two in [16, 2048]");
}

template <typename TF>
int inv_typed(const Plan &p, const TF *four, const TF *four2, double *grid, double *gradlon,
              double *gradlat, int nf, hipStream_t s) {
  InvOutputs<TF> io{};
  int n = 0;
  if (grid) { io.src[n] = four; io.dst[n] = grid; io.mode[n] = 0; ++n; }
  if (gradlon) { io.src[n] = four; io.dst[n] = gradlon; io.mode[n] = 1; ++n; }
  if (gradlat) { io.src[n] = four2; io.dst[n] = gradlat; io.mode[n] = 2; ++n; }
  if (n == 0) return SHT_OK;
  switch (p.im) {
    case 16: return inv_dispatch<16>(p, io, n, nf, s);
    case 32: return inv_dispatch<32>(p, io, n, nf, s);
    case 64: return inv_dispatch<64>(p, io, n, nf, s);
    case 128: return inv_dispatch<128>(p, io, n, nf, s);
    case 256: return inv_dispatch<256>(p, io, n, nf, s);
    case 512: return inv_dispatch<512>(p, io, n, nf, s);
    case 1024: return inv_dispatch<1024>(p, io, n, nf, s);
    case 2048: return inv_dispatch<2048>(p, io, n, nf, s);
  }
  return fail(SHT_ERR_UNSUPPORTED, "im must be a power of two in [16, 2048]");
}

}  // namespace

template <int L>
void capacities(int device, int (&out)[2]) {
  out[0] = fwd_resident<L, kFftBig>(device);
  if constexpr (L <= kFftSmallMaxIm) out[1] = fwd_resident<L, kFftSmall>(device);
}

void fft_fwd_capacity(const Plan &p, int (&out)[2]) {
  out[0] = out[1] = 0;
  switch (p.im) {
    case 16: return capacities<16>(p.device, out);
    case 32: return capacities<32>(p.device, out);
    case 64: return capacities<64>(p.device, out);
    case 128: return capacities<128>(p.device, out);
    case 256: return capacities<256>(p.device, out);
    case 512: return capacities<512>(p.device, out);
    case 1024: return capacities<1024>(p.device, out);
    case 2048: return capacities<2048>(p.device, out);
  }
}

int launch_fft_fwd(const Plan &p, const double *grid, void *four, bool f32, int nf, hipStream_t s) {
  if (f32) {
    FwdIO<float> io{{grid, nullptr}, {static_cast<float *>(four), nullptr}, {0, 0}};
    return fwd_typed(p, io, 1, nf, s);
  }
  FwdIO<double> io{{grid, nullptr}, {static_cast<double *>(four), nullptr}, {0, 0}};
  return fwd_typed(p, io, 1, nf, s);
}

int launch_fft_fwd_div(const Plan &p, const double *a, double *fa, const double *b, double *fb,
                       int nf, hipStream_t s) {
  FwdIO<double> io{};
  int n = 0;
  if (a) { io.src[n] = a; io.dst[n] = fa; io.mode[n] = 1; ++n; }
  if (b) { io.src[n] = b; io.dst[n] = fb; io.mode[n] = 2; ++n; }
  return fwd_typed(p, io, n, nf, s);
}

int launch_fft_inv(const Plan &p, const void *four, const void *four2, bool f32, double *grid,
                   double *gradlon, double *gradlat, int nf, hipStream_t s) {
  if (f32)
    return inv_typed(p, static_cast<const float *>(four), static_cast<const float *>(four2), grid,
                     gradlon, gradlat, nf, s);
  return inv_typed(p, static_cast<const double *>(four), static_cast<const double *>(four2), grid,
                   gradlon, gradlat, nf, s);
}

}  // namespace sht

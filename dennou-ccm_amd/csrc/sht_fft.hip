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
#include "sht_internal.h"

// No implicit fused multiply-adds in this file: the compiler picks which
// product of a*b - c*d to fuse per instantiation, so the direct and
// LDS-staged forward kernels (the same arithmetic) would round differently
// and a field's spectrum would depend on the batch size.  The complex
// multiply fuses explicitly instead (same operation count, fixed rounding).
#pragma clang fp contract(off)

namespace sht {
namespace {

constexpr int kThreads = 256;
constexpr int kLdsComplex = 2048;
// LDS holds one pad element per 16 complex values -- the radix-16 pass writes
// 16 consecutive outputs per thread (a 256-byte lane stride), which without the
// pad lands every lane of a ds_write on the same banks -- and one more per
// transform, because the Fourier-side loops run the G transforms of a block in
// consecutive lanes (element m of transform g), a stride of L(1 + 1/16)
// elements that is otherwise a multiple of the bank count.
constexpr int kLdsPadded = kLdsComplex + kLdsComplex / 16 + kLdsComplex / 16;
template <int L>
__device__ __forceinline__ int px(int i) { return i + (i >> 4) + i / L; }

// Regular stores: the Legendre step that reads the Fourier data next finds
// them in the Infinity Cache (streaming them made it slower at every size).
__device__ __forceinline__ void store_pair(double *o, double2 v) { *reinterpret_cast<double2 *>(o) = v; }
__device__ __forceinline__ void store_pair(float *o, double2 v) {
  *reinterpret_cast<float2 *>(o) = make_float2((float)v.x, (float)v.y);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// e^{+2 pi i k/16}, k = 0..15
__constant__ double kW16[16][2] = {
    {1.0, 0.0}, {0.92387953251128674, 0.38268343236508977}, {0.70710678118654752, 0.70710678118654752},
    {0.38268343236508977, 0.92387953251128674}, {0.0, 1.0}, {-0.38268343236508977, 0.92387953251128674},
    {-0.70710678118654752, 0.70710678118654752}, {-0.92387953251128674, 0.38268343236508977}, {-1.0, 0.0},
    {-0.92387953251128674, -0.38268343236508977}, {-0.70710678118654752, -0.70710678118654752},
    {-0.38268343236508977, -0.92387953251128674}, {0.0, -1.0}, {0.38268343236508977, -0.92387953251128674},
    {0.70710678118654752, -0.70710678118654752}, {0.92387953251128674, -0.38268343236508977}};

// In-register DFT of R = 2, 4, 8 or 16 points (exponent sign SIGN), radix-2
// decimation in time over bit-reversed inputs; every index and constant
// twiddle is resolved at compile time.
template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(double2 (&v)[R]) {
  constexpr int LOG = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  double2 t[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int rev = 0;
#pragma unroll
    for (int b = 0; b < LOG; ++b) rev |= ((i >> b) & 1) << (LOG - 1 - b);
    t[i] = v[rev];
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += len)
#pragma unroll
      for (int k = 0; k < len / 2; ++k) {
        const int wi = k * (16 / len);
        const double2 w = make_double2(kW16[wi][0], SIGN * kW16[wi][1]);
        const double2 b = (k == 0) ? t[i + k + len / 2] : cmul(t[i + k + len / 2], w);
        const double2 a = t[i + k];
        t[i + k] = cadd(a, b);
        t[i + k + len / 2] = csub(a, b);
      }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = t[i];
}

// One Stockham pass of radix R over G transforms of length L, Ns the product
// of the radices already applied.  Butterfly j of transform g reads
// x[j + r L/R] (in(g, i) returns element i of transform g), applies
// w^(r k L/(Ns R)) (k = j mod Ns; the powers are built by repeated
// multiplication from one LDS twiddle), runs a register DFT and writes
// y[(j/Ns) Ns R + k + r Ns] through out(g, i, v).  A barrier separates the
// reads from the writes (the passes run in place in buf) and follows them.
template <int L, int R, int Ns, int SIGN, class In, class Out>
__device__ __forceinline__ void stockham_pass(const double2 *tw, In in, Out out) {
  constexpr int G = kLdsComplex / L;
  constexpr int NB = G * (L / R);                       // butterflies in the block
  constexpr int QB = (NB + kThreads - 1) / kThreads;    // per thread
  const int tid = threadIdx.x;
  double2 v[QB][R];
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int idx = tid + q * kThreads;
    if (NB % kThreads != 0 && idx >= NB) continue;
    const int g = idx / (L / R), j = idx % (L / R), k = j % Ns;
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = in(g, j + r * (L / R));
    if constexpr (Ns > 1) {
      double2 w1 = tw[k * (L / (Ns * R))];
      w1.y *= SIGN;
      double2 w = w1;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        v[q][r] = cmul(v[q][r], w);
        if (r + 1 < R) w = cmul(w, w1);
      }
    }
    dft_reg<R, SIGN>(v[q]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int idx = tid + q * kThreads;
    if (NB % kThreads != 0 && idx >= NB) continue;
    const int g = idx / (L / R), j = idx % (L / R), k = j % Ns;
    const int d0 = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out(g, d0 + r * Ns, v[q][r]);
  }
  __syncthreads();
}

// Complex FFT of G = kLdsComplex / L transforms, exponent sign SIGN (-1
// forward, +1 inverse), unnormalised; tw (in LDS) holds e^{+2 pi i t / L}.
// The first pass reads through first(g, i) and the last writes through
// last(g, i, v), so a caller can feed the first butterflies straight from
// global memory, skip known zeros, or store the results straight out, and
// save the LDS round trips; the passes between run in place in buf.
// Radix plan: 16 = 16, 32 = 8*4, 64 = 8*8, 128 = 16*8, 256 = 16*16,
// 512 = 8*8*8, 1024 = 16*16*4, 2048 = 16*16*8.  Ends with a barrier.
template <int L, int SIGN, class First, class Last>
__device__ __forceinline__ void fft_lds(double2 *buf, const double2 *tw, First first, Last last) {
  auto rd = [&](int g, int i) { return buf[px<L>(g * L + i)]; };
  auto wr = [&](int g, int i, double2 v) { buf[px<L>(g * L + i)] = v; };
  if constexpr (L == 16) {
    stockham_pass<L, 16, 1, SIGN>(tw, first, last);
  } else if constexpr (L == 32) {
    stockham_pass<L, 8, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 4, 8, SIGN>(tw, rd, last);
  } else if constexpr (L == 64) {
    stockham_pass<L, 8, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 8, 8, SIGN>(tw, rd, last);
  } else if constexpr (L == 128) {
    stockham_pass<L, 16, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 8, 16, SIGN>(tw, rd, last);
  } else if constexpr (L == 256) {
    stockham_pass<L, 16, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 16, 16, SIGN>(tw, rd, last);
  } else if constexpr (L == 512) {
    stockham_pass<L, 8, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 8, 8, SIGN>(tw, rd, wr);
    stockham_pass<L, 8, 64, SIGN>(tw, rd, last);
  } else if constexpr (L == 1024) {
    stockham_pass<L, 16, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 16, 16, SIGN>(tw, rd, wr);
    stockham_pass<L, 4, 256, SIGN>(tw, rd, last);
  } else {
    static_assert(L == 2048, "unsupported ring length");
    stockham_pass<L, 16, 1, SIGN>(tw, first, wr);
    stockham_pass<L, 16, 16, SIGN>(tw, rd, wr);
    stockham_pass<L, 8, 256, SIGN>(tw, rd, last);
  }
}

template <typename TF>
struct FwdIO {
  const double *src[2];
  TF *dst[2];
  int mode[2];  // 0 value, 1 divergence lon term (x i m / cos), 2 lat term (x 1 / cos)
};

// grid -> four (S, A halves).  Blocks: P2 * ceil(nf / G) x inputs.
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
template <int L, typename TF, bool DIRECT>
__global__ __launch_bounds__(kThreads) void fft_fwd_kernel(FwdIO<TF> io,
                                                           const double2 *__restrict__ tw,
                                                           const double *__restrict__ coslat,
                                                           int P2, int M1, int nf, int jm, int spread) {
  const double *__restrict__ grid = io.src[blockIdx.y];
  TF *__restrict__ four = io.dst[blockIdx.y];
  const int mode = io.mode[blockIdx.y];
  constexpr int G = kLdsComplex / L;
  __shared__ double2 buf[kLdsPadded];
  __shared__ double2 twl[L];
  const int tile = spread ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile % P2, fgr = tile / P2, f0 = fgr * G;
  const int nfg = (nf + G - 1) / G;
  const int jn = P2 + p, js = P2 - 1 - p;
  const int tid = threadIdx.x;
  for (int q = tid; q < L; q += kThreads) twl[q] = tw[q];
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
    fft_lds<L, -1>(buf, twl, ring, keep);
  } else {
    constexpr int NQ = kLdsComplex / kThreads;  // samples per thread
    double2 z[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int idx = tid + q * kThreads;
      z[q] = ring(idx / L, idx % L);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) buf[px<L>(tid + q * kThreads)] = z[q];
    __syncthreads();
    fft_lds<L, -1>(buf, twl, [&](int g, int i) { return buf[px<L>(g * L + i)]; }, keep);
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
    TF *o = four + ((((size_t)p * nfg + fgr) * M1 + m) * 2) * (2 * G) + 2 * g;
    store_pair(o, S);
    store_pair(o + 2 * G, A);
  }
}


template <typename TF>
struct InvOutputs {
  const TF *src[3];
  double *dst[3];
  int mode[3]; // 0 value, 1 GradLon, 2 GradLat
};

// four (even/odd Legendre sums) -> grid.  Blocks: P2 * ceil(nf / G) x noutputs.
// A thread's Fourier loads (at most 4 (m, field) pairs: M1 * G <= L/2 * G =
// 1024) and its twiddles are issued before any LDS traffic.  Only the
// coefficients m < M1 and L - m are placed in LDS: the first FFT pass takes
// the band between them as zeros without reading it, and the last pass
// stores the rings straight to the grid (non-temporal).
template <int L, typename TF>
__global__ __launch_bounds__(kThreads) void fft_inv_kernel(InvOutputs<TF> io,
                                                           const double2 *__restrict__ tw,
                                                           const double *__restrict__ coslat,
                                                           int P2, int M1, int nf, int jm, int spread) {
  constexpr int G = kLdsComplex / L;
  constexpr int NM = (kLdsComplex / 2) / kThreads; // max (m, field) pairs per thread
  __shared__ double2 buf[kLdsPadded];
  __shared__ double2 twl[L];
  const int tile = spread ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile % P2, fgr = tile / P2, f0 = fgr * G;
  const int nfg = (nf + G - 1) / G;
  const int jn = P2 + p, js = P2 - 1 - p;
  const int tid = threadIdx.x;
  const int sel = blockIdx.y;
  const TF *__restrict__ src = io.src[sel];
  double *__restrict__ dst = io.dst[sel];
  const int mode = io.mode[sel];
  const double rc = (mode == 0) ? 1.0 : 1.0 / coslat[jn];
  double2 E[NM], O[NM];
#pragma unroll
  for (int q = 0; q < NM; ++q) {
    const int idx = tid + q * kThreads;
    const int m = idx / G, g = idx % G, f = f0 + g;
    E[q] = O[q] = make_double2(0.0, 0.0);
    if (m < M1 && f < nf) {
      const TF *e = src + ((((size_t)p * nfg + fgr) * M1 + m) * 2) * (2 * G) + 2 * g;
      E[q] = make_double2((double)e[0], (double)e[1]);
      O[q] = make_double2((double)e[2 * G], (double)e[2 * G + 1]);
    }
  }
  for (int q = tid; q < L; q += kThreads) twl[q] = tw[q];
#pragma unroll
  for (int q = 0; q < NM; ++q) {
    const int idx = tid + q * kThreads;
    const int m = idx / G, g = idx % G, f = f0 + g;
    if (m >= M1 || f >= nf) continue;
    double2 xn = cadd(E[q], O[q]), xs = csub(E[q], O[q]);
    if (mode == 1) { // i m / cos(lat)
      const double s = m * rc;
      xn = make_double2(-xn.y * s, xn.x * s);
      xs = make_double2(-xs.y * s, xs.x * s);
    } else if (mode == 2) {
      xn = make_double2(xn.x * rc, xn.y * rc);
      xs = make_double2(xs.x * rc, xs.y * rc);
    }
    if (m == 0) {
      buf[px<L>(g * L)] = make_double2(xn.x, xs.x);
    } else {
      buf[px<L>(g * L + m)] = make_double2(xn.x - xs.y, xn.y + xs.x);
      buf[px<L>(g * L + L - m)] = make_double2(xn.x + xs.y, -xn.y + xs.x);
    }
  }
  __syncthreads();
  // transforms past nf hold stale LDS values: their rings are not stored
  auto spectrum = [&](int g, int i) {
    return (i < M1 || i > L - M1) ? buf[px<L>(g * L + i)] : make_double2(0.0, 0.0);
  };
  auto rings = [&](int g, int i, double2 z) {
    const int f = f0 + g;
    if (f < nf) {
      __builtin_nontemporal_store(z.x, dst + ((size_t)f * jm + jn) * L + i);
      __builtin_nontemporal_store(z.y, dst + ((size_t)f * jm + js) * L + i);
    }
  };
  fft_lds<L, +1>(buf, twl, spectrum, rings);
}

template <int L, typename TF>
int fwd_dispatch(const Plan &p, const FwdIO<TF> &io, int nin, int nf, hipStream_t s) {
  constexpr int G = kLdsComplex / L;
  const int nblk = p.P2 * ((nf + G - 1) / G);
  const bool resident = nblk * nin <= p.fft_fwd_resident;
  auto kern = resident ? fft_fwd_kernel<L, TF, true> : fft_fwd_kernel<L, TF, false>;
  launch_kernel(kern, dim3(nblk, nin), kThreads, 0, s,
      io, reinterpret_cast<const double2 *>(p.twid), p.coslat, p.P2, p.N + 1, nf, p.jm, resident ? 1 : 0);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int L>
int fwd_resident(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fft_fwd_kernel<L, double, true>, kThreads, 0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * ncu;
}

template <int L, typename TF>
int inv_dispatch(const Plan &p, const InvOutputs<TF> &io, int nout, int nf, hipStream_t s) {
  constexpr int G = kLdsComplex / L;
  const int nblk = p.P2 * ((nf + G - 1) / G);
  // same LDS and block size as the forward kernel, hence the same residency
  const bool resident = nblk * nout <= p.fft_fwd_resident;
  launch_kernel(fft_inv_kernel<L, TF>, dim3(nblk, nout), kThreads, 0, s,
      io, reinterpret_cast<const double2 *>(p.twid), p.coslat, p.P2, p.N + 1, nf, p.jm, resident ? 1 : 0);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
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

int fft_fwd_capacity(const Plan &p) {
  switch (p.im) {
    case 16: return fwd_resident<16>(p.device);
    case 32: return fwd_resident<32>(p.device);
    case 64: return fwd_resident<64>(p.device);
    case 128: return fwd_resident<128>(p.device);
    case 256: return fwd_resident<256>(p.device);
    case 512: return fwd_resident<512>(p.device);
    case 1024: return fwd_resident<1024>(p.device);
    case 2048: return fwd_resident<2048>(p.device);
  }
  return 0;
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

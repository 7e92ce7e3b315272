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
// 2G-double runs on the Fourier side; the reordering happens in LDS.
// The complex FFT is a Stockham autosort (radix 4, one radix-2 pass when log2 L
// is odd) in a single 32 KB LDS buffer holding G*L = 2048 complex doubles.
#include "sht_internal.h"

namespace sht {
namespace {

constexpr int kThreads = 256;
constexpr int kLdsComplex = 2048;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by SIGN * i
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 a) {
  return SIGN > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// In-place complex FFT of G = kLdsComplex / L transforms stored back to back in
// buf; exponent sign SIGN (-1 forward, +1 inverse), unnormalised.  tw holds
// e^{+2 pi i t / L}.  Ends with a barrier.
template <int L, int SIGN>
__device__ __forceinline__ void fft_lds(double2 *buf, const double2 *__restrict__ tw) {
  constexpr int G = kLdsComplex / L;
  constexpr int Q4 = G * (L / 4) / kThreads; // radix-4 butterflies per thread
  const int tid = threadIdx.x;
  int Ns = 1;
  for (; Ns * 4 <= L; Ns *= 4) {
    double2 v[Q4][4];
#pragma unroll
    for (int q = 0; q < Q4; ++q) {
      const int idx = tid + q * kThreads;
      const int g = idx / (L / 4), j = idx % (L / 4), k = j % Ns;
      const double2 *src = buf + g * L;
#pragma unroll
      for (int r = 0; r < 4; ++r) v[q][r] = src[j + r * (L / 4)];
      const int ts = k * (L / (Ns * 4));
#pragma unroll
      for (int r = 1; r < 4; ++r) {
        double2 w = tw[(r * ts) & (L - 1)];
        w.y *= SIGN;
        v[q][r] = cmul(v[q][r], w);
      }
      const double2 t0 = cadd(v[q][0], v[q][2]), t1 = csub(v[q][0], v[q][2]);
      const double2 t2 = cadd(v[q][1], v[q][3]), t3 = mul_si<SIGN>(csub(v[q][1], v[q][3]));
      v[q][0] = cadd(t0, t2);
      v[q][2] = csub(t0, t2);
      v[q][1] = cadd(t1, t3);
      v[q][3] = csub(t1, t3);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q4; ++q) {
      const int idx = tid + q * kThreads;
      const int g = idx / (L / 4), j = idx % (L / 4), k = j % Ns;
      double2 *dst = buf + g * L + (j / Ns) * Ns * 4 + k;
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[r * Ns] = v[q][r];
    }
    __syncthreads();
  }
  if (Ns < L) { // remaining factor 2: Ns == L/2
    constexpr int Q2 = G * (L / 2) / kThreads;
    double2 v[Q2][2];
#pragma unroll
    for (int q = 0; q < Q2; ++q) {
      const int idx = tid + q * kThreads;
      const int g = idx / (L / 2), j = idx % (L / 2);
      const double2 *src = buf + g * L;
      double2 a = src[j], b = src[j + L / 2];
      double2 w = tw[j];
      w.y *= SIGN;
      b = cmul(b, w);
      v[q][0] = cadd(a, b);
      v[q][1] = csub(a, b);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q2; ++q) {
      const int idx = tid + q * kThreads;
      const int g = idx / (L / 2), j = idx % (L / 2);
      buf[g * L + j] = v[q][0];
      buf[g * L + j + L / 2] = v[q][1];
    }
    __syncthreads();
  }
}

// grid -> four (S, A halves).  Blocks: P2 * ceil(nf / G).
template <int L>
__global__ __launch_bounds__(kThreads) void fft_fwd_kernel(const double *__restrict__ grid,
                                                           double *__restrict__ four,
                                                           const double2 *__restrict__ tw, int P2,
                                                           int M1, int nf, int jm) {
  constexpr int G = kLdsComplex / L;
  __shared__ double2 buf[kLdsComplex];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile % P2, f0 = (tile / P2) * G;
  const int jn = P2 + p, js = P2 - 1 - p;
  const int tid = threadIdx.x;
  const size_t ncol = 2 * (size_t)nf;
  for (int idx = tid; idx < G * L; idx += kThreads) {
    const int g = idx / L, t = idx % L, f = f0 + g;
    double2 z = make_double2(0.0, 0.0);
    if (f < nf) {
      z.x = grid[((size_t)f * jm + jn) * L + t];
      z.y = grid[((size_t)f * jm + js) * L + t];
    }
    buf[idx] = z;
  }
  __syncthreads();
  fft_lds<L, -1>(buf, tw);
  const double sc = 0.5 / L;
  for (int idx = tid; idx < M1 * G; idx += kThreads) {
    const int m = idx / G, g = idx % G, f = f0 + g;
    if (f >= nf) continue;
    const double2 zm = buf[g * L + m];
    const double2 zc = buf[g * L + ((L - m) & (L - 1))];
    // X_N = (Z_m + conj Z_{L-m}) sc ;  X_S = -i (Z_m - conj Z_{L-m}) sc
    const double2 xn = make_double2((zm.x + zc.x) * sc, (zm.y - zc.y) * sc);
    const double2 xs = make_double2((zm.y + zc.y) * sc, -(zm.x - zc.x) * sc);
    double2 S = cadd(xn, xs), A = csub(xn, xs);
    if (m == 0) S.y = A.y = 0.0;
    double *o = four + ((size_t)(p * M1 + m) * 2) * ncol + 2 * (size_t)f;
    *reinterpret_cast<double2 *>(o) = S;
    *reinterpret_cast<double2 *>(o + ncol) = A;
  }
}

struct InvOutputs {
  const double *src[3];
  double *dst[3];
  int mode[3]; // 0 value, 1 GradLon, 2 GradLat
};

// four (even/odd Legendre sums) -> grid.  Blocks: P2 * ceil(nf / G) x noutputs.
template <int L>
__global__ __launch_bounds__(kThreads) void fft_inv_kernel(InvOutputs io,
                                                           const double2 *__restrict__ tw,
                                                           const double *__restrict__ coslat,
                                                           int P2, int M1, int nf, int jm) {
  constexpr int G = kLdsComplex / L;
  __shared__ double2 buf[kLdsComplex];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile % P2, f0 = (tile / P2) * G;
  const int jn = P2 + p, js = P2 - 1 - p;
  const int tid = threadIdx.x;
  const int sel = blockIdx.y;
  const double *__restrict__ src = io.src[sel];
  double *__restrict__ dst = io.dst[sel];
  const int mode = io.mode[sel];
  const double rc = (mode == 0) ? 1.0 : 1.0 / coslat[jn];
  const size_t ncol = 2 * (size_t)nf;
  for (int idx = tid; idx < kLdsComplex; idx += kThreads) buf[idx] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int idx = tid; idx < M1 * G; idx += kThreads) {
    const int m = idx / G, g = idx % G, f = f0 + g;
    if (f >= nf) continue;
    const double *e = src + ((size_t)(p * M1 + m) * 2) * ncol + 2 * (size_t)f;
    const double2 E = *reinterpret_cast<const double2 *>(e);
    const double2 O = *reinterpret_cast<const double2 *>(e + ncol);
    double2 xn = cadd(E, O), xs = csub(E, O);
    if (mode == 1) { // i m / cos(lat)
      const double s = m * rc;
      xn = make_double2(-xn.y * s, xn.x * s);
      xs = make_double2(-xs.y * s, xs.x * s);
    } else if (mode == 2) {
      xn = make_double2(xn.x * rc, xn.y * rc);
      xs = make_double2(xs.x * rc, xs.y * rc);
    }
    double2 *z = buf + g * L;
    if (m == 0) {
      z[0] = make_double2(xn.x, xs.x);
    } else {
      z[m] = make_double2(xn.x - xs.y, xn.y + xs.x);
      z[L - m] = make_double2(xn.x + xs.y, -xn.y + xs.x);
    }
  }
  __syncthreads();
  fft_lds<L, +1>(buf, tw);
  for (int idx = tid; idx < G * L; idx += kThreads) {
    const int g = idx / L, t = idx % L, f = f0 + g;
    if (f >= nf) continue;
    const double2 z = buf[idx];
    dst[((size_t)f * jm + jn) * L + t] = z.x;
    dst[((size_t)f * jm + js) * L + t] = z.y;
  }
}

template <int L>
int fwd_dispatch(const Plan &p, const double *grid, double *four, int nf, hipStream_t s) {
  constexpr int G = kLdsComplex / L;
  const int nblk = p.P2 * ((nf + G - 1) / G);
  fft_fwd_kernel<L><<<nblk, kThreads, 0, s>>>(grid, four, reinterpret_cast<const double2 *>(p.twid),
                                              p.P2, p.N + 1, nf, p.jm);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int L>
int inv_dispatch(const Plan &p, const InvOutputs &io, int nout, int nf, hipStream_t s) {
  constexpr int G = kLdsComplex / L;
  const int nblk = p.P2 * ((nf + G - 1) / G);
  fft_inv_kernel<L><<<dim3(nblk, nout), kThreads, 0, s>>>(
      io, reinterpret_cast<const double2 *>(p.twid), p.coslat, p.P2, p.N + 1, nf, p.jm);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace

int launch_fft_fwd(const Plan &p, const double *grid, double *four, int nf, hipStream_t s) {
  switch (p.im) {
    case 16: return fwd_dispatch<16>(p, grid, four, nf, s);
    case 32: return fwd_dispatch<32>(p, grid, four, nf, s);
    case 64: return fwd_dispatch<64>(p, grid, four, nf, s);
    case 128: return fwd_dispatch<128>(p, grid, four, nf, s);
    case 256: return fwd_dispatch<256>(p, grid, four, nf, s);
    case 512: return fwd_dispatch<512>(p, grid, four, nf, s);
    case 1024: return fwd_dispatch<1024>(p, grid, four, nf, s);
    case 2048: return fwd_dispatch<2048>(p, grid, four, nf, s);
  }
  return fail(SHT_ERR_UNSUPPORTED, "im must be a power of two in [16, 2048]");
}

int launch_fft_inv(const Plan &p, const double *four, const double *four2, double *grid,
                   double *gradlon, double *gradlat, int nf, hipStream_t s) {
  InvOutputs io{};
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

}  // namespace sht

// sht_legendre.hip -- the Gauss-Legendre step in FP64 on the matrix cores.
//
// Per zonal wavenumber m the transform is a pair of small GEMMs, one per
// equatorial parity of n - m (ISPACK's symmetric/antisymmetric Legendre
// transform behind SPML wa_xya / xya_wa):
//   inverse  E[p][c] = sum_k Pt_{m+2k}^m(mu_p)   s[m+2k][c]
//            O[p][c] = sum_k Pt_{m+2k+1}^m(mu_p) s[m+2k+1][c]
//   forward  s[m+2k+par][c] = sum_p (w_p/2) Pt^m(mu_p) SA_par[p][c]
// where c runs over 2 * nfield real columns (complex spectra, fields
// innermost), so levels x fields are the GEMMs' batch dimension.
//
// Each workgroup computes one (m, 16-column tile) with
// v_mfma_f64_16x16x4f64: A = table fragment (16 rows x 4; lane l holds row
// l%16, k l/16), B = data fragment (4 x 16 columns; lane l holds k l/16,
// column l%16), D = 16 x 16 accumulators, 4 doubles per lane: item v of lane l
// is row 4v + l/16, column l%16 (measured on gfx950 with a one-hot lane
// probe).  All global fragments are 128-byte runs.
// Work per m shrinks with m (N - m + 1 degrees), so tiles are ordered heaviest
// first and dealt to XCDs in contiguous ranges (xcd_tile).
#include "sht_internal.h"

namespace sht {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kWaves = 4;   // wavefronts per workgroup, splitting the GEMM's K dimension
constexpr int kSteps = 4;   // K-steps (of 4) per wave whose fragments are loaded at once

// Inverse: spec -> four (tab 0, value table) and/or four2 (tab 1, derivative
// table, written with the parity slots swapped because (1-mu^2) dPt/dmu has the
// opposite equatorial parity to Pt).
// Workgroup = one (m, 16-column tile); wave w takes the K-steps w, w+4, ...,
// issues all their table and spectral fragments at once, runs its MFMAs over
// RT row tiles (16 latitude pairs each), and the four partial sums are added
// in FP64 through LDS.
template <int RT>
__global__ __launch_bounds__(256) void leg_inv_f64_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv,
    const int *__restrict__ poff, int N, int P2, int P2p, int ncol, int nct, int tab0, int ntab,
    FourGeom geo) {
  extern __shared__ double red[]; // [kWaves][RT][4][64]
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = b % nct;
  const int rest = b / nct;
  const int tab = tab0 + rest % ntab;
  const int m = rest / ntab;
  const int M1 = N + 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP = (ke + 3) & ~3, nsteps = KP >> 2;
  const double *__restrict__ T = (tab ? Hinv : Pinv) + poff[m];
  double *__restrict__ out = tab ? four2 : four;
  const int se = tab ? 1 : 0, so = 1 - se;
  const double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;

  for (int rt0 = 0; rt0 < P2p / 16; rt0 += RT) {
    // both parities' fragments are issued in the same round of loads
    d4 acc[2][RT];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < RT; ++r) acc[q][r] = d4{0.0, 0.0, 0.0, 0.0};
    const double *__restrict__ Tp = T + rt0 * 16 + r16;
    const double *__restrict__ sp = sm + col;
    for (int j0 = 0; wave + kWaves * j0 < nsteps; j0 += kSteps) {
      double a[2][kSteps][RT], bv[2][kSteps];
#pragma unroll
      for (int j = 0; j < kSteps; ++j) {
        const int st = wave + kWaves * (j0 + j), k = st * 4 + kq;
        const bool ok = st < nsteps;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          bv[q][j] = (ok && cok && k < (q ? ko : ke)) ? sp[(size_t)(2 * k + q) * ncol] : 0.0;
#pragma unroll
          for (int r = 0; r < RT; ++r)
            a[q][j][r] = ok ? Tp[(size_t)(q * KP + k) * P2p + r * 16] : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < kSteps; ++j)
        if (wave + kWaves * (j0 + j) < nsteps)
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < RT; ++r) acc[q][r] = mfma_f64(a[q][j][r], bv[q][j], acc[q][r]);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) red[((wave * RT + r) * 4 + i) * 64 + lane] = acc[q][r][i];
      __syncthreads();
      const int slot = q ? so : se;
      for (int e = threadIdx.x; e < RT * 256; e += 256) {
        const double v = red[e] + red[e + RT * 256] + red[e + 2 * RT * 256] + red[e + 3 * RT * 256];
        const int ln = e & 63, ri = e >> 6, i = ri & 3, r = ri >> 2;
        const int p = (rt0 + r) * 16 + i * 4 + (ln >> 4), c = ct * 16 + (ln & 15);
        if (p < P2 && c < ncol) out[geo.at(p, m, slot, c)] = v;
      }
      __syncthreads();
    }
  }
}

// Forward: four (S, A halves) -> spec.  Workgroup = one (m, 16-column tile);
// wave w takes the latitude-pair steps w, w+4, ... (4 pairs each) of the
// quadrature, the partial sums over its pairs meet in LDS and are added in
// FP64.  MAXRT = row tiles of 16 degrees per parity at m = 0.
template <int MAXRT>
__global__ __launch_bounds__(256) void leg_fwd_f64_kernel(const double *__restrict__ four,
                                                          double *__restrict__ spec,
                                                          const double *__restrict__ Pfwd,
                                                          const int *__restrict__ poff, int N,
                                                          int P2, int P2p, int ncol, int nct,
                                                          FourGeom geo) {
  extern __shared__ double red[]; // [kWaves][MAXRT][4][64]
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = b % nct, m = b / nct;
  const int M1 = N + 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  const int nsteps = P2p >> 2;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
  const int kcnt[2] = {ke, nn >> 1};
  d4 acc[2][MAXRT];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int r = 0; r < MAXRT; ++r) acc[q][r] = d4{0.0, 0.0, 0.0, 0.0};
  const double *__restrict__ W = Pfwd + poff[m] + r16;
  const int csafe = cok ? col : 0;
  // Loads go to clamped, always-valid addresses and are masked afterwards, so
  // the compiler issues all of them before a single wait (guarded loads become
  // branches with a wait each).
  for (int j0 = 0; wave + kWaves * j0 < nsteps; j0 += kSteps) {
    double a[2][kSteps][MAXRT], bv[2][kSteps];
#pragma unroll
    for (int j = 0; j < kSteps; ++j) {
      const int st = wave + kWaves * (j0 + j), p = st * 4 + kq;
      const bool ok = st < nsteps;
      const bool vb = ok && cok && p < P2;
      const int pb = vb ? p : 0, pa = ok ? p : 0;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const double x = four[geo.at(pb, m, q, csafe)];
        bv[q][j] = vb ? x : 0.0;
#pragma unroll
        for (int r = 0; r < MAXRT; ++r)
          a[q][j][r] = W[((size_t)q * P2p + pa) * KR + (r < nrt ? r : 0) * 16];
      }
    }
#pragma unroll
    for (int j = 0; j < kSteps; ++j)
      if (wave + kWaves * (j0 + j) < nsteps)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int r = 0; r < MAXRT; ++r)
            if (r < nrt) acc[q][r] = mfma_f64(a[q][j][r], bv[q][j], acc[q][r]);
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
#pragma unroll
    for (int r = 0; r < MAXRT; ++r)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[((wave * MAXRT + r) * 4 + i) * 64 + lane] = acc[q][r][i];
    __syncthreads();
    for (int e = threadIdx.x; e < nrt * 256; e += 256) {
      const double v = red[e] + red[e + MAXRT * 256] + red[e + 2 * MAXRT * 256] + red[e + 3 * MAXRT * 256];
      const int ln = e & 63, ri = e >> 6, i = ri & 3, r = ri >> 2;
      const int k = r * 16 + i * 4 + (ln >> 4), c = ct * 16 + (ln & 15);
      if (k < kcnt[q] && c < ncol) sm[(size_t)(2 * k + q) * ncol + c] = v;
    }
    __syncthreads();
  }
}

template <int RT>
int inv_launch(const Plan &p, const double *spec, double *four, double *four2, int tab0, int ntab,
               int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nblk = (p.N + 1) * ntab * nct;
  const size_t lds = sizeof(double) * kWaves * RT * 256;
  leg_inv_f64_kernel<RT><<<nblk, 256, lds, s>>>(spec, four, four2, p.Pinv, p.Hinv, p.poff_inv, p.N,
                                                p.P2, p.P2p, ncol, nct, tab0, ntab,
                                                four_geom(p.N, p.im, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int MAXRT>
int fwd_launch(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nblk = (p.N + 1) * nct;
  const size_t lds = sizeof(double) * kWaves * MAXRT * 256;
  leg_fwd_f64_kernel<MAXRT><<<nblk, 256, lds, s>>>(four, spec, p.Pfwd, p.poff_fwd, p.N, p.P2,
                                                   p.P2p, ncol, nct, four_geom(p.N, p.im, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace

int launch_leg_inv_f64(const Plan &p, const double *spec, double *four, double *four2,
                       bool value_table, bool deriv_table, int nf, hipStream_t s) {
  if (!value_table && !deriv_table) return SHT_OK;
  const int tab0 = value_table ? 0 : 1;
  const int ntab = (value_table && deriv_table) ? 2 : 1;
  const int rtiles = p.P2p / 16;
  int rt = 8;
  while (rtiles % rt) --rt;
  switch (rt) {
    case 8: return inv_launch<8>(p, spec, four, four2, tab0, ntab, nf, s);
    case 7: return inv_launch<7>(p, spec, four, four2, tab0, ntab, nf, s);
    case 6: return inv_launch<6>(p, spec, four, four2, tab0, ntab, nf, s);
    case 5: return inv_launch<5>(p, spec, four, four2, tab0, ntab, nf, s);
    case 4: return inv_launch<4>(p, spec, four, four2, tab0, ntab, nf, s);
    case 3: return inv_launch<3>(p, spec, four, four2, tab0, ntab, nf, s);
    case 2: return inv_launch<2>(p, spec, four, four2, tab0, ntab, nf, s);
    default: return inv_launch<1>(p, spec, four, four2, tab0, ntab, nf, s);
  }
}

int launch_leg_fwd_f64(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  const int rt = (ceil_half(p.N + 1) + 15) / 16;
  switch (rt) {
    case 1: return fwd_launch<1>(p, four, spec, nf, s);
    case 2: return fwd_launch<2>(p, four, spec, nf, s);
    case 3: return fwd_launch<3>(p, four, spec, nf, s);
    case 4: return fwd_launch<4>(p, four, spec, nf, s);
    case 5: return fwd_launch<5>(p, four, spec, nf, s);
    case 6: return fwd_launch<6>(p, four, spec, nf, s);
    case 7: return fwd_launch<7>(p, four, spec, nf, s);
    case 8: return fwd_launch<8>(p, four, spec, nf, s);
  }
  return fail(SHT_ERR_UNSUPPORTED, "nmax above 255 is not instantiated");
}

}  // namespace sht

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
// Each workgroup is one wavefront computing one (m, 16-column tile) with
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

// Inverse: spec -> four (tab 0, value table) and/or four2 (tab 1, derivative
// table, written with the parity slots swapped because (1-mu^2) dPt/dmu has the
// opposite equatorial parity to Pt).  RT row tiles (16 pairs each) per pass.
template <int RT>
__global__ __launch_bounds__(64) void leg_inv_f64_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv,
    const int *__restrict__ poff, int N, int P2, int P2p, int ncol, int nct, int tab0, int ntab) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = b % nct;
  const int rest = b / nct;
  const int tab = tab0 + rest % ntab;
  const int m = rest / ntab;
  const int M1 = N + 1;
  const int lane = threadIdx.x, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP = (ke + 3) & ~3;
  const double *__restrict__ T = (tab ? Hinv : Pinv) + poff[m];
  double *__restrict__ out = tab ? four2 : four;
  const int se = tab ? 1 : 0, so = 1 - se;
  const double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;

  for (int rt0 = 0; rt0 < P2p / 16; rt0 += RT) {
    d4 ae[RT], ao[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      ae[r] = d4{0.0, 0.0, 0.0, 0.0};
      ao[r] = d4{0.0, 0.0, 0.0, 0.0};
    }
    const double *te = T + (size_t)kq * P2p + rt0 * 16 + r16;
    const double *to = te + (size_t)KP * P2p;
#pragma unroll 2
    for (int k4 = 0; k4 < KP; k4 += 4) {
      const int k = k4 + kq;
      const double be = (cok && k < ke) ? sm[(size_t)(2 * k) * ncol + col] : 0.0;
      const double bo = (cok && k < ko) ? sm[(size_t)(2 * k + 1) * ncol + col] : 0.0;
      double a_e[RT], a_o[RT];
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        a_e[r] = te[(size_t)k4 * P2p + r * 16];
        a_o[r] = to[(size_t)k4 * P2p + r * 16];
      }
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        ae[r] = mfma_f64(a_e[r], be, ae[r]);
        ao[r] = mfma_f64(a_o[r], bo, ao[r]);
      }
    }
    if (cok) {
#pragma unroll
      for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int p = (rt0 + r) * 16 + i * 4 + kq;
          if (p < P2) {
            double *dst = out + ((size_t)(p * M1 + m) * 2) * ncol + col;
            dst[se * ncol] = ae[r][i];
            dst[so * ncol] = ao[r][i];
          }
        }
    }
  }
}

// Forward: four (S, A halves) -> spec.  MAXRT = row tiles of 16 degrees per
// parity at m = 0.
template <int MAXRT>
__global__ __launch_bounds__(64) void leg_fwd_f64_kernel(const double *__restrict__ four,
                                                         double *__restrict__ spec,
                                                         const double *__restrict__ Pfwd,
                                                         const int *__restrict__ poff, int N,
                                                         int P2, int P2p, int ncol, int nct) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int ct = b % nct, m = b / nct;
  const int M1 = N + 1;
  const int lane = threadIdx.x, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
#pragma unroll 1
  for (int par = 0; par < 2; ++par) {
    const int kc = par ? (nn >> 1) : ke;
    d4 acc[MAXRT];
#pragma unroll
    for (int r = 0; r < MAXRT; ++r) acc[r] = d4{0.0, 0.0, 0.0, 0.0};
    const double *__restrict__ W = Pfwd + poff[m] + (size_t)par * P2p * KR + (size_t)kq * KR + r16;
    const double *__restrict__ src = four + (size_t)par * ncol + col;
#pragma unroll 2
    for (int p4 = 0; p4 < P2p; p4 += 4) {
      const int p = p4 + kq;
      const double bv = (cok && p < P2) ? src[(size_t)(p * M1 + m) * 2 * ncol] : 0.0;
      double a[MAXRT];
#pragma unroll
      for (int r = 0; r < MAXRT; ++r) a[r] = (r < nrt) ? W[(size_t)p4 * KR + r * 16] : 0.0;
#pragma unroll
      for (int r = 0; r < MAXRT; ++r)
        if (r < nrt) acc[r] = mfma_f64(a[r], bv, acc[r]);
    }
    if (cok) {
#pragma unroll
      for (int r = 0; r < MAXRT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = r * 16 + i * 4 + kq;
          if (r < nrt && k < kc) sm[(size_t)(2 * k + par) * ncol + col] = acc[r][i];
        }
    }
  }
}

template <int RT>
int inv_launch(const Plan &p, const double *spec, double *four, double *four2, int tab0, int ntab,
               int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nblk = (p.N + 1) * ntab * nct;
  leg_inv_f64_kernel<RT><<<nblk, 64, 0, s>>>(spec, four, four2, p.Pinv, p.Hinv, p.poff_inv, p.N,
                                             p.P2, p.P2p, ncol, nct, tab0, ntab);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int MAXRT>
int fwd_launch(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nblk = (p.N + 1) * nct;
  leg_fwd_f64_kernel<MAXRT><<<nblk, 64, 0, s>>>(four, spec, p.Pfwd, p.poff_fwd, p.N, p.P2, p.P2p,
                                                ncol, nct);
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
  if (rtiles % 4 == 0) return inv_launch<4>(p, spec, four, four2, tab0, ntab, nf, s);
  if (rtiles % 2 == 0) return inv_launch<2>(p, spec, four, four2, tab0, ntab, nf, s);
  return inv_launch<1>(p, spec, four, four2, tab0, ntab, nf, s);
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

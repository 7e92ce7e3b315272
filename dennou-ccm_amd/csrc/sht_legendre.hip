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
// first and dealt to the XCDs round robin in launch order (consecutive
// workgroups go to consecutive XCDs), which gives every XCD an equal share of
// heavy and light m; contiguous ranges per XCD would give the first XCD twice
// the average work.
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
template <int RT, bool NT>
__global__ __launch_bounds__(256) void leg_inv_f64_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv,
    const int *__restrict__ poff, const int *__restrict__ mlist, int mlen, int N, int P2, int P2p,
    int ncol, int nct, int tab0, int ntab, FourGeom geo) {
  extern __shared__ double red[]; // [kWaves][RT][4][64]
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, ntab * nct, m, t)) return;
  const int ct = t % nct;
  const int tab = tab0 + t / nct;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
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
        if (p < P2 && c < ncol) put_four<NT>(out + geo.at(p, m, slot, c), v);
      }
      __syncthreads();
    }
  }
}

// Forward: four (S, A halves) -> spec.  Workgroup = one (m, 16-column tile);
// with swap the parity-q degrees read slot 1-q (tables of opposite parity,
// e.g. the divergence H table), with accum the result is added to spec.
// wave w takes the latitude-pair steps w, w+4, ... (4 pairs each) of the
// quadrature, the partial sums over its pairs meet in LDS and are added in
// FP64.  MAXRT = row tiles of 16 degrees per parity at m = 0.
template <int MAXRT>
__global__ __launch_bounds__(256) void leg_fwd_f64_kernel(const double *__restrict__ four,
                                                          double *__restrict__ spec,
                                                          const double *__restrict__ Pfwd,
                                                          const int *__restrict__ poff,
                                                          const int *__restrict__ mlist, int mlen,
                                                          int N, int P2, int P2p, int ncol, int nct,
                                                          FourGeom geo, int swap, int accum) {
  extern __shared__ double red[]; // [kWaves][MAXRT][4][64]
  int m, ct;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, nct, m, ct)) return;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
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
        const double x = four[geo.at(pb, m, q ^ swap, csafe)];
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
      if (k < kcnt[q] && c < ncol) {
        double *o = sm + (size_t)(2 * k + q) * ncol + c;
        put_spec(o, accum ? *o + v : v);
      }
    }
    __syncthreads();
  }
}

// Wave-tile shape: one wavefront owns one 16 x 16 output tile for both parities
// and runs the whole K loop itself, so there is no cross-wave reduction, no
// LDS and no barrier, and a wave retires as soon as its tile is stored.  The
// K loop is software-pipelined over chunks of D steps (4 K values each): the
// fragments of chunk i+1 are in flight while chunk i runs on the matrix core.
// Loads go to clamped, always-valid addresses and the B operand is masked, so
// every chunk is straight-line code and the compiler keeps a running wait count.
// Units are numbered with the column tile fastest, so the four waves of a
// workgroup share table fragments (same m and row tile) in L1.

// Inverse, unit = (m, table, row tile of 16 pairs, column tile); heaviest m first.
template <int D, bool NT>
__global__ __launch_bounds__(256) void leg_inv_f64_wave_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv,
    const int *__restrict__ poff, const int *__restrict__ mlist, int mlen, int N, int P2, int P2p,
    int ncol, int nct, int tab0, int ntab, FourGeom geo) {
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int rtiles = P2p >> 4;
  const int upm = rtiles * ntab * nct;  // waves per m
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (upm + 3) >> 2, m, t)) return;
  const int u = t * 4 + wave;
  if (u >= upm) return;
  const int ct = u % nct;
  const int rest = u / nct;
  const int rt = rest % rtiles;
  const int tab = tab0 + rest / rtiles;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP = (ke + 3) & ~3, nsteps = KP >> 2;
  const double *__restrict__ T = (tab ? Hinv : Pinv) + poff[m] + rt * 16 + r16;
  double *__restrict__ out = tab ? four2 : four;
  const int se = tab ? 1 : 0, so = 1 - se;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;
  const double *__restrict__ sp = spec + (size_t)coef_off(N, m) * ncol + (cok ? col : 0);

  double a[2][D][2], b[2][D][2];
  auto load = [&](int buf, int s0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int st = s0 + j, k = st * 4 + kq;
      const int kc = st < nsteps ? k : kq;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const bool vb = cok && k < (q ? ko : ke);
        a[buf][j][q] = T[(size_t)(q * KP + kc) * P2p];
        const double x = sp[(size_t)(vb ? 2 * k + q : 0) * ncol];
        b[buf][j][q] = vb ? x : 0.0;
      }
    }
  };
  d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  auto run = [&](int buf, int s0) {
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (s0 + j < nsteps) {
        acc0 = mfma_f64(a[buf][j][0], b[buf][j][0], acc0);
        acc1 = mfma_f64(a[buf][j][1], b[buf][j][1], acc1);
      }
  };
  load(0, 0);
  for (int s0 = 0; s0 < nsteps; s0 += 2 * D) {
    load(1, s0 + D);
    run(0, s0);
    if (s0 + D >= nsteps) break;
    load(0, s0 + 2 * D);
    run(1, s0 + D);
  }
  if (cok) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = rt * 16 + i * 4 + kq;
      if (p < P2) {
        put_four<NT>(out + geo.at(p, m, se, col), acc0[i]);
        put_four<NT>(out + geo.at(p, m, so, col), acc1[i]);
      }
    }
  }
}

// Forward, unit = (m, 16-degree row tile, column tile) from the plan's unit
// list (m << 8 | row tile, heaviest m first); K runs over latitude pairs.
template <int D>
__global__ __launch_bounds__(256) void leg_fwd_f64_wave_kernel(
    const double *__restrict__ four, double *__restrict__ spec, const double *__restrict__ Pfwd,
    const int *__restrict__ poff, const int *__restrict__ units, int N, int P2, int P2p, int ncol,
    int nct, int nunits, FourGeom geo) {
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int u = blockIdx.x * 4 + wave;
  if (u >= nunits) return;
  const int ct = u % nct;
  const int mk = units[u / nct];
  const int m = mk >> 8, kt = mk & 255;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KR = (ke + 15) & ~15;
  const int nsteps = P2p >> 2;
  const double *__restrict__ W = Pfwd + poff[m] + kt * 16 + r16;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;
  const int csafe = cok ? col : 0;

  double a[2][D][2], b[2][D][2];
  auto load = [&](int buf, int s0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int st = s0 + j, p = st * 4 + kq;
      const int pa = st < nsteps ? p : kq;
      const bool vb = cok && p < P2;
      const int pb = vb ? p : 0;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[buf][j][q] = W[(size_t)(q * P2p + pa) * KR];
        const double x = four[geo.at(pb, m, q, csafe)];
        b[buf][j][q] = vb ? x : 0.0;
      }
    }
  };
  d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  auto run = [&](int buf, int s0) {
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (s0 + j < nsteps) {
        acc0 = mfma_f64(a[buf][j][0], b[buf][j][0], acc0);
        acc1 = mfma_f64(a[buf][j][1], b[buf][j][1], acc1);
      }
  };
  load(0, 0);
  for (int s0 = 0; s0 < nsteps; s0 += 2 * D) {
    load(1, s0 + D);
    run(0, s0);
    if (s0 + D >= nsteps) break;
    load(0, s0 + 2 * D);
    run(1, s0 + D);
  }
  if (cok) {
    double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol + col;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kt * 16 + i * 4 + kq;
      if (k < ke) put_spec(sm + (size_t)(2 * k) * ncol, acc0[i]);
      if (k < ko) put_spec(sm + (size_t)(2 * k + 1) * ncol, acc1[i]);
    }
  }
}

// Tiled shape with a per-m K split.  A wave holds RT row tiles (RT*16 latitude
// pairs) of one 16-column tile in registers, so each spectral fragment it
// loads feeds RT MFMAs per parity.  The four waves of a workgroup split the K
// loop of one column tile ks ways and cover 4/ks column tiles, with ks chosen
// per m from the K length: 4 for the heavy low wavenumbers (short dependency
// chains), 1 for the light high ones (no reduction, more columns per group).
// Workgroups are numbered m-major, heaviest first: [0, m4) use ks = 4,
// [m4, m2) ks = 2, [m2, N] ks = 1 (SHT_SHAPE_TILE: m4 = m2 = 0, no split;
// SHT_SHAPE_TILE_SPLIT: thresholds on the K length).  Partial sums of a split
// meet in LDS in FP64.
template <int RT, int D, bool NT>
__global__ __launch_bounds__(256) void leg_inv_f64_tile_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv,
    const int *__restrict__ poff, const double *__restrict__ zero, int N, int P2, int P2p,
    int ncol, int nct, int tab0, int ntab, KSplit kspl, const int *__restrict__ mlist, int mlen,
    FourGeom geo) {
  extern __shared__ double red[];  // [4 waves][RT][4][64], one parity at a time
  const int ngr = (P2p >> 4) / RT;  // row groups
  const int w4 = ngr * ntab * nct, w2 = ngr * ntab * ((nct + 1) >> 1), w1 = ngr * ntab * ((nct + 3) >> 2);
  int b = blockIdx.x;
  int ks, m, cpw;
  if (kspl.m2 == 0) {  // no split: whole m per XCD from the balanced lists
    ks = 1; cpw = 4;
    if (!xcd_m_block(b, mlist, mlen, w1, m, b)) return;
  } else if (b < kspl.m4 * w4) {  // split regions: round robin over the XCDs
    ks = 4; cpw = 1; m = b / w4; b %= w4;
  } else if ((b -= kspl.m4 * w4) < (kspl.m2 - kspl.m4) * w2) {
    ks = 2; cpw = 2; m = kspl.m4 + b / w2; b %= w2;
  } else {
    b -= (kspl.m2 - kspl.m4) * w2;
    ks = 1; cpw = 4; m = kspl.m2 + b / w1; b %= w1;
  }
  const int ncg = (nct + cpw - 1) / cpw;
  const int ctg = b % ncg;
  b /= ncg;
  const int tab = tab0 + b % ntab;
  const int rg = b / ntab;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int kw = wave % ks, cw = wave / ks;
  const int ct = ctg * cpw + cw;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP = (ke + 3) & ~3, nsteps = KP >> 2;
  const double *__restrict__ T = (tab ? Hinv : Pinv) + poff[m] + rg * RT * 16 + r16;
  double *__restrict__ out = tab ? four2 : four;
  const int se = tab ? 1 : 0, so = 1 - se;
  const int col = ct * 16 + r16;
  const bool cok = ct < nct && col < ncol;
  const double *__restrict__ sp = spec + (size_t)coef_off(N, m) * ncol + (cok ? col : 0);

  d4 acc[2][RT];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int r = 0; r < RT; ++r) acc[q][r] = d4{0.0, 0.0, 0.0, 0.0};
  // This wave's K steps are kw, kw + ks, ... < nsteps, taken in chunks of D
  // whose loads are all issued before the chunk's first MFMA.  Steps past the
  // end load clamped addresses with a zero B operand, so every MFMA is
  // unconditional (adding exact zeros) and no load can sink below a branch.
  double a[1][D][2][RT], bv[1][D][2];
  auto load = [&](int buf, int j0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int st = kw + ks * (j0 + j), k = st * 4 + kq;
      const int kc = st < nsteps ? k : kq;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const bool vb = cok && k < (q ? ko : ke);
        bv[buf][j][q] = *(vb ? sp + (size_t)(2 * k + q) * ncol : zero);
#pragma unroll
        for (int r = 0; r < RT; ++r) a[buf][j][q][r] = T[(size_t)(q * KP + kc) * P2p + r * 16];
      }
    }
  };
  auto run = [&](int buf) {
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < RT; ++r) acc[q][r] = mfma_f64(a[buf][j][q][r], bv[buf][j][q], acc[q][r]);
  };
  for (int j0 = 0; kw + ks * j0 < nsteps; j0 += D) {
    load(0, j0);
    run(0);
  }
  if (ks == 1) {
    if (cok) {
#pragma unroll
      for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int p = (rg * RT + r) * 16 + i * 4 + kq;
          if (p < P2) {
            put_four<NT>(out + geo.at(p, m, se, col), acc[0][r][i]);
            put_four<NT>(out + geo.at(p, m, so, col), acc[1][r][i]);
          }
        }
    }
    return;
  }
  // ks = 2 or 4 (uniform over the workgroup): FP64 sums of the ks partials
  const int per = RT * 256;  // doubles one wave holds per parity
#pragma unroll
  for (int q = 0; q < 2; ++q) {
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[wave * per + (r * 4 + i) * 64 + lane] = acc[q][r][i];
    __syncthreads();
    const int slot = q ? so : se;
    for (int e = threadIdx.x; e < cpw * per; e += 256) {
      const int g = e / per, el = e % per;
      double v = red[(g * ks) * per + el];
      for (int t = 1; t < ks; ++t) v += red[(g * ks + t) * per + el];
      const int ln = el & 63, ri = el >> 6, i = ri & 3, r = ri >> 2;
      const int p = (rg * RT + r) * 16 + i * 4 + (ln >> 4);
      const int c = (ctg * cpw + g) * 16 + (ln & 15);
      if (p < P2 && c < ncol) put_four<NT>(out + geo.at(p, m, slot, c), v);
    }
    __syncthreads();
  }
}

// Splits from the per-m K length: ks = 4 where nsteps(m) >= t4, 2 where >= t2.
inline KSplit ksplit_inv(const Plan &p, int t4, int t2) {
  KSplit k{0, 0};
  for (int m = 0; m <= p.N; ++m) {
    const int nsteps = roundup(ceil_half(p.N - m + 1), 4) / 4;
    if (nsteps >= t4) k.m4 = m + 1;
    if (nsteps >= t2) k.m2 = m + 1;
  }
  if (k.m2 < k.m4) k.m2 = k.m4;
  return k;
}

template <int RT, int D>
int inv_tile_launch(const Plan &p, KSplit ks, const double *spec, double *four, double *four2,
                    int tab0, int ntab, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int ngr = (p.P2p / 16) / RT;
  const long w4 = (long)ngr * ntab * nct, w2 = (long)ngr * ntab * ((nct + 1) / 2),
             w1 = (long)ngr * ntab * ((nct + 3) / 4);
  const long nblk = ks.m2 == 0 ? xcd_m_blocks(p, 0, (int)w1)
                               : ks.m4 * w4 + (long)(ks.m2 - ks.m4) * w2 + (long)(p.N + 1 - ks.m2) * w1;
  const size_t lds = ks.m2 > 0 ? sizeof(double) * 4 * RT * 256 : 0;
  const FourGeom geo = four_geom_out(p, nf, sizeof(double));
  launch_kernel(geo.stream ? leg_inv_f64_tile_kernel<RT, D, true> : leg_inv_f64_tile_kernel<RT, D, false>,
      (int)nblk, 256, lds, s, spec, four, four2, p.Pinv, p.Hinv, p.poff_inv, p.zero, p.N, p.P2,
      p.P2p, ncol, nct, tab0, ntab, ks, p.xcd_m[0], p.xcd_mlen[0], geo);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int D>
int inv_wave_launch(const Plan &p, const double *spec, double *four, double *four2, int tab0,
                    int ntab, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int upm = ntab * (p.P2p / 16) * nct;
  const FourGeom geo = four_geom_out(p, nf, sizeof(double));
  launch_kernel(geo.stream ? leg_inv_f64_wave_kernel<D, true> : leg_inv_f64_wave_kernel<D, false>,
      xcd_m_blocks(p, 0, (upm + 3) / 4), 256, 0, s, spec, four, four2, p.Pinv, p.Hinv, p.poff_inv,
      p.xcd_m[0], p.xcd_mlen[0], p.N, p.P2, p.P2p, ncol, nct, tab0, ntab, geo);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

// Forward counterpart: a wave holds all 16-degree row tiles of its m (at most
// MAXRT) for one column tile, so each Fourier fragment feeds every row tile;
// the latitude-pair loop is split ks ways per m as in the inverse, with ks
// chosen from the number of row tiles (the work of an m).
template <int MAXRT, int D>
__global__ __launch_bounds__(256) void leg_fwd_f64_tile_kernel(
    const double *__restrict__ four, double *__restrict__ spec, const double *__restrict__ Pfwd,
    const int *__restrict__ poff, const double *__restrict__ zero, int N, int P2, int P2p,
    int ncol, int nct, KSplit kspl, const int *__restrict__ mlist, int mlen, FourGeom geo) {
  extern __shared__ double red[];  // [4 waves][MAXRT][4][64], one parity at a time
  const int w4 = nct, w2 = (nct + 1) >> 1, w1 = (nct + 3) >> 2;
  int b = blockIdx.x;
  int ks, m, cpw;
  if (kspl.m2 == 0) {  // no split: whole m per XCD from the balanced lists
    ks = 1; cpw = 4;
    if (!xcd_m_block(b, mlist, mlen, w1, m, b)) return;
  } else if (b < kspl.m4 * w4) {  // split regions: round robin over the XCDs
    ks = 4; cpw = 1; m = b / w4; b %= w4;
  } else if ((b -= kspl.m4 * w4) < (kspl.m2 - kspl.m4) * w2) {
    ks = 2; cpw = 2; m = kspl.m4 + b / w2; b %= w2;
  } else {
    b -= (kspl.m2 - kspl.m4) * w2;
    ks = 1; cpw = 4; m = kspl.m2 + b / w1; b %= w1;
  }
  const int ctg = b;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int kw = wave % ks, cw = wave / ks;
  const int ct = ctg * cpw + cw;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  const int nsteps = P2p >> 2;
  const double *__restrict__ W = Pfwd + poff[m] + r16;
  const int col = ct * 16 + r16;
  const bool cok = ct < nct && col < ncol;
  const int csafe = cok ? col : 0;

  d4 acc[2][MAXRT];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int r = 0; r < MAXRT; ++r) acc[q][r] = d4{0.0, 0.0, 0.0, 0.0};
  // Chunks of D latitude-pair steps as in the inverse.  Row tiles past this
  // m's last (r >= nrt) reuse tile 0's fragment and their results are never
  // stored: unconditional MFMAs keep every load early.
  double a[1][D][2][MAXRT], bv[1][D][2];
  auto load = [&](int buf, int j0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int st = kw + ks * (j0 + j), p = st * 4 + kq;
      const int pa = st < nsteps ? p : kq;
      const bool vb = cok && p < P2;
      const int pb = vb ? p : 0;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        bv[buf][j][q] = *(vb ? four + geo.at(pb, m, q, csafe) : zero);
#pragma unroll
        for (int r = 0; r < MAXRT; ++r)
          a[buf][j][q][r] = W[(size_t)(q * P2p + pa) * KR + (r < nrt ? r : 0) * 16];
      }
    }
  };
  auto run = [&](int buf) {
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < MAXRT; ++r) acc[q][r] = mfma_f64(a[buf][j][q][r], bv[buf][j][q], acc[q][r]);
  };
  for (int j0 = 0; kw + ks * j0 < nsteps; j0 += D) {
    load(0, j0);
    run(0);
  }
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
  const int kcnt[2] = {ke, ko};
  if (ks == 1) {
    if (cok) {
#pragma unroll
      for (int r = 0; r < MAXRT; ++r)
        if (r < nrt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int k = r * 16 + i * 4 + kq;
            if (k < ke) put_spec(sm + (size_t)(2 * k) * ncol + col, acc[0][r][i]);
            if (k < ko) put_spec(sm + (size_t)(2 * k + 1) * ncol + col, acc[1][r][i]);
          }
    }
    return;
  }
  const int per = nrt * 256;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
#pragma unroll
    for (int r = 0; r < MAXRT; ++r)
      if (r < nrt)
#pragma unroll
        for (int i = 0; i < 4; ++i) red[wave * per + (r * 4 + i) * 64 + lane] = acc[q][r][i];
    __syncthreads();
    for (int e = threadIdx.x; e < cpw * per; e += 256) {
      const int g = e / per, el = e % per;
      double v = red[(g * ks) * per + el];
      for (int t = 1; t < ks; ++t) v += red[(g * ks + t) * per + el];
      const int ln = el & 63, ri = el >> 6, i = ri & 3, r = ri >> 2;
      const int k = r * 16 + i * 4 + (ln >> 4);
      const int c = (ctg * cpw + g) * 16 + (ln & 15);
      if (k < kcnt[q] && c < ncol) put_spec(sm + (size_t)(2 * k + q) * ncol + c, v);
    }
    __syncthreads();
  }
}

// ks = 4 where the m has at least t4 row tiles, 2 where at least t2.
inline KSplit ksplit_fwd(const Plan &p, int t4, int t2) {
  KSplit k{0, 0};
  for (int m = 0; m <= p.N; ++m) {
    const int nrt = roundup(ceil_half(p.N - m + 1), 16) / 16;
    if (nrt >= t4) k.m4 = m + 1;
    if (nrt >= t2) k.m2 = m + 1;
  }
  if (k.m2 < k.m4) k.m2 = k.m4;
  return k;
}

template <int MAXRT, int D>
int fwd_tile_launch(const Plan &p, KSplit ks, const double *four, double *spec, int nf,
                    hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const long nblk = ks.m2 == 0 ? xcd_m_blocks(p, 1, (nct + 3) / 4)
                               : (long)ks.m4 * nct + (long)(ks.m2 - ks.m4) * ((nct + 1) / 2) +
                                     (long)(p.N + 1 - ks.m2) * ((nct + 3) / 4);
  const size_t lds = ks.m2 > 0 ? sizeof(double) * 4 * MAXRT * 256 : 0;
  launch_kernel(leg_fwd_f64_tile_kernel<MAXRT, D>, (int)nblk, 256, lds, s,
      four, spec, p.Pfwd, p.poff_fwd, p.zero, p.N, p.P2, p.P2p, ncol, nct, ks, p.xcd_m[1],
      p.xcd_mlen[1], four_geom(p, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int D>
int fwd_wave_launch(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nunits = p.n_fwd_units * nct;
  launch_kernel(leg_fwd_f64_wave_kernel<D>, (nunits + 3) / 4, 256, 0, s,
      four, spec, p.Pfwd, p.poff_fwd, p.fwd_units, p.N, p.P2, p.P2p, ncol, nct, nunits,
      four_geom(p, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

// Staged shape: the workgroup copies the data block of its (m, CT column
// tiles) into LDS once -- for the forward all latitude pairs and both
// parities of the Fourier data, for the inverse all degrees of the spectra --
// and wave w then takes the (parity, 16-row tile) units w, w+NW, ... and runs
// each over the whole K loop: table fragments from global memory (one per
// step, used for all CT column tiles), data fragments from LDS.  No partial
// sums cross waves, and a wave holds only CT accumulators and its table
// pipeline, so occupancy is set by the LDS block alone.  Built for the large
// truncations, where the other shapes re-read the data tile per row tile or
// hold every row tile's accumulators at once.  Per output tile the MFMA
// sequence depends neither on NW nor on CT.
//
// NW waves share the block: from P2p = 128 latitude pairs (the T170 grid)
// eight, and the inverse there takes two column tiles (each table fragment
// feeds two MFMAs, halving the table traffic from L2); below that four waves
// and one column tile.  Measured at T170, 776 fields: inverse 248 against
// 273 us with four waves and one tile, forward 277 against 295 us (two tiles
// for the forward: 276 us, but slower at 97 and 388 fields); eight waves were
// slower at T42 and T85.  Measured no faster: table loads 8 or 16 steps
// ahead instead of 4, two accumulator chains per unit, staging rows
// parity-fastest or in one round of loads; slower: holding a wave's results
// until its last unit with a chunk's LDS data read ahead (T170 forward 320
// against 296 us).
//
// Copies rows of 16 columns with NT threads: element e of the block is the
// column pair c2 = e % 8 of row r = (e / 8) % nrow of column tile
// j = e / (8 nrow); src(j, r, col) returns columns col, col + 1 (col even) of
// the row, zeros outside the data.  Loads are issued NQ at a time before
// their LDS stores.
template <int NQ, int NT, class Src>
__device__ __forceinline__ void stage_rows(double *sb, int nrow, int nj, Src src) {
  const int nel = nj * nrow * 8;
  for (int e0 = threadIdx.x; e0 < nel; e0 += NQ * NT) {
    double2 v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = e0 + q * NT, c2 = e & 7, r = (e >> 3) % nrow, j = (e >> 3) / nrow;
      v[q] = e < nel ? src(j, r, 2 * c2) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = e0 + q * NT;
      if (e < nel) *reinterpret_cast<double2 *>(sb + 2 * e) = v[q];
    }
  }
}

// Runs one unit: acc[c] += A(s) B_c(s) over the steps s < nsteps of 4 K
// values each, in order; a_at(s) loads the lane's table value of step s from
// global memory (called with s < nsteps only: steps past the end load step
// 0 and are not run), b_at(c, s) reads column tile c's data from LDS.  Table
// loads run D steps ahead.
template <int CT, int D, class ALoad, class BLoad>
__device__ __forceinline__ void staged_unit(int nsteps, ALoad a_at, BLoad b_at, d4 (&acc)[CT]) {
  double a[2][D];
  auto load = [&](int buf, int s0) {
#pragma unroll
    for (int j = 0; j < D; ++j) a[buf][j] = a_at(s0 + j < nsteps ? s0 + j : 0);
  };
  auto run = [&](int buf, int s0) {
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (s0 + j < nsteps)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = mfma_f64(a[buf][j], b_at(c, s0 + j), acc[c]);
  };
  load(0, 0);
  for (int s0 = 0; s0 < nsteps; s0 += 2 * D) {
    load(1, s0 + D);
    run(0, s0);
    if (s0 + D >= nsteps) break;
    load(0, s0 + 2 * D);
    run(1, s0 + D);
  }
}

// Forward, workgroup = (m, CT column tiles); LDS [CT][parity][P2p pairs][16].
template <int CT, int D, int NW>
__global__ __launch_bounds__(NW * 64) void leg_fwd_f64_staged_kernel(
    const double *__restrict__ four, double *__restrict__ spec, const double *__restrict__ Pfwd,
    const int *__restrict__ poff, const int *__restrict__ mlist, int mlen, int N, int P2, int P2p,
    int ncol, int nct, FourGeom geo) {
  extern __shared__ double sb[];
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + CT - 1) / CT, m, t)) return;
  const int ct0 = t * CT;
  // the Fourier scratch is 16-byte aligned and its column runs are even
  stage_rows<4, NW * 64>(sb, 2 * P2p, CT, [&](int j, int r, int c) {
    const int par = r / P2p, p = r - par * P2p, col = (ct0 + j) * 16 + c;
    return (p < P2 && col < ncol) ? *reinterpret_cast<const double2 *>(four + geo.at(p, m, par, col))
                                  : make_double2(0.0, 0.0);
  });
  __syncthreads();
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
  for (int u = wave; u < 2 * nrt; u += NW) {
    const int q = u & 1, rt = u >> 1;
    const double *__restrict__ W = Pfwd + poff[m] + (size_t)q * P2p * KR + rt * 16 + r16;
    const double *B = sb + (q * P2p + kq) * 16 + r16;
    d4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
    staged_unit<CT, D>(
        P2p >> 2, [&](int s) { return W[(size_t)(s * 4 + kq) * KR]; },
        [&](int c, int s) { return B[(c * 2 * P2p + s * 4) * 16]; }, acc);
    const int kn = q ? ko : ke;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int col = (ct0 + c) * 16 + r16;
      if (col < ncol)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = rt * 16 + i * 4 + kq;
          if (k < kn) put_spec(sm + (size_t)(2 * k + q) * ncol + col, acc[c][i]);
        }
    }
  }
}

// Inverse, workgroup = (m, table, CT column tiles); LDS [CT][parity][KP][16]
// with KP = this m's degrees per parity rounded up to 4 (rows past a parity's
// last degree hold zeros).
template <int CT, int D, int NW, bool NT>
__global__ __launch_bounds__(NW * 64) void leg_inv_f64_staged_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv, const int *__restrict__ poff,
    const int *__restrict__ mlist, int mlen, int N, int P2, int P2p, int ncol, int nct, int tab0,
    int ntab, FourGeom geo) {
  extern __shared__ double sb[];
  const int ncg = (nct + CT - 1) / CT;
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, ntab * ncg, m, t)) return;
  const int tab = tab0 + t % ntab, ct0 = (t / ntab) * CT;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP = (ke + 3) & ~3;
  const double *__restrict__ sp = spec + (size_t)coef_off(N, m) * ncol;
  // the caller's spectra may be only 8-byte aligned: two loads per pair
  stage_rows<4, NW * 64>(sb, 2 * KP, CT, [&](int j, int r, int c) {
    const int par = r / KP, k = r - par * KP, col = (ct0 + j) * 16 + c;
    if (k >= (par ? ko : ke) || col >= ncol) return make_double2(0.0, 0.0);
    const double *x = sp + (size_t)(2 * k + par) * ncol + col;
    return make_double2(x[0], x[1]);
  });
  __syncthreads();
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int rtiles = P2p >> 4;
  double *__restrict__ out = tab ? four2 : four;
  const int se = tab ? 1 : 0;  // the derivative table's parities are swapped
  for (int u = wave; u < 2 * rtiles; u += NW) {
    const int q = u & 1, rt = u >> 1;
    const double *__restrict__ T = (tab ? Hinv : Pinv) + poff[m] + (size_t)q * KP * P2p + rt * 16 + r16;
    const double *B = sb + (q * KP + kq) * 16 + r16;
    d4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
    staged_unit<CT, D>(
        KP >> 2, [&](int s) { return T[(size_t)(s * 4 + kq) * P2p]; },
        [&](int c, int s) { return B[(c * 2 * KP + s * 4) * 16]; }, acc);
    const int slot = q ^ se;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int col = (ct0 + c) * 16 + r16;
      if (col < ncol)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int p = rt * 16 + i * 4 + kq;
          if (p < P2) put_four<NT>(out + geo.at(p, m, slot, col), acc[c][i]);
        }
    }
  }
}

template <int CT, int NW>
int fwd_staged_launch(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const size_t lds = sizeof(double) * CT * 2 * p.P2p * 16;  // <= 64 KB: P2p <= 256
  launch_kernel(leg_fwd_f64_staged_kernel<CT, 4, NW>, xcd_m_blocks(p, 1, (nct + CT - 1) / CT), NW * 64, lds,
      s, four, spec, p.Pfwd, p.poff_fwd, p.xcd_m[1], p.xcd_mlen[1], p.N, p.P2, p.P2p, ncol, nct,
      four_geom(p, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int CT, int NW>
int inv_staged_launch(const Plan &p, const double *spec, double *four, double *four2, int tab0,
                      int ntab, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int kp0 = roundup(ceil_half(p.N + 1), 4);
  const size_t lds = sizeof(double) * CT * 2 * kp0 * 16;  // <= 64 KB: kp0 <= 128
  const FourGeom geo = four_geom_out(p, nf, sizeof(double));
  launch_kernel(geo.stream ? leg_inv_f64_staged_kernel<CT, 4, NW, true>
                           : leg_inv_f64_staged_kernel<CT, 4, NW, false>,
      xcd_m_blocks(p, 0, ntab * ((nct + CT - 1) / CT)), NW * 64, lds, s, spec, four, four2, p.Pinv, p.Hinv,
      p.poff_inv, p.xcd_m[0], p.xcd_mlen[0], p.N, p.P2, p.P2p, ncol, nct, tab0, ntab, geo);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

// Staged in two parts: the same shape for large batches on the T170 grid.
// The K range (latitude pairs for the forward, degree steps of each parity
// for the inverse) is staged and run in two halves, so the LDS block of half
// the K range holds more column tiles -- two for the forward, four for the
// inverse -- and each table fragment feeds that many MFMAs.  The steps of an
// output tile still run in order from a zero accumulator, so results are
// bitwise those of the one-part kernels; a wave's MAXU units keep their
// accumulators across the halves and store after the second.  Measured at
// T170 (eight waves): 776 fields forward 261 against 280 us, inverse 234
// against 254 us; 388 fields 144 / 136.5 against 148 / 143.5 us; at 97 fields
// slower (37-40 against 34-39 us), hence the batch threshold at dispatch.
// Forward: LDS [CT][parity][P2p / 2 pairs][16].
template <int CT, int D, int NW, int MAXU>
__global__ __launch_bounds__(NW * 64) void leg_fwd_f64_halves_kernel(
    const double *__restrict__ four, double *__restrict__ spec, const double *__restrict__ Pfwd,
    const int *__restrict__ poff, const int *__restrict__ mlist, int mlen, int N, int P2, int P2p,
    int ncol, int nct, FourGeom geo) {
  extern __shared__ double sb[];
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + CT - 1) / CT, m, t)) return;
  const int ct0 = t * CT, H = P2p >> 1;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  d4 acc[MAXU][CT];
#pragma unroll
  for (int i = 0; i < MAXU; ++i)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[i][c] = d4{0.0, 0.0, 0.0, 0.0};
  for (int h = 0; h < 2; ++h) {
    if (h) __syncthreads();
    stage_rows<4, NW * 64>(sb, 2 * H, CT, [&](int j, int r, int c) {
      const int par = r / H, p = h * H + (r - par * H), col = (ct0 + j) * 16 + c;
      return (p < P2 && col < ncol) ? *reinterpret_cast<const double2 *>(four + geo.at(p, m, par, col))
                                    : make_double2(0.0, 0.0);
    });
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXU; ++i) {
      const int u = wave + NW * i, q = u & 1, rt = u >> 1;
      if (u >= 2 * nrt) break;
      const double *__restrict__ W = Pfwd + poff[m] + ((size_t)q * P2p + h * H) * KR + rt * 16 + r16;
      const double *B = sb + (q * H + kq) * 16 + r16;
      staged_unit<CT, D>(
          H >> 2, [&](int s) { return W[(size_t)(s * 4 + kq) * KR]; },
          [&](int c, int s) { return B[(c * 2 * H + s * 4) * 16]; }, acc[i]);
    }
  }
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
#pragma unroll
  for (int i = 0; i < MAXU; ++i) {
    const int u = wave + NW * i, q = u & 1, rt = u >> 1;
    if (u >= 2 * nrt) break;
    const int kn = q ? ko : ke;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int col = (ct0 + c) * 16 + r16;
      if (col < ncol)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = rt * 16 + j * 4 + kq;
          if (k < kn) put_spec(sm + (size_t)(2 * k + q) * ncol + col, acc[i][c][j]);
        }
    }
  }
}

// Inverse: steps S = KP / 4 of each parity split S0 = ceil(S / 2) + the rest;
// LDS [CT][parity][4 S0 rows][16].
template <int CT, int D, int NW, int MAXU, bool NT>
__global__ __launch_bounds__(NW * 64) void leg_inv_f64_halves_kernel(
    const double *__restrict__ spec, double *__restrict__ four, double *__restrict__ four2,
    const double *__restrict__ Pinv, const double *__restrict__ Hinv, const int *__restrict__ poff,
    const int *__restrict__ mlist, int mlen, int N, int P2, int P2p, int ncol, int nct, int tab0,
    int ntab, FourGeom geo) {
  extern __shared__ double sb[];
  const int ncg = (nct + CT - 1) / CT;
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, ntab * ncg, m, t)) return;
  const int tab = tab0 + t % ntab, ct0 = (t / ntab) * CT;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP = (ke + 3) & ~3, S = KP >> 2, S0 = (S + 1) >> 1;
  const double *__restrict__ sp = spec + (size_t)coef_off(N, m) * ncol;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int rtiles = P2p >> 4;
  double *__restrict__ out = tab ? four2 : four;
  const int se = tab ? 1 : 0;  // the derivative table's parities are swapped
  d4 acc[MAXU][CT];
#pragma unroll
  for (int i = 0; i < MAXU; ++i)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[i][c] = d4{0.0, 0.0, 0.0, 0.0};
  for (int h = 0; h < 2; ++h) {
    const int slo = h ? S0 : 0, ns = h ? S - S0 : S0, rows = 4 * ns;
    if (ns == 0) break;
    if (h) __syncthreads();
    stage_rows<4, NW * 64>(sb, 2 * rows, CT, [&](int j, int r, int c) {
      const int par = r / rows, k = 4 * slo + (r - par * rows), col = (ct0 + j) * 16 + c;
      if (k >= (par ? ko : ke) || col >= ncol) return make_double2(0.0, 0.0);
      const double *x = sp + (size_t)(2 * k + par) * ncol + col;
      return make_double2(x[0], x[1]);
    });
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXU; ++i) {
      const int u = wave + NW * i, q = u & 1, rt = u >> 1;
      if (u >= 2 * rtiles) break;
      const double *__restrict__ T =
          (tab ? Hinv : Pinv) + poff[m] + ((size_t)q * KP + 4 * slo) * P2p + rt * 16 + r16;
      const double *B = sb + (q * rows + kq) * 16 + r16;
      staged_unit<CT, D>(
          ns, [&](int s) { return T[(size_t)(s * 4 + kq) * P2p]; },
          [&](int c, int s) { return B[(c * 2 * rows + s * 4) * 16]; }, acc[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < MAXU; ++i) {
    const int u = wave + NW * i, q = u & 1, rt = u >> 1;
    if (u >= 2 * rtiles) break;
    const int slot = q ^ se;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int col = (ct0 + c) * 16 + r16;
      if (col < ncol)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int p = rt * 16 + j * 4 + kq;
          if (p < P2) put_four<NT>(out + geo.at(p, m, slot, col), acc[i][c][j]);
        }
    }
  }
}

// Field count from which the staged shape runs in two parts on the T170 grid.
constexpr int kHalvesMinFields = 192;

// MAXU = 2: at most 2 nrt(0) <= 16 forward units, and 2 (P2p / 16) = 16
// inverse units at P2p = 128, over eight waves.
template <int CT, int MAXU>
int fwd_halves_launch(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const size_t lds = sizeof(double) * CT * 2 * (p.P2p / 2) * 16;
  launch_kernel(leg_fwd_f64_halves_kernel<CT, 4, 8, MAXU>, xcd_m_blocks(p, 1, (nct + CT - 1) / CT), 512, lds,
      s, four, spec, p.Pfwd, p.poff_fwd, p.xcd_m[1], p.xcd_mlen[1], p.N, p.P2, p.P2p, ncol, nct,
      four_geom(p, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}
template <int CT, int MAXU>
int inv_halves_launch(const Plan &p, const double *spec, double *four, double *four2, int tab0,
                      int ntab, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int S = roundup(ceil_half(p.N + 1), 4) / 4, S0 = (S + 1) / 2;
  const size_t lds = sizeof(double) * CT * 2 * (4 * S0) * 16;
  const FourGeom geo = four_geom_out(p, nf, sizeof(double));
  launch_kernel(geo.stream ? leg_inv_f64_halves_kernel<CT, 4, 8, MAXU, true>
                           : leg_inv_f64_halves_kernel<CT, 4, 8, MAXU, false>,
      xcd_m_blocks(p, 0, ntab * ((nct + CT - 1) / CT)), 512, lds, s, spec, four, four2, p.Pinv, p.Hinv,
      p.poff_inv, p.xcd_m[0], p.xcd_mlen[0], p.N, p.P2, p.P2p, ncol, nct, tab0, ntab, geo);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int RT>
int inv_launch(const Plan &p, const double *spec, double *four, double *four2, int tab0, int ntab,
               int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nblk = xcd_m_blocks(p, 0, ntab * nct);
  const size_t lds = sizeof(double) * kWaves * RT * 256;
  const FourGeom geo = four_geom_out(p, nf, sizeof(double));
  launch_kernel(geo.stream ? leg_inv_f64_kernel<RT, true> : leg_inv_f64_kernel<RT, false>, nblk, 256,
                lds, s, spec, four, four2, p.Pinv, p.Hinv, p.poff_inv, p.xcd_m[0], p.xcd_mlen[0], p.N,
                p.P2, p.P2p, ncol, nct, tab0, ntab, geo);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int MAXRT>
int fwd_launch(const Plan &p, const double *four, double *spec, int nf, hipStream_t s,
               const double *table = nullptr, bool swap = false, bool accum = false) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const int nblk = xcd_m_blocks(p, 1, nct);
  const size_t lds = sizeof(double) * kWaves * MAXRT * 256;
  launch_kernel(leg_fwd_f64_kernel<MAXRT>, nblk, 256, lds, s, four, spec, table ? table : p.Pfwd,
                p.poff_fwd, p.xcd_m[1], p.xcd_mlen[1], p.N, p.P2, p.P2p, ncol, nct,
                four_geom(p, nf), swap ? 1 : 0, accum ? 1 : 0);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace

KSplit default_classes(const Plan &p, int dir) {
  if (dir == 0) {
    // halve the K loop of the wavenumbers with at least 70 % of the heaviest
    // one's K steps (the waves that set the critical path); no four-way
    // split.  Measured at T85: 97 fields 10.9 us (thresholds 6/3 gave 15.4,
    // TILE 13.5); T42, 388 fields 7.2 us (9.1, TILE 12.1).
    const int nsmax = roundup(ceil_half(p.N + 1), 4) / 4;
    return ksplit_inv(p, 1 << 30, (7 * nsmax + 9) / 10);
  }
  // halve the pair loop of the wavenumbers with the most row tiles only
  // (thresholds 2/1, which split every m, were slower everywhere)
  return ksplit_fwd(p, 1 << 30, roundup(ceil_half(p.N + 1), 16) / 16);
}

KSplit resolve_classes(const Plan &p, int nf, int dir) {
  for (const Plan::ShapeEntry &e : p.shapes)
    if (e.nf == nf) {
      const int m4 = dir ? e.t.fwd_m4 : e.t.inv_m4, m2 = dir ? e.t.fwd_m2 : e.t.inv_m2;
      if (m4 >= 0 && m2 >= m4 && m2 <= p.N + 1) return KSplit{m4, m2};
      break;
    }
  return default_classes(p, dir);
}

int resolve_shape(const Plan &p, int nf, int dir) {
  // an entry for nf, unless its half for this direction is AUTO
  for (const Plan::ShapeEntry &e : p.shapes)
    if (e.nf == nf) {
      const int s = dir ? e.t.fwd : e.t.inv;
      if (s != SHT_SHAPE_AUTO) return s;
      break;
    }
  const int d = dir ? p.shape_fwd : p.shape_inv;
  if (d != SHT_SHAPE_AUTO) return d;
  // mixed forms: the shapes give bitwise equal results, so the choice is by
  // speed alone -- two column tiles per workgroup from 128 latitude pairs
  // (T170: faster in both directions at 97 to 776 fields), one below.  On
  // the 128-pair grid the inverse streams (STAGED): always in the
  // six-product form, up to 8 fields in the two-piece forms, which run
  // KSPLIT up to 24 fields.  T170 inverse Legendre step, KSPLIT / TILE /
  // STAGED, us: x6 9 fields 8.5 / 10.7 / 8.1, 33 fields 15.9 / 14.2 /
  // 13.6, 97 fields 30.9 / 29.7 / 25.9, 388 fields 107 / 98.8 / 82.7, 776
  // fields 263 / 234 / 213; BF16 x3 4 fields 6.3 / 6.4 / 4.9, 17 fields
  // 8.8 / 9.8 / 9.0, 97 fields 23.4 / 21.1 / 21.9, 776 fields 221 / 182 /
  // 190.  Below 128 pairs the four idle table waves cost more than the
  // stream saves (T85 x6, 97 fields: KSPLIT 8.3, STAGED 10.1 us).
  if (p.prec != SHT_PREC_FP64) {
    if (dir == 0 && p.P2p == 128 && mixed_stream_available(p)) {
      if (p.prec == SHT_PREC_SPLIT_BF16X6 || nf <= 8) return SHT_SHAPE_STAGED;
      if (nf <= 24) return SHT_SHAPE_KSPLIT;
    }
    return p.P2p >= 128 ? SHT_SHAPE_TILE : SHT_SHAPE_KSPLIT;
  }
  // measured on MI355X from T21 to T170 and 16 to 388 fields: the staged
  // shape wins the forward nearly everywhere and the inverse from 32 fields;
  // below that a wavefront per tile fills the chip better.  STAGED and WAVE
  // run the same MFMA sequence per output tile, so the switch with the field
  // count changes no result.
  if (dir == 1) return SHT_SHAPE_STAGED;
  return nf >= 32 ? SHT_SHAPE_STAGED : SHT_SHAPE_WAVE;
}

int launch_leg_inv_f64_shape(const Plan &p, int shape, const double *spec, double *four,
                             double *four2, bool value_table, bool deriv_table, int nf,
                             hipStream_t s, KSplit cls) {
  if (!value_table && !deriv_table) return SHT_OK;
  const int tab0 = value_table ? 0 : 1;
  const int ntab = (value_table && deriv_table) ? 2 : 1;
  if (shape == SHT_SHAPE_WAVE) return inv_wave_launch<4>(p, spec, four, four2, tab0, ntab, nf, s);
  // two-part staging for large batches on the 128-pair grid (T170)
  if (shape == SHT_SHAPE_STAGED && p.P2p == 128 && nf >= kHalvesMinFields)
    return inv_halves_launch<4, 2>(p, spec, four, four2, tab0, ntab, nf, s);
  if (shape == SHT_SHAPE_STAGED)
    return p.P2p >= 128 ? inv_staged_launch<2, 8>(p, spec, four, four2, tab0, ntab, nf, s)
                        : inv_staged_launch<1, 4>(p, spec, four, four2, tab0, ntab, nf, s);
  if (shape == SHT_SHAPE_TILE || shape == SHT_SHAPE_TILE_SPLIT) {
    const int rtiles = p.P2p / 16;
    int rt = 4;
    while (rtiles % rt) --rt;
    // TILE_SPLIT: the given wavenumber classes, else the built-in ones
    const KSplit k = shape == SHT_SHAPE_TILE ? KSplit{0, 0} : (cls.m2 >= 0 ? cls : default_classes(p, 0));
    switch (rt) {
      case 4: return inv_tile_launch<4, 2>(p, k, spec, four, four2, tab0, ntab, nf, s);
      case 3: return inv_tile_launch<3, 2>(p, k, spec, four, four2, tab0, ntab, nf, s);
      case 2: return inv_tile_launch<2, 2>(p, k, spec, four, four2, tab0, ntab, nf, s);
      default: return inv_tile_launch<1, 2>(p, k, spec, four, four2, tab0, ntab, nf, s);
    }
  }
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

int launch_leg_fwd_f64_shape(const Plan &p, int shape, const double *four, double *spec, int nf,
                             hipStream_t s, KSplit cls) {
  if (shape == SHT_SHAPE_WAVE) return fwd_wave_launch<4>(p, four, spec, nf, s);
  if (shape == SHT_SHAPE_STAGED && p.P2p == 128 && nf >= kHalvesMinFields)
    return fwd_halves_launch<2, 2>(p, four, spec, nf, s);
  if (shape == SHT_SHAPE_STAGED)
    return p.P2p >= 128 ? fwd_staged_launch<1, 8>(p, four, spec, nf, s) : fwd_staged_launch<1, 4>(p, four, spec, nf, s);
  const int rt = (ceil_half(p.N + 1) + 15) / 16;
  if (shape == SHT_SHAPE_TILE || shape == SHT_SHAPE_TILE_SPLIT) {
    // TILE_SPLIT: the given wavenumber classes, else the built-in ones
    const KSplit k = shape == SHT_SHAPE_TILE ? KSplit{0, 0} : (cls.m2 >= 0 ? cls : default_classes(p, 1));
    switch (rt) {
      case 1: return fwd_tile_launch<1, 2>(p, k, four, spec, nf, s);
      case 2: return fwd_tile_launch<2, 2>(p, k, four, spec, nf, s);
      case 3: return fwd_tile_launch<3, 2>(p, k, four, spec, nf, s);
      case 4: return fwd_tile_launch<4, 2>(p, k, four, spec, nf, s);
      case 5: return fwd_tile_launch<5, 2>(p, k, four, spec, nf, s);
      case 6: return fwd_tile_launch<6, 2>(p, k, four, spec, nf, s);
      case 7: return fwd_tile_launch<7, 2>(p, k, four, spec, nf, s);
      case 8: return fwd_tile_launch<8, 2>(p, k, four, spec, nf, s);
    }
    return fail(SHT_ERR_UNSUPPORTED, "nmax above 255 is not instantiated");
  }
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

int launch_leg_fwd_f64_table(const Plan &p, const double *table, const double *four, double *spec,
                             bool swap, bool accumulate, int nf, hipStream_t s) {
  switch ((ceil_half(p.N + 1) + 15) / 16) {
    case 1: return fwd_launch<1>(p, four, spec, nf, s, table, swap, accumulate);
    case 2: return fwd_launch<2>(p, four, spec, nf, s, table, swap, accumulate);
    case 3: return fwd_launch<3>(p, four, spec, nf, s, table, swap, accumulate);
    case 4: return fwd_launch<4>(p, four, spec, nf, s, table, swap, accumulate);
    case 5: return fwd_launch<5>(p, four, spec, nf, s, table, swap, accumulate);
    case 6: return fwd_launch<6>(p, four, spec, nf, s, table, swap, accumulate);
    case 7: return fwd_launch<7>(p, four, spec, nf, s, table, swap, accumulate);
    case 8: return fwd_launch<8>(p, four, spec, nf, s, table, swap, accumulate);
  }
  return fail(SHT_ERR_UNSUPPORTED, "nmax above 255 is not instantiated");
}

int launch_leg_inv_f64(const Plan &p, const double *spec, double *four, double *four2,
                       bool value_table, bool deriv_table, int nf, hipStream_t s) {
  return launch_leg_inv_f64_shape(p, resolve_shape(p, nf, 0), spec, four, four2, value_table,
                                  deriv_table, nf, s, resolve_classes(p, nf, 0));
}

int launch_leg_fwd_f64(const Plan &p, const double *four, double *spec, int nf, hipStream_t s) {
  return launch_leg_fwd_f64_shape(p, resolve_shape(p, nf, 1), four, spec, nf, s, resolve_classes(p, nf, 1));
}

}  // namespace sht

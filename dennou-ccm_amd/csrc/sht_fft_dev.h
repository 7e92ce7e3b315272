// sht_fft_dev.h -- device code of the ring FFT stage, shared by the FFT
// kernels (sht_fft.hip) and the fused Legendre + FFT kernels
// (sht_legendre_mixed.hip): the LDS Stockham transform and the inverse's
// per-tile body.
//
// No implicit fused multiply-adds in this code: the compiler picks which
// product of a*b - c*d to fuse per instantiation, so the direct and
// LDS-staged forward kernels (the same arithmetic) would round differently
// and a field's spectrum would depend on the batch size.  The complex
// multiply fuses explicitly instead (same operation count, fixed rounding).
// Each function with floating-point arithmetic turns contraction off in its
// own body, so the setting holds whichever file includes this header and
// does not leak into it.
#pragma once

#include "sht_internal.h"

namespace sht {
namespace {

// A block transforms C complex values (G = C / L ring pairs) with C / 8
// threads: kFftBig (256 threads) or kFftSmall (64 threads, one wave), see
// fft_block.
template <int C>
constexpr int kThreadsOf = C / 8;
// LDS holds one pad element per 16 complex values -- the radix-16 pass writes
// 16 consecutive outputs per thread (a 256-byte lane stride), which without the
// pad lands every lane of a ds_write on the same banks -- and one more per
// transform, because the Fourier-side loops run the G transforms of a block in
// consecutive lanes (element m of transform g), a stride of L(1 + 1/16)
// elements that is otherwise a multiple of the bank count.
template <int C>
constexpr int kPaddedOf = C + C / 16 + C / 16;
// the big block, as the fused inverse uses it
constexpr int kThreads = kThreadsOf<kFftBig>;
constexpr int kLdsComplex = kFftBig;
constexpr int kLdsPadded = kPaddedOf<kFftBig>;
template <int L>
__device__ __forceinline__ int px(int i) { return i + (i >> 4) + i / L; }

// Regular stores: the Legendre step that reads the Fourier data next finds
// them in the Infinity Cache (streaming them made it slower at every size).
__device__ __forceinline__ void store_pair(double *o, double2 v) { *reinterpret_cast<double2 *>(o) = v; }
__device__ __forceinline__ void store_pair(float *o, double2 v) {
  *reinterpret_cast<float2 *>(o) = make_float2((float)v.x, (float)v.y);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(a.x - b.x, a.y - b.y);
}
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
#pragma clang fp contract(off)
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

// FFT tiles over the Fourier intermediate's field groups: a block of G fields
// covers 1 / KS of a group of layout_group(G) fields (KS = 1 up to 256
// longitudes).  Tile t is (pair p, FFT group fgr) with the KS blocks of one
// group consecutive, so they run on one XCD at about the same time and the
// parts of each 64-byte run meet in its L2.
template <int G>
struct FourTile {
  static constexpr int GL = layout_group(G), KS = GL / G;
  int p, fgr;
  __device__ __forceinline__ FourTile(int tile, int P2) {
    const int sub = tile % KS, rest = tile / KS;
    p = rest % P2;
    fgr = (rest / P2) * KS + sub;
  }
  // offset of (m, parity 0, the block's first field) for nf fields; parity
  // 1 follows 2 GL values later
  __device__ __forceinline__ size_t off(int m, int M1, int nf) const {
    const int nfg = (nf + GL - 1) / GL;
    return ((((size_t)p * nfg + fgr / KS) * M1 + m) * 2) * (2 * GL) + (fgr % KS) * (2 * G);
  }
};
// FFT tiles of a call with nf fields
template <int G>
__host__ __device__ inline int four_tiles(int P2, int nf) {
  constexpr int GL = layout_group(G);
  return P2 * ((nf + GL - 1) / GL) * (GL / G);
}

// One Stockham pass of radix R over G transforms of length L, Ns the product
// of the radices already applied.  Butterfly j of transform g reads
// x[j + r L/R] (in(g, i) returns element i of transform g), applies
// w^(r k L/(Ns R)) (k = j mod Ns; the powers are built by repeated
// multiplication from one LDS twiddle), runs a register DFT and writes
// y[(j/Ns) Ns R + k + r Ns] through out(g, i, v).  A barrier separates the
// reads from the writes (the passes run in place in buf) and follows them.
struct NoHook {
  __device__ void operator()() const {}
};

template <int L, int R, int Ns, int SIGN, int C, class In, class Out, class Hook = NoHook>
__device__ __forceinline__ void stockham_pass(const double2 *tw, In in, Out out, int tid, Hook hook = {}) {
#pragma clang fp contract(off)
  constexpr int G = C / L;
  constexpr int kThreads = kThreadsOf<C>;
  constexpr int NB = G * (L / R);                       // butterflies in the block
  constexpr int QB = (NB + kThreads - 1) / kThreads;    // per thread
  double2 v[QB][R];
  // all of the pass's reads first, then hook() (work whose own loads then
  // share one wait with them), then the butterflies
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int idx = tid + q * kThreads;
    if (NB % kThreads != 0 && idx >= NB) continue;
    const int g = idx / (L / R), j = idx % (L / R);
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = in(g, j + r * (L / R));
  }
  hook();
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int idx = tid + q * kThreads;
    if (NB % kThreads != 0 && idx >= NB) continue;
    const int j = idx % (L / R), k = j % Ns;
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

// Complex FFT of G = C / L transforms, exponent sign SIGN (-1
// forward, +1 inverse), unnormalised; tw (in LDS) holds e^{+2 pi i t / L}.
// The first pass reads through first(g, i) and the last writes through
// last(g, i, v), so a caller can feed the first butterflies straight from
// global memory, skip known zeros, or store the results straight out, and
// save the LDS round trips; the passes between run in place in buf.
// Radix plan: 16 = 16, 32 = 8*4, 64 = 8*8, 128 = 16*8, 256 = 16*16,
// 512 = 8*8*8, 1024 = 16*16*4, 2048 = 16*16*8.  Ends with a barrier.
// hook() runs in the first pass once its reads are issued (see stockham_pass).
template <int L, int SIGN, int C, class First, class Last, class Hook = NoHook>
__device__ __forceinline__ void fft_lds(double2 *buf, const double2 *tw, First first, Last last, int tid,
                                        Hook hook = {}) {
#pragma clang fp contract(off)
  auto rd = [&](int g, int i) { return buf[px<L>(g * L + i)]; };
  auto wr = [&](int g, int i, double2 v) { buf[px<L>(g * L + i)] = v; };
  if constexpr (L == 16) {
    stockham_pass<L, 16, 1, SIGN, C>(tw, first, last, tid, hook);
  } else if constexpr (L == 32) {
    stockham_pass<L, 8, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 4, 8, SIGN, C>(tw, rd, last, tid);
  } else if constexpr (L == 64) {
    stockham_pass<L, 8, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 8, 8, SIGN, C>(tw, rd, last, tid);
  } else if constexpr (L == 128) {
    stockham_pass<L, 16, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 8, 16, SIGN, C>(tw, rd, last, tid);
  } else if constexpr (L == 256) {
    stockham_pass<L, 16, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 16, 16, SIGN, C>(tw, rd, last, tid);
  } else if constexpr (L == 512) {
    stockham_pass<L, 8, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 8, 8, SIGN, C>(tw, rd, wr, tid);
    stockham_pass<L, 8, 64, SIGN, C>(tw, rd, last, tid);
  } else if constexpr (L == 1024) {
    stockham_pass<L, 16, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 16, 16, SIGN, C>(tw, rd, wr, tid);
    stockham_pass<L, 4, 256, SIGN, C>(tw, rd, last, tid);
  } else {
    static_assert(L == 2048, "unsupported ring length");
    stockham_pass<L, 16, 1, SIGN, C>(tw, first, wr, tid, hook);
    stockham_pass<L, 16, 16, SIGN, C>(tw, rd, wr, tid);
    stockham_pass<L, 8, 256, SIGN, C>(tw, rd, last, tid);
  }
}


// Loads of the Fourier intermediate, one (re, im) pair in its stored type
// (float2 or double2): COH = device-coherent (agent-scope relaxed atomic
// loads, 8 bytes each), for data another workgroup of the same kernel wrote;
// plain loads otherwise.  The pair is 8- or 16-byte aligned.
template <bool COH>
__device__ __forceinline__ float2 load_four_raw(const float *e) {
  if constexpr (COH) {
    const unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(e),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(float2, u);
  } else {
    return *reinterpret_cast<const float2 *>(e);
  }
}
template <bool COH>
__device__ __forceinline__ double2 load_four_raw(const double *e) {
  if constexpr (COH) {
    const unsigned long long a = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(e),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long b = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(e) + 1,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_double2(__builtin_bit_cast(double, a), __builtin_bit_cast(double, b));
  } else {
    return *reinterpret_cast<const double2 *>(e);
  }
}
__device__ __forceinline__ double2 widen(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 widen(double2 v) { return v; }

// One inverse tile: four (even/odd Legendre sums) -> grid for latitude pair
// p and field group fgr (tile = fgr * P2 + p).  A thread's Fourier loads (at
// most 4 (m, field) pairs: M1 * G <= L/2 * G = C/2) and, with tw not NULL,
// the twiddles into twl are issued before any LDS traffic (with tw NULL the
// caller has filled twl and synchronised).  Only the coefficients m < M1 and
// L - m are placed in LDS: the first FFT pass takes the band between them as
// zeros without reading it, and the last pass stores the rings straight to
// the grid (non-temporal).  mode: 0 value, 1 GradLon (x i m / cos lat),
// 2 GradLat (x 1 / cos lat).  Ends with a barrier.
template <int L, typename TF, bool COH, int C = kFftBig>
__device__ __forceinline__ void fft_inv_tile(double2 *buf, double2 *twl, const double2 *__restrict__ tw,
                                             const TF *__restrict__ src, double *__restrict__ dst, int mode,
                                             const double *__restrict__ coslat, int P2, int M1, int nf, int jm,
                                             int tile, int tid) {
#pragma clang fp contract(off)
  constexpr int G = C / L;
  constexpr int kThreads = kThreadsOf<C>;
  constexpr int NM = (C / 2) / kThreads;  // max (m, field) pairs per thread
  const FourTile<G> ft(tile, P2);
  const int p = ft.p, f0 = ft.fgr * G;
  const int jn = P2 + p, js = P2 - 1 - p;
  const double rc = (mode == 0) ? 1.0 : 1.0 / coslat[jn];
  // every load issued unconditionally (pairs past M1 or nf read the tile's
  // first pair and are not used), converted to FP64 only after the
  // twiddles: converting inside a branch per pair made the FP32 form wait
  // for each pair before issuing the next
  using Raw = decltype(load_four_raw<COH>(src));
  Raw Er[NM], Or[NM];
#pragma unroll
  for (int q = 0; q < NM; ++q) {
    const int idx = tid + q * kThreads;
    const int m = idx / G, g = idx % G, f = f0 + g;
    const bool ok = m < M1 && f < nf;
    const TF *e = src + ft.off(ok ? m : 0, M1, nf) + (ok ? 2 * g : 0);
    Er[q] = load_four_raw<COH>(e);
    Or[q] = load_four_raw<COH>(e + 2 * FourTile<G>::GL);
  }
  if (tw)
    for (int q = tid; q < L; q += kThreads) twl[q] = tw[q];
#pragma unroll
  for (int q = 0; q < NM; ++q) {
    const int idx = tid + q * kThreads;
    const int m = idx / G, g = idx % G, f = f0 + g;
    if (m >= M1 || f >= nf) continue;
    const double2 Eq = widen(Er[q]), Oq = widen(Or[q]);
    double2 xn = cadd(Eq, Oq), xs = csub(Eq, Oq);
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
  fft_lds<L, +1, C>(buf, twl, spectrum, rings, tid);
}

}  // namespace
}  // namespace sht


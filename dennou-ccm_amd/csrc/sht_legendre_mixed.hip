// sht_legendre_mixed.hip -- the Gauss-Legendre step on the 16-bit matrix
// cores (v_mfma_f32_16x16x32_bf16 / _f16) with split operands.
//
// Each FP64 operand x is carried as a sum of 16-bit pieces:
//   SPLIT_BF16    x ~ hi + lo               (bf16, |x - hi - lo| <= 2^-18 |x|)
//   SPLIT_BF16X6  x ~ hi + mid + lo         (bf16, <= 2^-27 |x|)
//   SCALED_FP16   x 2^e ~ hi + lo 2^-11     (fp16, <= 2^-22 |x 2^e|)
// (bounds for the tables; the data operands are rounded to FP32 before the
// split, adding up to 2^-25 |x|), and a product A.B is formed from the piece
// products whose magnitude is not
// below the split residual: hi.hi + hi.lo + lo.hi (three products) for the
// two-piece forms, hi.hi + hi.mid + mid.hi + hi.lo + lo.hi + mid.mid (six) for
// BF16X6.  Each v_mfma covers a 32-wide slice of the GEMM's K dimension and
// sums in FP32; the slices are combined (and FP16 scalings undone) in FP64.
//
// FP16's exponent range is why that form is scaled: the table block of each
// (m, parity) carries one power of two (plan time), and the data of each
// (column, K slice) another, chosen in the kernel from the slice's largest
// magnitude when the workgroup splits its data tile into LDS.  The lo piece is
// stored 2^11 larger so it stays a normal FP16 number; its products are kept
// in their own accumulator.
//
// Table layouts (16-bit, K contiguous so one lane loads its 8 K values with a
// single 16-byte load, matching the 16x16x32 operand layout: lane l holds row
// l%16, K = 8*(l/16) .. +7):
//   Linv per m: [par][piece][P2p pair rows][KP32 degrees]
//   Lfwd per m: [par][piece][KR degree rows][P2p32 pairs]   (w/2 folded in)
// The Fourier intermediate of mixed-precision plans is FP32.
#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "sht_fft_dev.h"
#include "sht_internal.h"

namespace sht {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

template <int MODE>
struct Form;
template <>
struct Form<SHT_PREC_SPLIT_BF16> {
  static constexpr int NP = 2;
  static constexpr bool kScaled = false;
};
template <>
struct Form<SHT_PREC_SPLIT_BF16X6> {
  static constexpr int NP = 3;
  static constexpr bool kScaled = false;
};
template <>
struct Form<SHT_PREC_SCALED_FP16> {
  static constexpr int NP = 2;
  static constexpr bool kScaled = true;
};

__device__ __forceinline__ uint16_t bf16_bits(double x) {
  __bf16 h = (__bf16)(float)x;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ double bf16_val(uint16_t b) {
  return (double)(float)__builtin_bit_cast(__bf16, b);
}
__device__ __forceinline__ uint16_t f16_bits(double x) {
  _Float16 h = (_Float16)(float)x;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ double f16_val(uint16_t b) {
  return (double)(float)__builtin_bit_cast(_Float16, b);
}

// Splits x (already scaled for FP16) into NP 16-bit pieces.
template <int MODE>
__device__ __forceinline__ void split(double x, uint16_t (&pc)[3]) {
  if constexpr (MODE == SHT_PREC_SCALED_FP16) {
    pc[0] = f16_bits(x);
    pc[1] = f16_bits((x - f16_val(pc[0])) * 2048.0);
    pc[2] = 0;
  } else {
    pc[0] = bf16_bits(x);
    double r = x - bf16_val(pc[0]);
    pc[1] = bf16_bits(r);
    pc[2] = bf16_bits(r - bf16_val(pc[1]));
  }
}

// The same pieces for an FP32 value, in FP32 arithmetic: x has 24
// significant bits, so every residual is exact in FP32 -- the pieces equal
// those the FP64 split above gives for (double)x -- and hi + mid + lo = x
// exactly for BF16X6.  Both data operands are split this way: the forward's
// Fourier data are FP32, and the inverse rounds each FP64 coefficient to
// FP32 first, which keeps the bound of the three-piece form (2^-25 |x| from
// the rounding, none from the split) at a fraction of the FP64 split's cost.
// The FP64 split above serves the tables.
// Round to nearest even on the FP32 bits (finite x).  Written with integer
// operations on purpose: given (__bf16)(float)d for an FP64 d, the compiler
// fuses the two roundings into one double -> bf16 rounding, which differs
// from rounding the FP32 value when that value is a bf16 tie, and it did so
// in one kernel shape but not in another.
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_valf(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
template <int MODE>
__device__ __forceinline__ void split(float x, uint16_t (&pc)[3]) {
  if constexpr (MODE == SHT_PREC_SCALED_FP16) {
    const _Float16 h = (_Float16)x;
    pc[0] = __builtin_bit_cast(uint16_t, h);
    pc[1] = __builtin_bit_cast(uint16_t, (_Float16)((x - (float)h) * 2048.0f));
    pc[2] = 0;
  } else {
    pc[0] = bf16_bits(x);
    const float r = x - bf16_valf(pc[0]);
    pc[1] = bf16_bits(r);
    pc[2] = bf16_bits(r - bf16_valf(pc[1]));
  }
}

template <int MODE>
__device__ __forceinline__ f4 mfma16(const uint4 &a, const uint4 &b, f4 c) {
  if constexpr (MODE == SHT_PREC_SCALED_FP16)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Products of one K slice, in two independent FP32 chains (the small cross
// terms, then the large ones) so that consecutive MFMAs do not wait on each
// other; the chains meet in FP64.
template <int MODE>
struct SliceChains {
  f4 small, big;
};

template <int MODE>
__device__ __forceinline__ SliceChains<MODE> slice_products(const uint4 (&pa)[3], const uint4 (&pb)[3]) {
  const f4 z = {0.f, 0.f, 0.f, 0.f};
  SliceChains<MODE> c;
  if constexpr (MODE == SHT_PREC_SCALED_FP16) {
    // small: the lo pieces (stored 2^11 larger), big: hi.hi
    c.small = mfma16<MODE>(pa[0], pb[1], z);
    c.big = mfma16<MODE>(pa[0], pb[0], z);
    c.small = mfma16<MODE>(pa[1], pb[0], c.small);
  } else if constexpr (MODE == SHT_PREC_SPLIT_BF16X6) {
    c.small = mfma16<MODE>(pa[1], pb[1], z);
    c.big = mfma16<MODE>(pa[0], pb[1], z);
    c.small = mfma16<MODE>(pa[0], pb[2], c.small);
    c.big = mfma16<MODE>(pa[1], pb[0], c.big);
    c.small = mfma16<MODE>(pa[2], pb[0], c.small);
    c.big = mfma16<MODE>(pa[0], pb[0], c.big);
  } else {
    c.small = mfma16<MODE>(pa[0], pb[1], z);
    c.big = mfma16<MODE>(pa[0], pb[0], z);
    c.small = mfma16<MODE>(pa[1], pb[0], c.small);
  }
  return c;
}

// Adds one slice's contribution, unscaled, to acc in FP64.
template <int MODE>
__device__ __forceinline__ void slice_add(const SliceChains<MODE> &c, double scale, d4 &acc) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if constexpr (MODE == SHT_PREC_SCALED_FP16)
      acc[i] += ((double)c.big[i] + (double)c.small[i] * (1.0 / 2048.0)) * scale;
    else
      acc[i] += (double)c.small[i] + (double)c.big[i];
  }
}

template <int MODE>
__device__ __forceinline__ void slice_mac(const uint4 (&pa)[3], const uint4 (&pb)[3], double scale,
                                          d4 &acc) {
  slice_add<MODE>(slice_products<MODE>(pa, pb), scale, acc);
}

// Two slices (the two parities of one row tile), products issued together so
// four independent chains alternate on the matrix core.
template <int MODE>
__device__ __forceinline__ void slice_mac2(const uint4 (&pa0)[3], const uint4 (&pb0)[3], double sc0,
                                           d4 &acc0, const uint4 (&pa1)[3], const uint4 (&pb1)[3],
                                           double sc1, d4 &acc1) {
  const SliceChains<MODE> c0 = slice_products<MODE>(pa0, pb0);
  const SliceChains<MODE> c1 = slice_products<MODE>(pa1, pb1);
  slice_add<MODE>(c0, sc0, acc0);
  slice_add<MODE>(c1, sc1, acc1);
}

// Data tile of one workgroup: for each parity and K slice, the 16 columns x 32
// K values split into pieces, laid out [group][piece][col][kRowPitch] (16-bit)
// so a lane's 8 K values are one 16-byte LDS read; ebits[group][col] holds the
// FP16 scaling exponent.  group = par * maxsl + slice.  A column row is 32
// values padded to 40 (80 bytes): eight lanes reading the same K octet of
// consecutive columns then hit eight different 16-byte bank groups.
constexpr int kRowPitch = 40;
// Every wave converts whole groups (wave w of NW takes groups w, w+NW, ...): lane l
// reads the 8 values K = 8 (l/16) .. +7 of column l % 16 (load(g, k, c) returns
// the FP64 value or 0 outside the tile), the FP16 exponent comes from the
// maximum over the group's 32 values (a 4-lane shuffle reduction), and the
// lane writes each of its pieces as one 16-byte LDS store -- already the
// matrix-core B layout.  A wave issues the loads of NIT groups before it
// converts any, so one memory latency covers them (groups past ngroups load
// group 0 again and are not stored).  Group g lands in LDS slot slot(g), so a
// caller can stage only the groups that hold data.
template <int MODE, int NIT, int NW, class Slot, class Load>
__device__ __forceinline__ void stage_data(uint16_t *bt, int *ebits, int ngroups, Slot slot, Load load,
                                           int tid) {
  constexpr int NP = Form<MODE>::NP;
  const int lane = tid & 63, c = lane & 15, k8 = lane >> 4;
  using T = decltype(load(0, 0, 0));  // double (spectra) or float (Fourier data)
  for (int g0 = __builtin_amdgcn_readfirstlane(tid >> 6); g0 < ngroups; g0 += NW * NIT) {
    T v[NIT][8];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int g = g0 + NW * it < ngroups ? g0 + NW * it : 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) v[it][i] = load(g, k8 * 8 + i, c);
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      if (g0 + NW * it >= ngroups) break;
      const int g = slot(g0 + NW * it);
      int e = 0;
      if constexpr (Form<MODE>::kScaled) {
        T mx = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) mx = fmax(mx, fabs(v[it][i]));
        mx = fmax(mx, __shfl_xor(mx, 16));
        mx = fmax(mx, __shfl_xor(mx, 32));
        e = mx > 0 ? 13 - ilogb(mx) : 0;
        if (k8 == 0) ebits[g * 16 + c] = e;
      }
      uint32_t w[3][4];
#pragma unroll
      for (int i = 0; i < 8; i += 2) {
        uint16_t p0[3], p1[3];
        split<MODE>(Form<MODE>::kScaled ? ldexp(v[it][i], e) : v[it][i], p0);
        split<MODE>(Form<MODE>::kScaled ? ldexp(v[it][i + 1], e) : v[it][i + 1], p1);
#pragma unroll
        for (int q = 0; q < 3; ++q) w[q][i >> 1] = p0[q] | ((uint32_t)p1[q] << 16);
      }
#pragma unroll
      for (int q = 0; q < NP; ++q)
        *reinterpret_cast<uint4 *>(bt + (((size_t)g * NP + q) * 16 + c) * kRowPitch + k8 * 8) =
            make_uint4(w[q][0], w[q][1], w[q][2], w[q][3]);
    }
  }
}

// Groups a wave of NW stages per memory round trip: its share of NGROUPS, at
// most two, so the staged values (16 registers per group) stay few.
template <int NGROUPS, int NW>
constexpr int stage_iters() {
  return (NGROUPS + NW - 1) / NW < 2 ? (NGROUPS + NW - 1) / NW : 2;
}

template <int MODE>
__device__ __forceinline__ void load_pb(const uint16_t *bt, int g, int kq, int r16, uint4 (&pb)[3]) {
  constexpr int NP = Form<MODE>::NP;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    pb[q] = (q < NP) ? *reinterpret_cast<const uint4 *>(bt + (((size_t)g * NP + q) * 16 + r16) * kRowPitch + kq * 8)
                     : make_uint4(0, 0, 0, 0);
}

// Row pitch (floats) of the inverse's output transpose: 32 values of a pair
// row plus 4, so the 4 lane groups of one ds_write hit different banks.
constexpr int kOutPitch = 36;

// NSL (template) = the kernel's K slices of 32 at m = 0: KP32/32 <= 4 for the
// inverse (N <= 255), P2p32/32 <= 8 for the forward (jm <= 512).  Registers and
// LDS are sized by it, so small truncations keep a high occupancy.

// Inverse: spec (FP64) -> four (FP32 even/odd Legendre sums).
// Workgroup = one (m, CT 16-column tiles) of NW waves.  Each wave first issues
// the table fragments of its first 16-pair row tile, then all threads split
// the spectral tiles into LDS; wave w then takes row tiles w, w+NW, ... and runs
// each against every column tile with the same table fragments.  CT = 2
// halves the workgroups, and with them the waves that wait on memory at
// once, and the table fragments fetched per column.  Results leave through a
// per-wave LDS transpose as 16-byte stores of 4 consecutive columns (the
// Fourier layout keeps a field group's 2G columns contiguous).
// One column tile per workgroup with at most two K slices (N <= 127) asks
// the compiler for 5 waves per SIMD, so a T85 batch of 97 fields (1144
// workgroups) is resident at once (measured at T85: neutral at 97 fields, 6 %
// faster at 388; the six-product form spills two registers for it).  With
// more slices the hint would spill heavily.
// Store of the FP32 Fourier intermediate: COH = device-coherent (agent-scope
// relaxed atomic stores, 8 bytes each), for data another workgroup of the
// same kernel reads; otherwise put_four.
template <bool NT, bool COH>
__device__ __forceinline__ void put_four4(float *dst, f4 v) {
  if constexpr (COH) {
    auto *u = reinterpret_cast<unsigned long long *>(dst);
    const unsigned long long lo = (unsigned long long)__float_as_uint(v[1]) << 32 | __float_as_uint(v[0]);
    const unsigned long long hi = (unsigned long long)__float_as_uint(v[3]) << 32 | __float_as_uint(v[2]);
    __hip_atomic_store(u, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(u + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    put_four<NT>(reinterpret_cast<f4 *>(dst), v);
  }
}
template <bool NT, bool COH>
__device__ __forceinline__ void put_four1(float *dst, float v) {
  if constexpr (COH)
    __hip_atomic_store(reinterpret_cast<unsigned *>(dst), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  else
    put_four<NT>(dst, v);
}

// One row tile rt (16 latitude pairs, its table fragments pa) against one
// staged column tile jt (groups par * NSL + s at bt, FP16 exponents at
// ebits): the nsl K slices' products summed in FP64 per parity, then the FP32
// sums to the Fourier intermediate.  ob is the wave's LDS patch for the
// output transpose.
template <int MODE, int NSL, bool NT, bool COH>
__device__ __forceinline__ void inv_tile(const uint16_t *bt, const int *ebits, float *ob,
                                         const uint4 (&pa)[2][NSL][3], double ea0, double ea1, int nsl,
                                         float *__restrict__ four, FourGeom geo, int P2, int ncol, int m,
                                         int rt, int jt, int lane) {
  const int r16 = lane & 15, kq = lane >> 4, col = jt * 16 + r16;
  d4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
  for (int s = 0; s < NSL; ++s) {
    if (s >= nsl) break;
    uint4 pb0[3], pb1[3];
    load_pb<MODE>(bt, s, kq, r16, pb0);
    load_pb<MODE>(bt, NSL + s, kq, r16, pb1);
    const double sc0 = Form<MODE>::kScaled ? ea0 * ldexp(1.0, -ebits[s * 16 + r16]) : 1.0;
    const double sc1 = Form<MODE>::kScaled ? ea1 * ldexp(1.0, -ebits[(NSL + s) * 16 + r16]) : 1.0;
    slice_mac2<MODE>(pa[0][s], pb0, sc0, acc[0], pa[1][s], pb1, sc1, acc[1]);
  }
  if (geo.lg >= 2) {
    // 4 consecutive columns per lane: through the wave's LDS patch
    // [16 pairs][2 parities][16 columns] (row pitch kOutPitch); the
    // padding columns of the last field group are written too (zeros)
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int i = 0; i < 4; ++i) ob[(kq * 4 + i) * kOutPitch + par * 16 + r16] = (float)acc[par][i];
    const int cend = geo.nfg << geo.lg;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = lane + 64 * h, pl = q >> 3, par = (q >> 2) & 1, c4 = (q & 3) * 4;
      const int p = rt * 16 + pl, cc = jt * 16 + c4;
      const f4 v = *reinterpret_cast<const f4 *>(ob + pl * kOutPitch + par * 16 + c4);
      if (p < P2 && cc < cend) put_four4<NT, COH>(four + geo.at(p, m, par, cc), v);
    }
  } else if (col < ncol) {
    // field groups of fewer than 4 columns: none since groups hold at least
    // 8 fields (layout_group), but without this path the compiler spills 94
    // registers of the T85 six-product inverse (2 with it)
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int p = rt * 16 + kq * 4 + i;
        if (p < P2) put_four1<NT, COH>(four + geo.at(p, m, par, col), (float)acc[par][i]);
      }
  }
}

// LDS of one inverse workgroup item (bytes, 16-byte aligned parts): the split
// data tiles, their FP16 exponents (scaled form only), the per-wave output
// transpose.
template <int MODE, int NSL, int CT, int NW>
struct InvItemLds {
  static constexpr int NP = Form<MODE>::NP, NG = 2 * NSL;
  static constexpr size_t bt = sizeof(uint16_t) * CT * NG * NP * 16 * kRowPitch;
  static constexpr size_t ebits = Form<MODE>::kScaled ? (sizeof(int) * CT * NG * 16 + 15) / 16 * 16 : 0;
  static constexpr size_t obuf = sizeof(float) * NW * 16 * kOutPitch;
  static constexpr size_t bytes = bt + ebits + obuf;
};

// Inverse: spec (FP64) -> four (FP32 even/odd Legendre sums).
// Workgroup = one (m, CT 16-column tiles) of NW waves.  Each wave first issues
// the table fragments of its first 16-pair row tile, then all threads split
// the spectral tiles into LDS; wave w then takes row tiles w, w+NW, ... and runs
// each against every column tile with the same table fragments.  CT = 2
// halves the workgroups, and with them the waves that wait on memory at
// once, and the table fragments fetched per column.  Results leave through a
// per-wave LDS transpose as 16-byte stores of 4 consecutive columns (the
// Fourier layout keeps a field group's 2G columns contiguous).  This is the
// work of one (m, column tile t) item, with its LDS at lds; COH stores the
// intermediate device-coherently (the fused kernel).
template <int MODE, int NSL, int CT, bool NT, int NW, bool COH>
__device__ __forceinline__ void leg_inv_mixed_item(
    unsigned char *lds, const double *__restrict__ spec, float *__restrict__ four,
    const uint16_t *__restrict__ L, const int *__restrict__ loff, const int *__restrict__ lexp, int N,
    int P2, int P2p, int ncol, int nct, FourGeom geo, int m, int t, int tid) {
  constexpr int NP = Form<MODE>::NP, NG = 2 * NSL;  // LDS groups per column tile
  using Lds = InvItemLds<MODE, NSL, CT, NW>;
  uint16_t *bt = reinterpret_cast<uint16_t *>(lds);
  int *ebits = reinterpret_cast<int *>(lds + Lds::bt);
  float *obuf = reinterpret_cast<float *>(lds + Lds::bt + Lds::ebits);
  const int ct0 = t * CT, nc = nct - ct0 < CT ? nct - ct0 : CT;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP32 = (ke + 31) & ~31, nsl = KP32 >> 5;
  const double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol + ct0 * 16;
  const uint16_t *__restrict__ Lm = L + loff[m];
  const int rtiles = P2p >> 4;

  auto a_frag = [&](int rt, int par, int s, int q) {
    return *reinterpret_cast<const uint4 *>(
        Lm + ((size_t)(par * NP + q) * P2p + rt * 16 + r16) * KP32 + s * 32 + kq * 8);
  };
  // table fragments of the nsl slices in use (the others stay zero, unused)
  uint4 pa[2][NSL][3];
  auto load_a = [&](int rt) {
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int s = 0; s < NSL; ++s)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          pa[par][s][q] = (q < NP && s < nsl) ? a_frag(rt, par, s, q) : make_uint4(0, 0, 0, 0);
  };
  if (wave < rtiles) load_a(wave);
  // only the nsl slices that hold degrees: group g = (j, parity, s) of
  // column tile j goes to slot j NG + parity NSL + s
  const int ng = 2 * nsl;
  auto slot = [&](int g) {
    const int j = g / ng, r = g - j * ng, par = r >= nsl;
    return j * NG + par * NSL + r - par * nsl;
  };
  stage_data<MODE, stage_iters<CT * NG, NW>(), NW>(bt, ebits, nc * ng, slot, [&](int g, int k, int c) {
    const int j = g / ng, r = g - j * ng, par = r >= nsl, s = r - par * nsl;
    const int kk = s * 32 + k, cc = j * 16 + c;
    const bool v = kk < (par ? ko : ke) && ct0 * 16 + cc < ncol;
    const double x = sm[(size_t)(v ? 2 * kk + par : 0) * ncol + (v ? cc : 0)];
    return v ? (float)x : 0.0f;  // split as FP32: see split(float)
  }, tid);
  __syncthreads();
  for (int rt = wave; rt < rtiles; rt += NW) {
    if (rt != wave) load_a(rt);
    const double ea0 = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m]) : 1.0;
    const double ea1 = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m + 1]) : 1.0;
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      if (j >= nc) break;
      inv_tile<MODE, NSL, NT, COH>(bt + (size_t)j * NG * NP * 16 * kRowPitch, ebits + j * NG * 16,
                                   obuf + wave * (16 * kOutPitch), pa, ea0, ea1, nsl, four, geo, P2, ncol,
                                   m, rt, ct0 + j, lane);
    }
  }
}

// One column tile per workgroup with at most two K slices (N <= 127) asks
// the compiler for 5 waves per SIMD, so a T85 batch of 97 fields (1144
// workgroups) is resident at once (measured at T85: neutral at 97 fields, 6 %
// faster at 388; the six-product form spills two registers for it).  With
// more slices the hint would spill heavily.
template <int MODE, int NSL, int CT, bool NT, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu((CT == 1 && NSL <= 2) ? 5 : 1))) void
leg_inv_mixed_kernel(
    const double *__restrict__ spec, float *__restrict__ four, const uint16_t *__restrict__ L,
    const int *__restrict__ loff, const int *__restrict__ lexp, const int *__restrict__ mlist,
    int mlen, int N, int P2, int P2p, int ncol, int nct, FourGeom geo) {
  __shared__ __align__(16) unsigned char lds[InvItemLds<MODE, NSL, CT, NW>::bytes];
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + CT - 1) / CT, m, t)) return;
  leg_inv_mixed_item<MODE, NSL, CT, NT, NW, false>(lds, spec, four, L, loff, lexp, N, P2, P2p, ncol, nct,
                                                   geo, m, t, threadIdx.x);
}

// ---- Fused inverse: Legendre step and FFT in one persistent launch.
//
// The inverse Legendre items (m, 16-column tile) of leg_inv_mixed_kernel and
// the FFT tiles (latitude pair, field group) of fft_inv_kernel run in one
// launch, so a call pays one kernel start and one ramp-up of the chip instead
// of two.  Each workgroup serves the queue of its XCD (blockIdx % 8): first
// the Legendre items of the wavenumbers dealt to that XCD (column tile by
// column tile, so the first column tiles complete first), then every eighth
// FFT tile in field-group order.  A Legendre item stores its part of the
// Fourier intermediate device-coherently (agent-scope atomic stores), waits
// for the stores to complete, and counts itself done on its column tile; an
// FFT tile starts when the column tiles of its field group count all N + 1
// wavenumbers, and reads the intermediate device-coherently.  A workgroup
// waiting on a count takes Legendre items no workgroup has taken yet, from
// any queue; Legendre items never wait, so every count completes whatever
// the number of workgroups the device runs at once.  A wait beyond one
// second (a broken device or counter state) gives up: the tile's grid rows
// are set to NaN and the sync area's error word counts it.  The last
// workgroup to finish resets the counters for the next call; calls on one
// plan must therefore not overlap (as for the plan's Fourier scratch).
struct FusedInv {
  const double *spec;
  float *four;
  double *grid;
  const uint16_t *L;
  const int *loff, *lexp, *mlist;
  const double2 *tw;
  const double *coslat;
  int *sync;
  int mlen, N, P2, P2p, ncol, nct, nf, jm;
  int need;             // completions a column tile waits for: N + 1
  uint64_t spin_ticks;  // bound of one wait (s_memrealtime ticks)
  FourGeom geo;
};

// Sync area: every counter in a 256-byte slot of its own (kSyncStride ints),
// so the polls of waiting workgroups and the queue atomics do not share
// cache lines: slots 0-7 Legendre queues, 8-15 FFT queues, 16 finished
// workgroups, 17 errors, kFusedSyncHead + c column tile c's count.
constexpr int kSyncL = 0, kSyncF = 8, kSyncDone = 16, kSyncErr = 17;
constexpr uint64_t kFusedSpinTicks = 100000000;  // s_memrealtime runs at 100 MHz: one second
// SHT_INJECT_FUSED_TIMEOUT: every wait of the plan's fused calls gives up
// after 10 us (the completion count asked for is one more than the items
// make), which exercises the timeout path: NaN rings and the error count
constexpr uint64_t kFusedInjectedSpinTicks = 1000;

__device__ __forceinline__ int sync_load(const int *a) {
  return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int sync_add(int *a, int v) {
  return __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Three waves per SIMD (three workgroups per CU; the LDS would hold four, the
// registers of the larger items not).
template <int L, int MODE, int NSL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void inv_fused_kernel(FusedInv a) {
  using Lds = InvItemLds<MODE, NSL, 1, 4>;
  constexpr size_t kFft = sizeof(double2) * kLdsPadded;
  __shared__ __align__(16) unsigned char lds[Lds::bytes > kFft ? Lds::bytes : kFft];
  __shared__ double2 twl[L];
  __shared__ int s_item, s_bad;
  constexpr int kDone = -1;
  const int tid = threadIdx.x, x = blockIdx.x & 7;
  int *const sy = a.sync;
  auto slot = [&](int k) { return sy + k * kSyncStride; };
  const int nL = a.mlen * a.nct;                // Legendre items per queue
  constexpr int GF = kLdsComplex / L;  // fields per FFT tile
  const int nF = four_tiles<GF>(a.P2, a.nf), nFx = (nF - x + 7) >> 3;
  const int need = a.need;
  for (int q = tid; q < L; q += 256) twl[q] = a.tw[q];
  // Next item, chosen by thread 0: >= 0 Legendre item (queue << 24 | index),
  // <= -2 FFT tile -2 - tile, kDone.  Thread 0 keeps its queue state.
  bool own_done = false, all_taken = false;
  int pend = -1;
  auto take_l = [&](int y) -> int {  // the next Legendre item of queue y, or -1
    for (;;) {
      const int i = sync_add(slot(kSyncL + y), 1);
      if (i >= nL) return -1;
      if (a.mlist[y * a.mlen + i % a.mlen] >= 0) return y << 24 | i;
    }
  };
  // a Legendre item of another queue that no workgroup has taken (one poll
  // of the eight queues, issued together), or -1
  auto steal = [&]() -> int {
    int v[8];
#pragma unroll
    for (int y = 0; y < 8; ++y) v[y] = sync_load(slot(kSyncL + y));
    bool any = false;
#pragma unroll
    for (int y = 0; y < 8; ++y) any = any || v[y] < nL;
    if (!any) {
      all_taken = true;
      return -1;
    }
    for (int k = 1; k <= 8; ++k) {
      const int y = (x + k) & 7;
      if (v[y] < nL) {
        const int it = take_l(y);
        if (it >= 0) return it;
      }
    }
    return -1;
  };
  auto next_item = [&](int &bad) -> int {
    bad = 0;
    if (pend < 0) {
      if (!own_done) {
        const int it = take_l(x);
        if (it >= 0) return it;
        own_done = true;
      }
      const int i = sync_add(slot(kSyncF + x), 1);
      if (i >= nFx) return kDone;
      pend = x + 8 * i;
    }
    const int f0 = FourTile<GF>(pend, a.P2).fgr * GF;  // the tile's columns 2 f0 .. 2 (f0 + GF) - 1
    const int c0 = (2 * f0) >> 4;
    const int c1 = min((2 * (f0 + GF) - 1) >> 4, a.nct - 1);
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (int poll = 0;; ++poll) {
      // at most 4 column tiles per field group (2G <= 64 columns)
      int cnt[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) cnt[k] = c0 + k <= c1 ? sync_load(slot(kFusedSyncHead + c0 + k)) : need;
      bool ready = cnt[0] >= need && cnt[1] >= need && cnt[2] >= need && cnt[3] >= need;
      if (!ready) {
        if (!all_taken && poll % 8 == 0) {
          const int it = steal();
          if (it >= 0) return it;
        }
        if (__builtin_amdgcn_s_memrealtime() - t0 > a.spin_ticks) {
          sync_add(slot(kSyncErr), 1);
          bad = 1;
          ready = true;
        }
      }
      if (ready) {
        const int f = pend;
        pend = -1;
        return -2 - f;
      }
      __builtin_amdgcn_s_sleep(16);
    }
  };
  for (;;) {
    if (tid == 0) {
      int bad;
      s_item = next_item(bad);
      s_bad = bad;
    }
    __syncthreads();
    const int item = s_item, bad = s_bad;
    __syncthreads();  // s_item is rewritten only after every wave has read it
    if (item == kDone) break;
    // the thread index, opaque to the compiler for this item: otherwise it
    // hoists both item bodies' index arithmetic out of the loop and keeps it
    // live throughout (twice the registers of either kernel alone)
    int ltid = tid;
    asm volatile("" : "+v"(ltid));
    if (item >= 0) {
      const int y = item >> 24, i = item & 0xffffff;
      const int t = i / a.mlen, m = a.mlist[y * a.mlen + i % a.mlen];
      leg_inv_mixed_item<MODE, NSL, 1, false, 4, true>(lds, a.spec, a.four, a.L, a.loff, a.lexp, a.N, a.P2,
                                                       a.P2p, a.ncol, a.nct, a.geo, m, t, ltid);
      // this thread's coherent stores have completed; after the barrier
      // every thread's have, and the column tile counts one more m
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0) sync_add(slot(kFusedSyncHead + t), 1);
    } else {
      const int tile = -2 - item;
      fft_inv_tile<L, float, true>(reinterpret_cast<double2 *>(lds), twl, nullptr, a.four, a.grid, 0,
                                   a.coslat, a.P2, a.N + 1, a.nf, a.jm, tile, ltid);
      if (bad) {
        constexpr int G = GF;
        const FourTile<G> ft(tile, a.P2);
        const int p = ft.p, f0 = ft.fgr * G;
        for (int q = tid; q < G * 2 * L; q += 256) {
          const int g = q / (2 * L), r = q % (2 * L), f = f0 + g;
          const int j = r < L ? a.P2 + p : a.P2 - 1 - p;
          if (f < a.nf) a.grid[((size_t)f * a.jm + j) * L + (r % L)] = __builtin_nan("");
        }
      }
    }
  }
  if (tid == 0 && sync_add(slot(kSyncDone), 1) == (int)gridDim.x - 1) {
    // every other workgroup has finished: reset for the next call
    for (int k = 0; k < kFusedSyncHead + a.nct; ++k)
      if (k != kSyncErr) __hip_atomic_store(slot(k), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Forward: four (FP32 S/A halves) -> spec (FP64).  Workgroup = one (m, CT
// 16-column tiles); the S/A tiles (all latitude pairs, both parities) are
// split once into LDS; wave w of NW takes the (parity, 16-degree row tile)
// units w, w+NW, ... and runs each against every column tile with the same
// table fragments.  NSL = P2p32 / 32 exactly (the dispatch picks it from the
// plan).
template <int MODE, int NSL, int CT, int NW>
__global__ __launch_bounds__(NW * 64) void leg_fwd_mixed_kernel(
    const float *__restrict__ four, double *__restrict__ spec, const uint16_t *__restrict__ L,
    const int *__restrict__ loff, const int *__restrict__ lexp, const int *__restrict__ mlist,
    int mlen, int N, int P2, int P2p32, int ncol, int nct, FourGeom geo) {
  constexpr int NP = Form<MODE>::NP, NG = 2 * NSL;  // LDS groups per column tile
  __shared__ __align__(16) uint16_t bt[CT * NG * NP * 16 * kRowPitch];
  __shared__ int ebits[CT * NG * 16];
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + CT - 1) / CT, m, t)) return;
  const int ct0 = t * CT, nc = nct - ct0 < CT ? nct - ct0 : CT;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int nn = N - m + 1, ke = (nn + 1) >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol + ct0 * 16;
  const uint16_t *__restrict__ Lm = L + loff[m];
  const int nunits = 2 * nrt;

  auto a_frag = [&](int u, int s, int q) {
    const int par = u / nrt, rt = u % nrt;
    return *reinterpret_cast<const uint4 *>(
        Lm + ((size_t)(par * NP + q) * KR + rt * 16 + r16) * P2p32 + s * 32 + kq * 8);
  };
  uint4 pa[NSL][3];
  if (wave < nunits) {
#pragma unroll
    for (int s = 0; s < NSL; ++s)
#pragma unroll
      for (int q = 0; q < 3; ++q) pa[s][q] = (q < NP) ? a_frag(wave, s, q) : make_uint4(0, 0, 0, 0);
  }
  stage_data<MODE, stage_iters<CT * NG, NW>(), NW>(bt, ebits, nc * NG, [](int g) { return g; }, [&](int g, int k, int c) {
    const int j = g / NG, par = (g % NG) / NSL, s = g % NSL, p = s * 32 + k, cc = ct0 * 16 + j * 16 + c;
    const bool v = p < P2 && cc < ncol;
    const float x = four[geo.at(v ? p : 0, m, par, v ? cc : ct0 * 16)];
    return v ? x : 0.0f;
  }, threadIdx.x);
  __syncthreads();
  for (int u = wave; u < nunits; u += NW) {
    const int par = u / nrt, rt = u % nrt;
    const int kc = par ? (nn >> 1) : ke;
    if (u != wave) {
#pragma unroll
      for (int s = 0; s < NSL; ++s)
#pragma unroll
        for (int q = 0; q < 3; ++q) pa[s][q] = (q < NP) ? a_frag(u, s, q) : make_uint4(0, 0, 0, 0);
    }
    const double ea = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m + par]) : 1.0;
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      if (j >= nc) break;
      d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        const int g = j * NG + par * NSL + s;
        uint4 pb[3];
        load_pb<MODE>(bt, g, kq, r16, pb);
        slice_mac<MODE>(pa[s], pb, Form<MODE>::kScaled ? ea * ldexp(1.0, -ebits[g * 16 + r16]) : 1.0,
                        acc);
      }
      if ((ct0 + j) * 16 + r16 < ncol) {
        double *__restrict__ sj = sm + j * 16 + r16;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = rt * 16 + kq * 4 + i;
          if (k < kc) put_spec(sj + (size_t)(2 * k + par) * ncol, acc[i]);
        }
      }
    }
  }
}

// Wave shape: each wave owns one 16-column tile of an m and holds that tile's
// data, split into pieces, in registers as the B operand: lane l reads the 8
// values K = 32 s + 8 (l/16) .. +7 of column l % 16 and splits them itself
// (the FP16 exponent is a shuffle maximum over the 4 lanes of a column), so
// there is no LDS staging and no barrier.  The wave then runs every row tile
// (inverse: 16 latitude pairs; forward: parity x 16 degrees), loading the
// table fragments, which the four waves of a workgroup share through L1.
template <int MODE, typename T>
__device__ __forceinline__ void split8(const T (&v)[8], int e, uint4 (&pb)[3]) {
  uint32_t w[3][4];
#pragma unroll
  for (int i = 0; i < 8; i += 2) {
    uint16_t p0[3], p1[3];
    split<MODE>(Form<MODE>::kScaled ? ldexp(v[i], e) : v[i], p0);
    split<MODE>(Form<MODE>::kScaled ? ldexp(v[i + 1], e) : v[i + 1], p1);
#pragma unroll
    for (int q = 0; q < 3; ++q) w[q][i >> 1] = p0[q] | ((uint32_t)p1[q] << 16);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) pb[q] = make_uint4(w[q][0], w[q][1], w[q][2], w[q][3]);
}

template <int MODE, typename T>
__device__ __forceinline__ int col_exponent(const T (&v)[8]) {
  if constexpr (!Form<MODE>::kScaled) return 0;
  T mx = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) mx = fmax(mx, fabs(v[i]));
  mx = fmax(mx, __shfl_xor(mx, 16));
  mx = fmax(mx, __shfl_xor(mx, 32));
  return mx > 0 ? 13 - ilogb(mx) : 0;
}

template <int MODE, int NSL, bool NT>
__global__ __launch_bounds__(256) void leg_inv_mixed_wave_kernel(
    const double *__restrict__ spec, float *__restrict__ four, const uint16_t *__restrict__ L,
    const int *__restrict__ loff, const int *__restrict__ lexp, const int *__restrict__ mlist,
    int mlen, const double *__restrict__ zero, int N, int P2, int P2p, int ncol, int nct,
    FourGeom geo) {
  constexpr int NP = Form<MODE>::NP;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + 3) >> 2, m, t)) return;
  const int ct = t * 4 + wave;
  if (ct >= nct) return;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP32 = (ke + 31) & ~31, nsl = KP32 >> 5;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;
  const double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol + (cok ? col : 0);
  const uint16_t *__restrict__ Lm = L + loff[m];
  uint4 pb[2][NSL][3];
  double bsc[2][NSL];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int sl = 0; sl < NSL; ++sl) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = sl * 32 + kq * 8 + i;
        const bool ok = sl < nsl && cok && k < (q ? ko : ke);
        v[i] = (float)*(ok ? sm + (size_t)(2 * k + q) * ncol : zero);
      }
      const int e = col_exponent<MODE>(v);
      bsc[q][sl] = ldexp(1.0, -e);
      split8<MODE>(v, e, pb[q][sl]);
    }
  const int rtiles = P2p >> 4;
  for (int rt = 0; rt < rtiles; ++rt) {
    uint4 pa[2][NSL][3];
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          pa[par][sl][q] = (q < NP) ? *reinterpret_cast<const uint4 *>(
                                          Lm + ((size_t)(par * NP + q) * P2p + rt * 16 + r16) * KP32 +
                                          (sl < nsl ? sl : 0) * 32 + kq * 8)
                                    : make_uint4(0, 0, 0, 0);
    const double ea0 = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m]) : 1.0;
    const double ea1 = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m + 1]) : 1.0;
    d4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int sl = 0; sl < NSL; ++sl) {
      if (sl >= nsl) break;
      slice_mac2<MODE>(pa[0][sl], pb[0][sl], Form<MODE>::kScaled ? ea0 * bsc[0][sl] : 1.0, acc[0],
                       pa[1][sl], pb[1][sl], Form<MODE>::kScaled ? ea1 * bsc[1][sl] : 1.0, acc[1]);
    }
    if (cok) {
#pragma unroll
      for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int p = rt * 16 + kq * 4 + i;
          if (p < P2) put_four<NT>(four + geo.at(p, m, par, col), (float)acc[par][i]);
        }
    }
  }
}

// Streaming inverse (shape STAGED of the mixed forms, grids of 128 latitude
// pairs): the table stays in registers while the data stream past it.
// Workgroup = (m, a chunk of cs column tiles) of eight waves; wave w holds the
// table fragments of row tile w for the whole chunk, so each m's table is
// read once per chunk instead of once per one or two column tiles.  The
// spectral tiles pass through two LDS buffers: wave w stages group w of each
// column tile (2 nsl <= 8 groups: parity, K slice) from registers loaded PD
// tiles ahead.  (The compiler's wait counting has each tile's conversion wait
// for every earlier load, so fewer tiles stay in flight than issued; forcing
// exact counts -- the table awaited first, no branch around the loads, the
// steps unrolled -- measured slower below 776 fields and equal at 776.)
// Per output tile the products,
// their order and the stores are those of leg_inv_mixed_item, so the results
// equal the other shapes' bitwise.  One workgroup per CU (the table takes
// 72 of the six-product form's ~210 registers at 128 pairs).
template <int MODE, int NSL, bool NT, int PD>
__global__ __launch_bounds__(512) void leg_inv_mixed_stream_kernel(
    const double *__restrict__ spec, float *__restrict__ four, const uint16_t *__restrict__ L,
    const int *__restrict__ loff, const int *__restrict__ lexp, const int *__restrict__ mlist,
    int mlen, int N, int P2, int P2p, int ncol, int nct, int cs, FourGeom geo) {
  constexpr int NW = 8, NP = Form<MODE>::NP;
  using One = InvItemLds<MODE, NSL, 1, NW>;
  constexpr size_t kBuf = One::bt + One::ebits;
  __shared__ __align__(16) unsigned char lds[2 * kBuf + One::obuf];
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + cs - 1) / cs, m, t)) return;
  const int wave = wave_id(), lane = threadIdx.x & 63, c = lane & 15, k8 = lane >> 4;
  const int j0 = t * cs, nj = nct - j0 < cs ? nct - j0 : cs;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KP32 = (ke + 31) & ~31, nsl = KP32 >> 5;
  const double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol;
  const uint16_t *__restrict__ Lm = L + loff[m];
  const bool mma = wave < (P2p >> 4);
  uint4 pa[2][NSL][3];
  if (mma) {
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int s = 0; s < NSL; ++s)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          pa[par][s][q] = (q < NP && s < nsl) ? *reinterpret_cast<const uint4 *>(
                                                    Lm + ((size_t)(par * NP + q) * P2p + wave * 16 + c) * KP32 +
                                                    s * 32 + k8 * 8)
                                              : make_uint4(0, 0, 0, 0);
  }
  // this wave's group: parity gpar, K slice gs, LDS slot gpar NSL + gs
  const bool stg = wave < 2 * nsl;
  const int gpar = wave >= nsl, gs = wave - gpar * nsl, kend = gpar ? ko : ke;
  auto load = [&](int j, double (&v)[8]) {
    const int col = (j0 + j) * 16 + c;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int kk = gs * 32 + k8 * 8 + i;
      const bool ok = kk < kend && col < ncol;
      v[i] = sm[(size_t)(ok ? 2 * kk + gpar : 0) * ncol + (ok ? col : 0)];
    }
  };
  auto put = [&](int j, const double (&v)[8]) {
    unsigned char *b = lds + (j & 1) * kBuf;
    const int col = (j0 + j) * 16 + c;
    float f[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int kk = gs * 32 + k8 * 8 + i;
      f[i] = kk < kend && col < ncol ? (float)v[i] : 0.0f;  // split as FP32: see split(float)
    }
    const int e = col_exponent<MODE>(f);
    uint4 pc[3];
    split8<MODE>(f, e, pc);
    const int g = gpar * NSL + gs;
    if constexpr (Form<MODE>::kScaled)
      if (k8 == 0) reinterpret_cast<int *>(b + One::bt)[g * 16 + c] = e;
    uint16_t *bt = reinterpret_cast<uint16_t *>(b);
#pragma unroll
    for (int q = 0; q < NP; ++q)
      *reinterpret_cast<uint4 *>(bt + ((size_t)(g * NP + q) * 16 + c) * kRowPitch + k8 * 8) = pc[q];
  };
  const double ea0 = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m]) : 1.0;
  const double ea1 = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m + 1]) : 1.0;
  float *ob = reinterpret_cast<float *>(lds + 2 * kBuf) + wave * (16 * kOutPitch);
  // column tile j: its staged values in v, convert them, then (barrier) load
  // tile j + PD into v and multiply tile j
  auto step = [&](int j, double (&v)[8]) {
    if (stg) put(j, v);
    __syncthreads();
    if (stg && j + PD < nj) load(j + PD, v);
    if (mma) {
      const unsigned char *b = lds + (j & 1) * kBuf;
      inv_tile<MODE, NSL, NT, false>(reinterpret_cast<const uint16_t *>(b), reinterpret_cast<const int *>(b + One::bt),
                                     ob, pa, ea0, ea1, nsl, four, geo, P2, ncol, m, wave, j0 + j, lane);
    }
  };
  double v[PD][8];
  if (stg) {
#pragma unroll
    for (int d = 0; d < PD; ++d)
      if (d < nj) load(d, v[d]);
  }
  for (int j = 0; j < nj; j += PD) {
#pragma unroll
    for (int d = 0; d < PD; ++d)
      if (j + d < nj) step(j + d, v[d]);
  }
}

template <int MODE, int NSL>
__global__ __launch_bounds__(256) void leg_fwd_mixed_wave_kernel(
    const float *__restrict__ four, double *__restrict__ spec, const uint16_t *__restrict__ L,
    const int *__restrict__ loff, const int *__restrict__ lexp, const int *__restrict__ mlist,
    int mlen, const double *__restrict__ zero, int N, int P2, int P2p32, int ncol, int nct,
    FourGeom geo) {
  constexpr int NP = Form<MODE>::NP;
  const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  int m, t;
  if (!xcd_m_block(blockIdx.x, mlist, mlen, (nct + 3) >> 2, m, t)) return;
  const int ct = t * 4 + wave;
  if (ct >= nct) return;
  const int nn = N - m + 1, ke = (nn + 1) >> 1, ko = nn >> 1;
  const int KR = (ke + 15) & ~15, nrt = KR >> 4;
  const int nsl = P2p32 >> 5;
  const int col = ct * 16 + r16;
  const bool cok = col < ncol;
  const int csafe = cok ? col : 0;
  const float *__restrict__ fz = reinterpret_cast<const float *>(zero);
  const uint16_t *__restrict__ Lm = L + loff[m];
  uint4 pb[2][NSL][3];
  double bsc[2][NSL];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int sl = 0; sl < NSL; ++sl) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int p = sl * 32 + kq * 8 + i;
        const bool ok = sl < nsl && cok && p < P2;
        v[i] = *(ok ? four + geo.at(p, m, q, csafe) : fz);
      }
      const int e = col_exponent<MODE>(v);
      bsc[q][sl] = ldexp(1.0, -e);
      split8<MODE>(v, e, pb[q][sl]);
    }
  double *__restrict__ sm = spec + (size_t)coef_off(N, m) * ncol + col;
#pragma unroll
  for (int par = 0; par < 2; ++par) {  // unrolled: pb[par] stays in registers
    const int kc = par ? ko : ke;
    const double ea = Form<MODE>::kScaled ? ldexp(1.0, -lexp[2 * m + par]) : 1.0;
    for (int kt = 0; kt < nrt; ++kt) {
      uint4 pa[NSL][3];
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          pa[sl][q] = (q < NP) ? *reinterpret_cast<const uint4 *>(
                                     Lm + ((size_t)(par * NP + q) * KR + kt * 16 + r16) * P2p32 +
                                     (sl < nsl ? sl : 0) * 32 + kq * 8)
                               : make_uint4(0, 0, 0, 0);
      d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl) {
        if (sl >= nsl) break;
        slice_mac<MODE>(pa[sl], pb[par][sl], Form<MODE>::kScaled ? ea * bsc[par][sl] : 1.0, acc);
      }
      if (cok) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = kt * 16 + kq * 4 + i;
          if (k < kc) put_spec(sm + (size_t)(2 * k + par) * ncol, acc[i]);
        }
      }
    }
  }
}

// Plan time: split and transpose one (m, parity) block of an FP64 table into
// the 16-bit piece layout.  For SCALED_FP16 the block's largest magnitude sets
// the power of two e (stored value = x 2^e), written to lexp[2m + par].
// inv = 1: src [KP][P2p] (row k, pair fastest)  -> dst [piece][P2p][KP32]
// inv = 0: src [P2p][KR] (row pair, k fastest)  -> dst [piece][KR][P2p32]
template <int MODE>
__global__ __launch_bounds__(256) void split_table_kernel(const double *__restrict__ src_tab,
                                                          const int *__restrict__ poff,
                                                          uint16_t *__restrict__ dst_tab,
                                                          const int *__restrict__ loff,
                                                          int *__restrict__ lexp, int N, int P2p,
                                                          int P2p32, int inv) {
  constexpr int NP = Form<MODE>::NP;
  const int m = blockIdx.x >> 1, par = blockIdx.x & 1;
  const int ke = ceil_half(N - m + 1);
  const int KP = roundup(ke, 4), KP32 = roundup(ke, 32), KR = roundup(ke, 16);
  const int rows = inv ? KP : P2p, cols = inv ? P2p : KR;  // source block shape
  const double *__restrict__ src = src_tab + poff[m] + (size_t)par * rows * cols;
  const int drows = inv ? P2p : KR, dcols = inv ? KP32 : P2p32;
  uint16_t *__restrict__ dst = dst_tab + loff[m] + (size_t)par * NP * drows * dcols;
  __shared__ double red[256];
  __shared__ int e_sh;
  double mx = 0.0;
  for (int t = threadIdx.x; t < rows * cols; t += 256) mx = fmax(mx, fabs(src[t]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int e = (Form<MODE>::kScaled && red[0] > 0.0) ? 13 - ilogb(red[0]) : 0;
    e_sh = e;
    lexp[2 * m + par] = e;
  }
  __syncthreads();
  const int e = e_sh;
  for (int t = threadIdx.x; t < rows * cols; t += 256) {
    const int r = t / cols, c = t % cols;
    const double x = Form<MODE>::kScaled ? ldexp(src[t], e) : src[t];
    uint16_t pc[3];
    split<MODE>(x, pc);
    // transpose: destination row = source column
    const size_t d = (size_t)c * dcols + r;
#pragma unroll
    for (int q = 0; q < NP; ++q) dst[(size_t)q * drows * dcols + d] = pc[q];
  }
}

template <int MODE>
int build_mixed_mode(Plan &p) {
  constexpr int NP = Form<MODE>::NP;
  const int N = p.N;
  const int P2p32 = roundup(p.P2p, 32);
  std::vector<int> li(N + 1), lf(N + 1);
  size_t ti = 0, tf = 0;
  for (int m = 0; m <= N; ++m) {
    const int ke = ceil_half(N - m + 1);
    li[m] = (int)ti;
    lf[m] = (int)tf;
    ti += (size_t)2 * NP * p.P2p * roundup(ke, 32);
    tf += (size_t)2 * NP * roundup(ke, 16) * P2p32;
  }
  if (ti > (size_t)INT32_MAX || tf > (size_t)INT32_MAX)
    return fail(SHT_ERR_UNSUPPORTED, "mixed-precision table exceeds 2^31 entries");
  p.n_linv = ti;
  p.n_lfwd = tf;
  SHT_PLAN_ALLOC(p, &p.Linv, sizeof(uint16_t) * ti);
  SHT_PLAN_ALLOC(p, &p.Lfwd, sizeof(uint16_t) * tf);
  SHT_PLAN_ALLOC(p, &p.loff_inv, sizeof(int) * (N + 1));
  SHT_PLAN_ALLOC(p, &p.loff_fwd, sizeof(int) * (N + 1));
  SHT_PLAN_ALLOC(p, &p.lexp_inv, sizeof(int) * 2 * (N + 1));
  SHT_PLAN_ALLOC(p, &p.lexp_fwd, sizeof(int) * 2 * (N + 1));
  SHT_HIP(hipMemcpy(p.loff_inv, li.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
  SHT_HIP(hipMemcpy(p.loff_fwd, lf.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
  SHT_HIP(hipMemset(p.Linv, 0, sizeof(uint16_t) * ti));
  SHT_HIP(hipMemset(p.Lfwd, 0, sizeof(uint16_t) * tf));
  split_table_kernel<MODE><<<2 * (N + 1), 256>>>(p.Pinv, p.poff_inv, static_cast<uint16_t *>(p.Linv),
                                                 p.loff_inv, p.lexp_inv, N, p.P2p, P2p32, 1);
  SHT_HIP(hipGetLastError());
  split_table_kernel<MODE><<<2 * (N + 1), 256>>>(p.Pfwd, p.poff_fwd, static_cast<uint16_t *>(p.Lfwd),
                                                 p.loff_fwd, p.lexp_fwd, N, p.P2p, P2p32, 0);
  SHT_HIP(hipGetLastError());
  SHT_HIP(hipDeviceSynchronize());
  return SHT_OK;
}

// Column tiles per workgroup of the streaming inverse: at most 8, and at
// least 3 workgroups per m once there are more than 3 tiles, the chunks
// balanced (one workgroup per CU runs, so a short last chunk or too few
// workgroups leaves CUs idle).  Measured at T170 (split-BF16 x6): 776
// fields 4 / 8 / 12 / 16 / 32 tiles 225 / 212 / 211 / 219 / 252 us; 200
// fields (25 tiles) 8 tiles 52.1 us, 6 tiles 44.1 us; 33 fields (5 tiles)
// 5 tiles 15.6 us, 2 tiles 13.6 us.
int stream_chunk(int nct) {
  if (nct <= 3) return nct;
  const int nchunks = std::max(3, (nct + 7) / 8);
  return (nct + nchunks - 1) / nchunks;
}

template <int MODE, int NSL>
int inv_nsl(const Plan &p, int shape, const double *spec, float *four, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  const FourGeom geo = four_geom_out(p, nf, sizeof(float));
  if (shape == SHT_SHAPE_WAVE) {
    launch_kernel(geo.stream ? leg_inv_mixed_wave_kernel<MODE, NSL, true> : leg_inv_mixed_wave_kernel<MODE, NSL, false>,
                  xcd_m_blocks(p, 0, (nct + 3) / 4), 256, 0, s, spec, four,
                  static_cast<const uint16_t *>(p.Linv), p.loff_inv, p.lexp_inv, p.xcd_m[0],
                  p.xcd_mlen[0], p.zero, p.N, p.P2, p.P2p, ncol, nct, geo);
    SHT_HIP(hipGetLastError());
    return SHT_OK;
  }
  if (shape == SHT_SHAPE_STAGED) {
    if (!mixed_stream_available(p)) return fail(SHT_ERR_UNSUPPORTED, "mixed inverse: streaming shape needs <= 128 latitude pairs");
    // loads 4 tiles ahead (T170, 776 fields, split-BF16 x6, 8 tiles per
    // workgroup: 2 / 3 / 4 tiles ahead 219 / 216 / 212 us)
    const int cs = stream_chunk(nct);
    launch_kernel(geo.stream ? leg_inv_mixed_stream_kernel<MODE, NSL, true, 4> : leg_inv_mixed_stream_kernel<MODE, NSL, false, 4>,
                  xcd_m_blocks(p, 0, (nct + cs - 1) / cs), 512, 0, s, spec, four,
                  static_cast<const uint16_t *>(p.Linv), p.loff_inv, p.lexp_inv, p.xcd_m[0],
                  p.xcd_mlen[0], p.N, p.P2, p.P2p, ncol, nct, cs, geo);
    SHT_HIP(hipGetLastError());
    return SHT_OK;
  }
  // KSPLIT: one column tile per workgroup; TILE: two.  Four waves (eight,
  // as the forward takes them at T170, measured slower here: 242 against
  // 217 us for TILE at T170, 776 fields).
  auto kern = shape == SHT_SHAPE_TILE
                  ? (geo.stream ? leg_inv_mixed_kernel<MODE, NSL, 2, true, 4> : leg_inv_mixed_kernel<MODE, NSL, 2, false, 4>)
                  : (geo.stream ? leg_inv_mixed_kernel<MODE, NSL, 1, true, 4> : leg_inv_mixed_kernel<MODE, NSL, 1, false, 4>);
  const int ct = shape == SHT_SHAPE_TILE ? 2 : 1;
  launch_kernel(kern, xcd_m_blocks(p, 0, (nct + ct - 1) / ct), 256, 0, s, spec, four,
                static_cast<const uint16_t *>(p.Linv), p.loff_inv, p.lexp_inv, p.xcd_m[0],
                p.xcd_mlen[0], p.N, p.P2, p.P2p, ncol, nct, geo);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int MODE>
int inv_mode(const Plan &p, int shape, const double *spec, float *four, int nf, hipStream_t s) {
  switch (roundup(ceil_half(p.N + 1), 32) / 32) {
    case 1: return inv_nsl<MODE, 1>(p, shape, spec, four, nf, s);
    case 2: return inv_nsl<MODE, 2>(p, shape, spec, four, nf, s);
    case 3: return inv_nsl<MODE, 3>(p, shape, spec, four, nf, s);
    case 4: return inv_nsl<MODE, 4>(p, shape, spec, four, nf, s);
  }
  return fail(SHT_ERR_UNSUPPORTED, "mixed inverse: nmax above 255");
}

// Fused inverse dispatch: f(L, MODE, NSL) as integral constants for the
// plan's grid, form and K slices; SHT_ERR_UNSUPPORTED outside the
// instantiated set (im 64..512, at most 3 slices, i.e. nmax < 192).  As
// im >= 2 nmax + 2, a grid of L longitudes needs at most L / 128 slices.
template <class F>
int fused_dispatch(const Plan &p, F &&f) {
  using std::integral_constant;
  const int nsl = roundup(ceil_half(p.N + 1), 32) / 32;
  auto by_nsl = [&](auto l, auto mode) -> int {
    constexpr int L = decltype(l)::value;
    if (nsl == 1) return f(l, mode, integral_constant<int, 1>{});
    if constexpr (L >= 256)
      if (nsl == 2) return f(l, mode, integral_constant<int, 2>{});
    if constexpr (L >= 512)
      if (nsl == 3) return f(l, mode, integral_constant<int, 3>{});
    return fail(SHT_ERR_UNSUPPORTED, "fused inverse: nmax above 191");
  };
  auto by_mode = [&](auto l) -> int {
    switch (p.prec) {
      case SHT_PREC_SPLIT_BF16: return by_nsl(l, integral_constant<int, SHT_PREC_SPLIT_BF16>{});
      case SHT_PREC_SCALED_FP16: return by_nsl(l, integral_constant<int, SHT_PREC_SCALED_FP16>{});
      case SHT_PREC_SPLIT_BF16X6: return by_nsl(l, integral_constant<int, SHT_PREC_SPLIT_BF16X6>{});
    }
    return fail(SHT_ERR_UNSUPPORTED, "fused inverse: mixed-precision plans only");
  };
  switch (p.im) {
    case 64: return by_mode(integral_constant<int, 64>{});
    case 128: return by_mode(integral_constant<int, 128>{});
    case 256: return by_mode(integral_constant<int, 256>{});
    case 512: return by_mode(integral_constant<int, 512>{});
  }
  return fail(SHT_ERR_UNSUPPORTED, "fused inverse: im must be 64 to 512");
}

bool fused_supported(const Plan &p) {
  return p.prec != SHT_PREC_FP64 && p.im >= 64 && p.im <= 512 && ceil_half(p.N + 1) <= 96;
}

// Sync area and residency of the fused inverse (plan creation).
int build_fused(Plan &p) {
  if (!fused_supported(p)) return SHT_OK;
  const size_t n = (kFusedSyncHead + (size_t)(2 * p.kmax + 15) / 16) * kSyncStride;
  SHT_PLAN_ALLOC(p, &p.fsync, sizeof(int) * n);
  SHT_HIP(hipMemset(p.fsync, 0, sizeof(int) * n));
  int ncu = 0;
  SHT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p.device));
  return fused_dispatch(p, [&](auto l, auto mode, auto nsl) -> int {
    int per_cu = 0;
    SHT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, inv_fused_kernel<decltype(l)::value, decltype(mode)::value, decltype(nsl)::value>, 256, 0));
    p.fused_blocks = per_cu * ncu;
    return SHT_OK;
  });
}

template <int MODE, int NSL>
int fwd_nsl(const Plan &p, int shape, const float *four, double *spec, int nf, hipStream_t s) {
  const int ncol = 2 * nf, nct = (ncol + 15) / 16;
  if (shape == SHT_SHAPE_WAVE) {
    launch_kernel(leg_fwd_mixed_wave_kernel<MODE, NSL>, xcd_m_blocks(p, 1, (nct + 3) / 4), 256, 0, s,
                  four, spec, static_cast<const uint16_t *>(p.Lfwd), p.loff_fwd, p.lexp_fwd,
                  p.xcd_m[1], p.xcd_mlen[1], p.zero, p.N, p.P2, roundup(p.P2p, 32), ncol, nct,
                  four_geom(p, nf));
    SHT_HIP(hipGetLastError());
    return SHT_OK;
  }
  // Eight waves share the data tile from 128 latitude pairs (T170: TILE 171
  // against 195 us at 776 fields, 26 against 31 us at 97; KSPLIT neutral),
  // four below.
  const bool eight = p.P2p >= 128;
  auto kern = shape == SHT_SHAPE_TILE
                  ? (eight ? leg_fwd_mixed_kernel<MODE, NSL, 2, 8> : leg_fwd_mixed_kernel<MODE, NSL, 2, 4>)
                  : (eight ? leg_fwd_mixed_kernel<MODE, NSL, 1, 8> : leg_fwd_mixed_kernel<MODE, NSL, 1, 4>);
  const int ct = shape == SHT_SHAPE_TILE ? 2 : 1;
  launch_kernel(kern, xcd_m_blocks(p, 1, (nct + ct - 1) / ct), eight ? 512 : 256, 0, s, four, spec,
                static_cast<const uint16_t *>(p.Lfwd), p.loff_fwd, p.lexp_fwd, p.xcd_m[1],
                p.xcd_mlen[1], p.N, p.P2, roundup(p.P2p, 32), ncol, nct, four_geom(p, nf));
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <int MODE>
int fwd_mode(const Plan &p, int shape, const float *four, double *spec, int nf, hipStream_t s) {
  switch (roundup(p.P2p, 32) / 32) {
    case 1: return fwd_nsl<MODE, 1>(p, shape, four, spec, nf, s);
    case 2: return fwd_nsl<MODE, 2>(p, shape, four, spec, nf, s);
    case 3: return fwd_nsl<MODE, 3>(p, shape, four, spec, nf, s);
    case 4: return fwd_nsl<MODE, 4>(p, shape, four, spec, nf, s);
    case 5: return fwd_nsl<MODE, 5>(p, shape, four, spec, nf, s);
    case 6: return fwd_nsl<MODE, 6>(p, shape, four, spec, nf, s);
    case 7: return fwd_nsl<MODE, 7>(p, shape, four, spec, nf, s);
    case 8: return fwd_nsl<MODE, 8>(p, shape, four, spec, nf, s);
  }
  return fail(SHT_ERR_UNSUPPORTED, "mixed forward: jm above 512");
}

}  // namespace

int build_mixed_tables(Plan &p) {
  int rc;
  switch (p.prec) {
    case SHT_PREC_SPLIT_BF16: rc = build_mixed_mode<SHT_PREC_SPLIT_BF16>(p); break;
    case SHT_PREC_SCALED_FP16: rc = build_mixed_mode<SHT_PREC_SCALED_FP16>(p); break;
    case SHT_PREC_SPLIT_BF16X6: rc = build_mixed_mode<SHT_PREC_SPLIT_BF16X6>(p); break;
    default: return fail(SHT_ERR_INVALID, "build_mixed_tables: not a mixed-precision plan");
  }
  return rc ? rc : build_fused(p);
}

int launch_leg_inv_mixed_shape(const Plan &p, int shape, const double *spec, float *four, int nf,
                               hipStream_t s) {
  switch (p.prec) {
    case SHT_PREC_SPLIT_BF16: return inv_mode<SHT_PREC_SPLIT_BF16>(p, shape, spec, four, nf, s);
    case SHT_PREC_SCALED_FP16: return inv_mode<SHT_PREC_SCALED_FP16>(p, shape, spec, four, nf, s);
    case SHT_PREC_SPLIT_BF16X6: return inv_mode<SHT_PREC_SPLIT_BF16X6>(p, shape, spec, four, nf, s);
  }
  return fail(SHT_ERR_INVALID, "launch_leg_inv_mixed: not a mixed-precision plan");
}

int launch_leg_fwd_mixed_shape(const Plan &p, int shape, const float *four, double *spec, int nf,
                               hipStream_t s) {
  switch (p.prec) {
    case SHT_PREC_SPLIT_BF16: return fwd_mode<SHT_PREC_SPLIT_BF16>(p, shape, four, spec, nf, s);
    case SHT_PREC_SCALED_FP16: return fwd_mode<SHT_PREC_SCALED_FP16>(p, shape, four, spec, nf, s);
    case SHT_PREC_SPLIT_BF16X6: return fwd_mode<SHT_PREC_SPLIT_BF16X6>(p, shape, four, spec, nf, s);
  }
  return fail(SHT_ERR_INVALID, "launch_leg_fwd_mixed: not a mixed-precision plan");
}

int launch_leg_inv_mixed(const Plan &p, const double *spec, float *four, int nf, hipStream_t s) {
  return launch_leg_inv_mixed_shape(p, resolve_shape(p, nf, 0), spec, four, nf, s);
}

bool fused_inv_available(const Plan &p) { return fused_supported(p) && p.fsync && p.fused_blocks >= 8; }

int fused_wait_failures(const Plan &p, long long *count) {
  *count = 0;
  if (!p.fsync) return SHT_OK;
  int v = 0;
  SHT_HIP(hipMemcpy(&v, p.fsync + kSyncErr * kSyncStride, sizeof(int), hipMemcpyDeviceToHost));
  *count = v;
  return SHT_OK;
}

int launch_inv_fused(const Plan &p, const double *spec, float *four, double *grid, int nf, hipStream_t s) {
  if (!fused_inv_available(p)) return fail(SHT_ERR_UNSUPPORTED, "fused inverse: not available for this plan");
  FusedInv a{};
  a.spec = spec;
  a.four = four;
  a.grid = grid;
  a.L = static_cast<const uint16_t *>(p.Linv);
  a.loff = p.loff_inv;
  a.lexp = p.lexp_inv;
  a.mlist = p.xcd_m[0];
  a.tw = reinterpret_cast<const double2 *>(p.twid);
  a.coslat = p.coslat;
  a.sync = p.fsync;
  a.mlen = p.xcd_mlen[0];
  a.N = p.N;
  a.P2 = p.P2;
  a.P2p = p.P2p;
  a.ncol = 2 * nf;
  a.nct = (a.ncol + 15) / 16;
  a.nf = nf;
  a.jm = p.jm;
  a.geo = four_geom_out(p, nf, sizeof(float), fft_group(p.im));  // the big FFT blocks
  a.need = p.N + 1 + (p.fused_inject ? 1 : 0);
  a.spin_ticks = p.fused_inject ? kFusedInjectedSpinTicks : kFusedSpinTicks;
  // no more workgroups than items, a multiple of 8 (one queue per XCD)
  const int gf = fft_group(p.im);
  const long items = 8L * a.mlen * a.nct + (long)p.P2 * a.geo.nfg * (layout_group(gf) / gf);
  const int nblk = (int)std::max(8L, std::min<long>(p.fused_blocks, items) / 8 * 8);
  return fused_dispatch(p, [&](auto l, auto mode, auto nsl) -> int {
    launch_kernel(inv_fused_kernel<decltype(l)::value, decltype(mode)::value, decltype(nsl)::value>, nblk, 256,
                  0, s, a);
    SHT_HIP(hipGetLastError());
    return SHT_OK;
  });
}

int launch_leg_fwd_mixed(const Plan &p, const float *four, double *spec, int nf, hipStream_t s) {
  return launch_leg_fwd_mixed_shape(p, resolve_shape(p, nf, 1), four, spec, nf, s);
}

}  // namespace sht

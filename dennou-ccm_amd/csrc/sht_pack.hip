// sht_pack.hip -- spectral layout conversion between SPML's real (N+1)^2
// packing (Fortran wa((N+1)**2, k), index l = n(n+1)+m+1 of SPML l_nm) and the
// canonical complex m-major layout the Legendre GEMMs read, plus the global
// mean AvrLonLat_xy (atm/dccm_atm_mod.f90:668-669, :774) and the diagonal
// spectral operators (SPML wa_Lapla_wa / wa_LaplaInv_wa and their powers, the
// hyperdiffusion of DCPAM's dynamics namelist, HDOrder in
// exp/APEI07Couple/common/dcpam/dcpam_APEI07Couple_T21L16.conf:205-230).
#include "sht_internal.h"

namespace sht {
namespace {

constexpr double kSqrt2 = 1.41421356237309504880;

// one thread per (canonical k, field), field fastest: canonical writes coalesce
__global__ void spml_to_canon_kernel(const double *__restrict__ wa, double *__restrict__ spec,
                                     int N, int nf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int K = (N + 1) * (N + 2) / 2;
  if (t >= (int64_t)K * nf) return;
  const int f = (int)(t % nf);
  const int k = (int)(t / nf);
  // invert k = off(m) + n - m: off(m) = m (2N + 3 - m) / 2, so the largest m
  // with off(m) <= k is the floor of the smaller root of m^2 - (2N+3) m + 2k,
  // corrected by a step where the square root rounds across an integer
  const double b = 2.0 * N + 3.0;
  int m = (int)floor(0.5 * (b - sqrt(b * b - 8.0 * (double)k)));
  m = m < 0 ? 0 : m > N ? N : m;
  while (m > 0 && coef_off(N, m) > k) --m;
  while (m < N && coef_off(N, m + 1) <= k) ++m;
  const int n = m + (k - (int)coef_off(N, m));
  const double *w = wa + (size_t)f * (N + 1) * (N + 1);
  double re, im;
  if (m == 0) {
    re = w[n * (n + 1)];
    im = 0.0;
  } else {
    re = w[n * (n + 1) + m] / kSqrt2;
    im = -w[n * (n + 1) - m] / kSqrt2;
  }
  spec[2 * t] = re;
  spec[2 * t + 1] = im;
}

// one thread per (field, l), l fastest: SPML-layout writes coalesce
__global__ void canon_to_spml_kernel(const double *__restrict__ spec, double *__restrict__ wa,
                                     int N, int nf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int L = (N + 1) * (N + 1);
  if (t >= (int64_t)L * nf) return;
  const int l = (int)(t % L);
  const int f = (int)(t / L);
  int n = (int)sqrt((double)l);
  while (n * n > l) --n;
  while ((n + 1) * (n + 1) <= l) ++n;
  const int m = l - n * (n + 1);
  const int am = m < 0 ? -m : m;
  const double *s = spec + ((size_t)(coef_off(N, am) + n - am) * nf + f) * 2;
  double v;
  if (m == 0) v = s[0];
  else if (m > 0) v = kSqrt2 * s[0];
  else v = -kSqrt2 * s[1];
  wa[t] = v;
}

// one block per field: (1/4pi) sum_ij (2pi/im) w_j f_ij
__global__ __launch_bounds__(256) void avr_kernel(const double *__restrict__ grid,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ avg, int im, int jm) {
  const int f = blockIdx.x;
  const double *g = grid + (size_t)f * jm * im;
  double acc = 0.0;
  for (int t = threadIdx.x; t < im * jm; t += blockDim.x) acc += w[t / im] * g[t];
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) avg[f] = red[0] / (2.0 * im);
}

// one thread per (canonical k, field) complex pair: out = (-n(n+1))^power in
__global__ void lapla_kernel(const double2 *__restrict__ in, double2 *__restrict__ out, int N,
                             int nf, int power) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int K = (N + 1) * (N + 2) / 2;
  if (t >= (int64_t)K * nf) return;
  const int k = (int)(t / nf);
  int m = 0;
  while (m < N && coef_off(N, m + 1) <= k) ++m;
  const int n = m + (k - (int)coef_off(N, m));
  const double e = -(double)n * (n + 1);
  double fac = 1.0;
  if (power < 0 && n == 0) {
    fac = 0.0;
  } else {
    const int ap = power < 0 ? -power : power;
    for (int i = 0; i < ap; ++i) fac = power < 0 ? fac / e : fac * e;
  }
  const double2 v = in[t];
  out[t] = make_double2(fac * v.x, fac * v.y);
}

}  // namespace

int launch_spec_lapla(const Plan &p, const double *in, double *out, int power, int nf,
                      hipStream_t s) {
  const int64_t n = (int64_t)p.K * nf;
  lapla_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
      reinterpret_cast<const double2 *>(in), reinterpret_cast<double2 *>(out), p.N, nf, power);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_spml_to_canon(const Plan &p, const double *wa, double *spec, int nf, hipStream_t s) {
  const int64_t n = (int64_t)p.K * nf;
  spml_to_canon_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(wa, spec, p.N, nf);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_canon_to_spml(const Plan &p, const double *spec, double *wa, int nf, hipStream_t s) {
  const int64_t n = (int64_t)(p.N + 1) * (p.N + 1) * nf;
  canon_to_spml_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(spec, wa, p.N, nf);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_avr_lonlat(const Plan &p, const double *grid, double *avg, int nf, hipStream_t s) {
  avr_kernel<<<nf, 256, 0, s>>>(grid, p.w, avg, p.im, p.jm);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace sht

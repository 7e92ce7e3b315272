// sht_zonal.hip -- the zonal (m = 0) Legendre transform of SPML's l_module:
// l_Initial(nm, jm), y_Lat, y_Lat_Weight, l_y (grid -> Legendre) and y_l
// (Legendre -> grid), as the energy-balance model calls them
// (exp/APEI07Couple/ebm/EBM_main.f90:64 l_initial(1024, 1024), :170 and :242
// l_y, with y_Lat_Weight at :200-204).
//
//   l_y:  c_n = (1/2) sum_j w_j Pt_n(mu_j) y_j,   y_l:  y_j = sum_n c_n Pt_n(mu_j)
//   Pt_n = sqrt(2n+1) P_n, (1/2) int Pt_n^2 dmu = 1 (the normalisation
//   EBM_main.f90:170-174 and :242-246 undo with sqrt(2n+1)).
//
// Unlike the wa plan, nm may equal jm (the EBM's own configuration): degree jm
// then vanishes at every Gauss node, so l_y returns 0 for it and y_l ignores it.
// Layouts: y[f][j] (Fortran y(1:jm, 1:k), latitudes south to north),
// l[f][n] (Fortran l(0:nm, 1:k)).
//
// Tables (FP64, computed on the GPU by the same three-term recurrence as the
// wa plan's m = 0 column): P2 = jm/2 latitude pairs, pair p = north node
// P2 + p and south node P2 - 1 - p; degrees split by parity, n = 2k + q.
//   Tinv [q][KP][P2p]  Pt_n(mu_p)          (inverse; pairs fastest)
//   Tfwd [q][P2p][KR]  (w_p/2) Pt_n(mu_p)  (forward; degrees fastest)
// Both directions are GEMMs on v_mfma_f64_16x16x4f64, one wavefront per
// 16 x 16 output tile (the EBM's single field uses one column of it).
#include <cstring>
#include <new>

#include "sht_internal.h"

namespace sht {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// A: lane l holds row l % 16, K index l / 16; B: K index l / 16, column
// l % 16; result item i of lane l is row 4 i + l / 16, column l % 16.
__device__ __forceinline__ d4 mfma16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double eps_n0(int n) {
  const double nn = n;
  return sqrt(nn * nn / (4.0 * nn * nn - 1.0));
}

// one thread per latitude pair: Pt_{n+1} = (mu Pt_n - eps_n Pt_{n-1}) / eps_{n+1}
__global__ void zonal_table_kernel(int N, int P2, int P2p, int KP, int KR, const double *__restrict__ mu,
                                   const double *__restrict__ w, double *__restrict__ Tinv,
                                   double *__restrict__ Tfwd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P2) return;
  const double x = mu[P2 + p], hw = 0.5 * w[P2 + p];
  double pc = 1.0, pp = 0.0;
  for (int n = 0; n <= N; ++n) {
    const double back = n > 0 ? eps_n0(n) * pp : 0.0;
    const double pn = (x * pc - back) / eps_n0(n + 1);
    const int q = n & 1, k = n >> 1;
    Tinv[(size_t)(q * KP + k) * P2p + p] = pc;
    Tfwd[(size_t)(q * P2p + p) * KR + k] = hw * pc;
    pp = pc;
    pc = pn;
  }
}

// l_y: wave = (parity q, 16-degree tile, 16-field tile); K runs over pairs.
__global__ __launch_bounds__(64) void zonal_fwd_kernel(const double *__restrict__ y, double *__restrict__ l,
                                                        const double *__restrict__ Tfwd, int N, int jm, int P2,
                                                        int P2p, int KR, int nrt, int nf) {
  const int b = blockIdx.x, q = b & 1, rt = (b >> 1) % nrt, ct = (b >> 1) / nrt;
  const int lane = threadIdx.x, r16 = lane & 15, kq = lane >> 4;
  const int f = ct * 16 + r16;
  const double *__restrict__ T = Tfwd + (size_t)q * P2p * KR + rt * 16 + r16;
  const double *__restrict__ yf = y + (size_t)(f < nf ? f : 0) * jm;
  const double sgn = q ? -1.0 : 1.0;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < (P2p >> 2); ++s) {
    const int p = s * 4 + kq;
    const double a = T[(size_t)p * KR];
    double v = 0.0;
    if (p < P2 && f < nf) v = yf[P2 + p] + sgn * yf[P2 - 1 - p];
    acc = mfma16x4(a, v, acc);
  }
  if (f >= nf) return;
  const int kn = q ? (N + 1) >> 1 : (N + 2) >> 1;  // degrees of this parity
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = rt * 16 + i * 4 + kq;
    if (k < kn) l[(size_t)f * (N + 1) + 2 * k + q] = acc[i];
  }
}

// y_l: wave = (16-pair tile, 16-field tile); even and odd degrees in two
// accumulators, then north = even + odd, south = even - odd.
__global__ __launch_bounds__(64) void zonal_inv_kernel(const double *__restrict__ l, double *__restrict__ y,
                                                        const double *__restrict__ Tinv, int N, int jm, int P2,
                                                        int P2p, int KP, int nrt, int nf) {
  const int b = blockIdx.x, rt = b % nrt, ct = b / nrt;
  const int lane = threadIdx.x, r16 = lane & 15, kq = lane >> 4;
  const int f = ct * 16 + r16;
  const double *__restrict__ lf = l + (size_t)(f < nf ? f : 0) * (N + 1);
  d4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double *__restrict__ T = Tinv + (size_t)q * KP * P2p + rt * 16 + r16;
    for (int s = 0; s < (KP >> 2); ++s) {
      const int k = s * 4 + kq, n = 2 * k + q;
      const double a = T[(size_t)k * P2p];
      const double v = (n <= N && f < nf) ? lf[n] : 0.0;
      acc[q] = mfma16x4(a, v, acc[q]);
    }
  }
  if (f >= nf) return;
  double *__restrict__ yf = y + (size_t)f * jm;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = rt * 16 + i * 4 + kq;
    if (p < P2) {
      yf[P2 + p] = acc[0][i] + acc[1][i];
      yf[P2 - 1 - p] = acc[0][i] - acc[1][i];
    }
  }
}

}  // namespace
}  // namespace sht

struct sht_zonal {
  int N = 0, jm = 0, P2 = 0, P2p = 0, KP = 0, KR = 0, kmax = 0;
  double *mu = nullptr, *w = nullptr, *coslat = nullptr;
  double *Tinv = nullptr, *Tfwd = nullptr;
  double *stage_a = nullptr, *stage_b = nullptr, *pinned = nullptr;
  hipStream_t host_stream = nullptr;
  std::vector<double> h_mu, h_w, h_coslat;
};

namespace sht {
namespace {

void zonal_release(sht_zonal *z) {
  void *ptrs[] = {z->mu, z->w, z->coslat, z->Tinv, z->Tfwd, z->stage_a, z->stage_b};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (z->pinned) (void)hipHostFree(z->pinned);
  if (z->host_stream) (void)hipStreamDestroy(z->host_stream);
  delete z;
}

int zonal_build(sht_zonal *z) {
  const size_t ninv = (size_t)2 * z->KP * z->P2p, nfwd = (size_t)2 * z->P2p * z->KR;
  SHT_HIP(hipMalloc(&z->mu, sizeof(double) * z->jm));
  SHT_HIP(hipMalloc(&z->w, sizeof(double) * z->jm));
  SHT_HIP(hipMalloc(&z->coslat, sizeof(double) * z->jm));
  SHT_HIP(hipMalloc(&z->Tinv, sizeof(double) * ninv));
  SHT_HIP(hipMalloc(&z->Tfwd, sizeof(double) * nfwd));
  SHT_HIP(hipMemset(z->Tinv, 0, sizeof(double) * ninv));
  SHT_HIP(hipMemset(z->Tfwd, 0, sizeof(double) * nfwd));
  int rc = launch_gauss(z->jm, z->mu, z->w, z->coslat, nullptr);
  if (rc) return rc;
  zonal_table_kernel<<<(z->P2 + 63) / 64, 64>>>(z->N, z->P2, z->P2p, z->KP, z->KR, z->mu, z->w, z->Tinv,
                                                 z->Tfwd);
  SHT_HIP(hipGetLastError());
  SHT_HIP(hipDeviceSynchronize());
  z->h_mu.resize(z->jm);
  z->h_w.resize(z->jm);
  z->h_coslat.resize(z->jm);
  SHT_HIP(hipMemcpy(z->h_mu.data(), z->mu, sizeof(double) * z->jm, hipMemcpyDeviceToHost));
  SHT_HIP(hipMemcpy(z->h_w.data(), z->w, sizeof(double) * z->jm, hipMemcpyDeviceToHost));
  SHT_HIP(hipMemcpy(z->h_coslat.data(), z->coslat, sizeof(double) * z->jm, hipMemcpyDeviceToHost));
  return SHT_OK;
}

int zonal_check(const sht_zonal *z, const void *a, const void *b, int nf, const char *who) {
  if (!z) return fail(SHT_ERR_INVALID, std::string(who) + ": plan is NULL");
  if (!a || !b) return fail(SHT_ERR_INVALID, std::string(who) + ": NULL array");
  if (nf < 1 || nf > z->kmax) return fail(SHT_ERR_INVALID, std::string(who) + ": nfield must be in [1, kmax]");
  return SHT_OK;
}

int zonal_fwd(const sht_zonal *z, const double *y, double *l, int nf, hipStream_t s) {
  const int nrt = z->KR / 16, nct = (nf + 15) / 16;
  zonal_fwd_kernel<<<2 * nrt * nct, 64, 0, s>>>(y, l, z->Tfwd, z->N, z->jm, z->P2, z->P2p, z->KR, nrt, nf);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

int zonal_inv(const sht_zonal *z, const double *l, double *y, int nf, hipStream_t s) {
  const int nrt = z->P2p / 16, nct = (nf + 15) / 16;
  zonal_inv_kernel<<<nrt * nct, 64, 0, s>>>(l, y, z->Tinv, z->N, z->jm, z->P2, z->P2p, z->KP, nrt, nf);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

// Synchronous host call through pinned staging: dir 0 = l_y, 1 = y_l.
int zonal_host(sht_zonal *z, const double *in, double *out, int nf, int dir, const char *who) {
  int rc = zonal_check(z, in, out, nf, who);
  if (rc) return rc;
  const size_t ny = (size_t)z->kmax * z->jm, nl = (size_t)z->kmax * (z->N + 1);
  if (!z->pinned) {
    SHT_HIP(hipStreamCreateWithFlags(&z->host_stream, hipStreamNonBlocking));
    SHT_HIP(hipMalloc(&z->stage_a, sizeof(double) * ny));
    SHT_HIP(hipMalloc(&z->stage_b, sizeof(double) * nl));
    SHT_HIP(hipHostMalloc(&z->pinned, sizeof(double) * (ny > nl ? ny : nl), hipHostMallocDefault));
  }
  hipStream_t s = z->host_stream;
  const size_t gy = (size_t)nf * z->jm, gl = (size_t)nf * (z->N + 1);
  const size_t nin = dir == 0 ? gy : gl, nout = dir == 0 ? gl : gy;
  double *din = dir == 0 ? z->stage_a : z->stage_b, *dout = dir == 0 ? z->stage_b : z->stage_a;
  std::memcpy(z->pinned, in, sizeof(double) * nin);
  SHT_HIP(hipMemcpyAsync(din, z->pinned, sizeof(double) * nin, hipMemcpyHostToDevice, s));
  rc = dir == 0 ? zonal_fwd(z, din, dout, nf, s) : zonal_inv(z, din, dout, nf, s);
  if (rc) return rc;
  SHT_HIP(hipMemcpyAsync(z->pinned, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
  SHT_HIP(hipStreamSynchronize(s));
  std::memcpy(out, z->pinned, sizeof(double) * nout);
  return SHT_OK;
}

}  // namespace
}  // namespace sht

using namespace sht;

extern "C" {

int sht_zonal_create(int nmax, int jm, int kmax, sht_zonal **out) {
  if (!out) return fail(SHT_ERR_INVALID, "sht_zonal_create: out is NULL");
  *out = nullptr;
  std::string msg;
  if (int rc = host::check_zonal_desc(nmax, jm, kmax, msg)) return fail(rc, msg);
  sht_zonal *z = new (std::nothrow) sht_zonal;
  if (!z) return fail(SHT_ERR_NOMEM, "sht_zonal_create: host allocation failed");
  z->N = nmax;
  z->jm = jm;
  z->P2 = jm / 2;
  z->P2p = roundup(z->P2, 16);
  z->KP = roundup(ceil_half(nmax + 1), 4);
  z->KR = roundup(ceil_half(nmax + 1), 16);
  z->kmax = kmax;
  int rc = zonal_build(z);
  if (rc) {
    std::string keep = sht_last_error();
    zonal_release(z);
    return fail(rc, keep);
  }
  *out = z;
  return SHT_OK;
}

void sht_zonal_destroy(sht_zonal *z) {
  if (z) zonal_release(z);
}

int sht_zonal_grid(const sht_zonal *z, double *lat, double *lat_weight) {
  if (!z) return fail(SHT_ERR_INVALID, "sht_zonal_grid: plan is NULL");
  for (int j = 0; j < z->jm; ++j) {
    if (lat) lat[j] = atan2(z->h_mu[j], z->h_coslat[j]);
    if (lat_weight) lat_weight[j] = z->h_w[j];
  }
  return SHT_OK;
}

int sht_l_y_dev(sht_zonal *z, const double *y, double *l, int nf, void *stream) {
  int rc = zonal_check(z, y, l, nf, "sht_l_y_dev");
  if (rc) return rc;
  return zonal_fwd(z, y, l, nf, (hipStream_t)stream);
}

int sht_y_l_dev(sht_zonal *z, const double *l, double *y, int nf, void *stream) {
  int rc = zonal_check(z, l, y, nf, "sht_y_l_dev");
  if (rc) return rc;
  return zonal_inv(z, l, y, nf, (hipStream_t)stream);
}

int sht_l_y(sht_zonal *z, const double *y, double *l, int nlev) {
  return zonal_host(z, y, l, nlev, 0, "sht_l_y");
}

int sht_y_l(sht_zonal *z, const double *l, double *y, int nlev) {
  return zonal_host(z, l, y, nlev, 1, "sht_y_l");
}

}  // extern "C"

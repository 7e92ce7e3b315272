// sht_remap.hip -- the coupler's grid remap on the GPU: the sparse
// interpolation of common/interpolation_data_latlon_mod.f90:274-306
// (interpolate_data_latlon), which Jcup applies to every exchanged field:
//   recv(:, d) = 0;  recv(r_i, d) += send(s_i, d) * coef_i,  i = 1..nnz
// for d = 1..ndata, fields contiguous per point (Fortran send_data(:, d)).
//
// The triplets are regrouped at creation by receiving point (CSR), keeping
// the original order within each row, and one thread sums one (row, field)
// with separately rounded multiply and add in that order: the same operations
// in the same order as the reference loop, so results are bitwise equal to an
// unfused evaluation of it (no atomics, no fused multiply-add).
#include <algorithm>
#include <new>
#include <numeric>

#include "sht_internal.h"

struct sht_remap {
  int nsend = 0, nrecv = 0, nnz = 0;
  int *row = nullptr;     // nrecv + 1 offsets
  int *col = nullptr;     // nnz send points, row-major
  double *val = nullptr;  // nnz coefficients
};

namespace sht {
namespace {

__global__ __launch_bounds__(256) void remap_kernel(const int *__restrict__ row,
                                                    const int *__restrict__ col,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ send,
                                                    double *__restrict__ recv, int nsend, int nrecv,
                                                    int ndata) {
  // separately rounded multiply and add: hipcc otherwise contracts the pair
  // into one fused multiply-add (the operations must be written here, in the
  // pragma's scope: __dmul_rn/__dadd_rn are contracted at their definition)
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrecv * ndata) return;
  const int r = (int)(t % nrecv), d = (int)(t / nrecv);
  const double *__restrict__ s = send + (size_t)d * nsend;
  double acc = 0.0;
  for (int i = row[r]; i < row[r + 1]; ++i) {
    const double prod = s[col[i]] * val[i];
    acc = acc + prod;
  }
  recv[(size_t)d * nrecv + r] = acc;
}

void remap_release(sht_remap *m) {
  if (m->row) (void)hipFree(m->row);
  if (m->col) (void)hipFree(m->col);
  if (m->val) (void)hipFree(m->val);
  delete m;
}

}  // namespace
}  // namespace sht

using namespace sht;

extern "C" {

int sht_remap_create(const int *send_index, const int *recv_index, const double *coef, int nnz,
                     int nsend, int nrecv, sht_remap **out) {
  if (!out) return fail(SHT_ERR_INVALID, "sht_remap_create: out is NULL");
  *out = nullptr;
  if (nnz < 0 || nsend < 1 || nrecv < 1 || (nnz > 0 && (!send_index || !recv_index || !coef)))
    return fail(SHT_ERR_INVALID, "sht_remap_create: bad sizes or NULL triplets");
  for (int i = 0; i < nnz; ++i)
    if (send_index[i] < 0 || send_index[i] >= nsend || recv_index[i] < 0 || recv_index[i] >= nrecv)
      return fail(SHT_ERR_INVALID, "sht_remap_create: index out of range at triplet " + std::to_string(i));
  // stable regrouping by receiving point keeps the reference's summation order
  std::vector<int> order(nnz);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return recv_index[a] < recv_index[b]; });
  std::vector<int> row(nrecv + 1, 0), col(nnz);
  std::vector<double> val(nnz);
  for (int i = 0; i < nnz; ++i) row[recv_index[i] + 1] += 1;
  for (int r = 0; r < nrecv; ++r) row[r + 1] += row[r];
  for (int i = 0; i < nnz; ++i) {
    col[i] = send_index[order[i]];
    val[i] = coef[order[i]];
  }
  sht_remap *m = new (std::nothrow) sht_remap;
  if (!m) return fail(SHT_ERR_NOMEM, "sht_remap_create: host allocation failed");
  m->nsend = nsend;
  m->nrecv = nrecv;
  m->nnz = nnz;
  const size_t nz = nnz > 0 ? (size_t)nnz : 1;
  if (hipMalloc(&m->row, sizeof(int) * (nrecv + 1)) != hipSuccess ||
      hipMalloc(&m->col, sizeof(int) * nz) != hipSuccess ||
      hipMalloc(&m->val, sizeof(double) * nz) != hipSuccess) {
    remap_release(m);
    return fail(SHT_ERR_NOMEM, "sht_remap_create: device allocation failed");
  }
  if (hipMemcpy(m->row, row.data(), sizeof(int) * (nrecv + 1), hipMemcpyHostToDevice) != hipSuccess ||
      (nnz > 0 && (hipMemcpy(m->col, col.data(), sizeof(int) * nnz, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(m->val, val.data(), sizeof(double) * nnz, hipMemcpyHostToDevice) != hipSuccess))) {
    remap_release(m);
    return fail(SHT_ERR_DEVICE, "sht_remap_create: upload failed");
  }
  *out = m;
  return SHT_OK;
}

void sht_remap_destroy(sht_remap *m) {
  if (m) remap_release(m);
}

int sht_remap_apply_dev(const sht_remap *m, const double *send, double *recv, int ndata, void *stream) {
  if (!m) return fail(SHT_ERR_INVALID, "sht_remap_apply_dev: remap is NULL");
  if (!send || !recv || ndata < 1) return fail(SHT_ERR_INVALID, "sht_remap_apply_dev: bad arguments");
  const int64_t n = (int64_t)m->nrecv * ndata;
  if (n > (int64_t)INT32_MAX * 128) return fail(SHT_ERR_INVALID, "sht_remap_apply_dev: too many outputs");
  remap_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      m->row, m->col, m->val, send, recv, m->nsend, m->nrecv, ndata);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // extern "C"

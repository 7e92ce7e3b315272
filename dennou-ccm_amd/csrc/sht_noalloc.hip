// sht_noalloc.hip -- spectral -> grid (xya_wa) through a static scratch, for
// callers that cannot allocate device memory at call time (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The Fourier intermediate between the Legendre and FFT kernels normally
// lives in the plan's scratch (allocated by sht_plan_create for kmax
// fields).  Here it lives in `scratch`, one file-scope device array of
// kScratchBytes per GPU, held until the process ends.  Its device address is
// looked up once per device with hipGetSymbolAddress (at the first
// sht_plan_create on that device, with that device current) and kept per
// device; the kernels get that address, never the array's host-side name.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (on different streams, or from different threads) race.  The kernels
// write every element of the intermediate they later read, so nothing needs
// clearing between calls.
#include <mutex>

#include "sht_internal.h"

namespace sht {
namespace {

constexpr size_t kScratchBytes = (size_t)32 << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

// Called by sht_plan_create with the plan's device current.
int noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(SHT_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return SHT_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(SHT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_ptr[device] = ptr;
  return SHT_OK;
}

int spec_to_grid_into(Plan &p, const double *spec, double *grid, int nf, void *four, hipStream_t s);

}  // namespace sht

using namespace sht;

extern "C" {

int sht_spec_to_grid_dev_noalloc(sht_plan *h, const double *spec, double *grid, int nf, void *stream) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_spec_to_grid_dev_noalloc: plan is NULL");
  if (!spec || !grid) return fail(SHT_ERR_INVALID, "sht_spec_to_grid_dev_noalloc: NULL array");
  Plan &p = h->p;
  if (nf < 1 || nf > p.kmax)
    return fail(SHT_ERR_INVALID, "sht_spec_to_grid_dev_noalloc: nfield must be in [1, kmax]");
  // the intermediate [P2][ceil(nf/G)][N+1][2][2G] in FP64, FP32 in the mixed forms
  // the largest layout any shape of the call can use (the fused inverse
  // always takes the big FFT blocks)
  const FourGeom g = four_geom_g(p.N, fft_group(p.im), nf);
  const size_t elem = p.prec == SHT_PREC_FP64 ? sizeof(double) : sizeof(float);
  const size_t need = (size_t)p.P2 * g.nfg * g.M1 * 2 * ((size_t)1 << g.lg) * elem;
  if (need > kScratchBytes)
    return fail(SHT_ERR_INVALID, "sht_spec_to_grid_dev_noalloc: " + std::to_string(need) +
                                     " bytes of Fourier intermediate exceed the 32 MiB static scratch");
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ptr = (p.device >= 0 && p.device < kMaxDevices) ? g_ptr[p.device] : nullptr;
  }
  if (!ptr) return fail(SHT_ERR_DEVICE, "sht_spec_to_grid_dev_noalloc: no scratch address for the plan's device");
  return spec_to_grid_into(p, spec, grid, nf, ptr, (hipStream_t)stream);
}

}  // extern "C"

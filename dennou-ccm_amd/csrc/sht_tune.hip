// sht_tune.hip -- Legendre kernel shapes per field count, and the autotuner
// that picks them by timing every shape on the plan's own device.
//
// The JSON cache is one object whose keys name the problem and whose values
// are the winning shapes:
//   {"T85_256x128_nf97_p0": [1, 1], ...}
// key = truncation, grid, field count and SHT_PREC_*; value = [inverse, forward].
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

#include "sht_internal.h"

namespace sht {
namespace {

// FP64 plans: every shape; mixed-precision plans: KSPLIT (shared LDS data
// tile of one column tile), WAVE (data tile per wave in registers) and TILE
// (shared LDS data tiles of two column tiles).
int nshapes(const Plan &p) { return p.prec == SHT_PREC_FP64 ? SHT_NSHAPE : 3; }

bool shape_ok(const Plan &p, int s) { return s == SHT_SHAPE_AUTO || (s >= 0 && s < nshapes(p)); }

void put_entry(Plan &p, int nf, int inv, int fwd) {
  for (Plan::ShapeEntry &e : p.shapes)
    if (e.nf == nf) {
      e.inv = inv;
      e.fwd = fwd;
      return;
    }
  p.shapes.push_back({nf, inv, fwd});
}

std::string cache_key(const Plan &p, int nf) {
  std::ostringstream k;
  k << "T" << p.N << "_" << p.im << "x" << p.jm << "_nf" << nf << "_p" << p.prec;
  return k.str();
}

// Reads "key": [a, b] pairs; anything else in the file is ignored.
std::map<std::string, std::pair<int, int>> read_cache(const char *path) {
  std::map<std::string, std::pair<int, int>> out;
  std::ifstream f(path);
  if (!f) return out;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string t = ss.str();
  size_t pos = 0;
  while ((pos = t.find('"', pos)) != std::string::npos) {
    const size_t end = t.find('"', pos + 1);
    if (end == std::string::npos) break;
    const std::string key = t.substr(pos + 1, end - pos - 1);
    int a = 0, b = 0;
    const size_t br = t.find('[', end);
    if (br != std::string::npos && sscanf(t.c_str() + br, "[%d , %d]", &a, &b) == 2) out[key] = {a, b};
    pos = end + 1;
  }
  return out;
}

int write_cache(const char *path, const std::map<std::string, std::pair<int, int>> &c) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return fail(SHT_ERR_INVALID, std::string("sht_plan_autotune: cannot write ") + tmp);
  fprintf(f, "{");
  bool first = true;
  for (const auto &kv : c) {
    fprintf(f, "%s\n  \"%s\": [%d, %d]", first ? "" : ",", kv.first.c_str(), kv.second.first,
            kv.second.second);
    first = false;
  }
  fprintf(f, "\n}\n");
  fclose(f);
  if (rename(tmp.c_str(), path) != 0)
    return fail(SHT_ERR_INVALID, std::string("sht_plan_autotune: cannot replace ") + path);
  return SHT_OK;
}

// Mean duration in microseconds of reps launches of one shape of one kernel.
template <class F>
int time_launches(hipStream_t s, int reps, F launch, double &us) {
  hipEvent_t a = nullptr, b = nullptr;
  SHT_HIP(hipEventCreate(&a));
  SHT_HIP(hipEventCreate(&b));
  int rc = launch();  // warm-up (code object load, caches)
  if (!rc) rc = launch();
  if (!rc && hipEventRecord(a, s) != hipSuccess) rc = fail(SHT_ERR_DEVICE, "autotune: hipEventRecord");
  for (int r = 0; r < reps && !rc; ++r) rc = launch();
  if (!rc && hipEventRecord(b, s) != hipSuccess) rc = fail(SHT_ERR_DEVICE, "autotune: hipEventRecord");
  if (!rc && hipEventSynchronize(b) != hipSuccess) rc = fail(SHT_ERR_DEVICE, "autotune: hipEventSynchronize");
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = fail(SHT_ERR_DEVICE, "autotune: hipEventElapsedTime");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  us = 1e3 * ms / reps;
  return rc;
}

}  // namespace
}  // namespace sht

using namespace sht;

extern "C" {

int sht_plan_set_shapes(sht_plan *h, int nf, int inv, int fwd) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_set_shapes: plan is NULL");
  Plan &p = h->p;
  if (nf < 0 || nf > p.kmax) return fail(SHT_ERR_INVALID, "sht_plan_set_shapes: nf out of range");
  if (!shape_ok(p, inv) || !shape_ok(p, fwd))
    return fail(SHT_ERR_UNSUPPORTED, "sht_plan_set_shapes: shape not available for this precision");
  if (nf == 0) {
    p.shape_inv = inv;
    p.shape_fwd = fwd;
  } else if (inv == SHT_SHAPE_AUTO && fwd == SHT_SHAPE_AUTO) {
    p.shapes.erase(std::remove_if(p.shapes.begin(), p.shapes.end(),
                                  [nf](const Plan::ShapeEntry &e) { return e.nf == nf; }),
                   p.shapes.end());
  } else {
    // an AUTO half stays AUTO: at call time it resolves to the plan default,
    // else to the built-in choice for nf fields (resolve_shape)
    put_entry(p, nf, inv, fwd);
  }
  return SHT_OK;
}

int sht_plan_get_shapes(const sht_plan *h, int nf, int *inv, int *fwd) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_get_shapes: plan is NULL");
  if (inv) *inv = resolve_shape(h->p, nf, 0);
  if (fwd) *fwd = resolve_shape(h->p, nf, 1);
  return SHT_OK;
}

int sht_plan_autotune(sht_plan *h, int nf, int reps, const char *cache_path, double *best_us) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_autotune: plan is NULL");
  Plan &p = h->p;
  if (nf < 1 || nf > p.kmax) return fail(SHT_ERR_INVALID, "sht_plan_autotune: nf out of range");
  if (reps < 1 || reps > 10000) return fail(SHT_ERR_INVALID, "sht_plan_autotune: reps out of range");
  if (best_us) best_us[0] = best_us[1] = 0.0;
  const std::string key = cache_key(p, nf);
  std::map<std::string, std::pair<int, int>> cache;
  if (cache_path) {
    cache = read_cache(cache_path);
    auto it = cache.find(key);
    if (it != cache.end() && shape_ok(p, it->second.first) && shape_ok(p, it->second.second) &&
        it->second.first >= 0 && it->second.second >= 0) {
      put_entry(p, nf, it->second.first, it->second.second);
      return SHT_OK;
    }
  }
  int best[2] = {0, 0};
  double tbest[2] = {0.0, 0.0};
  if (nshapes(p) > 1) {
    int cur = 0;
    SHT_HIP(hipGetDevice(&cur));
    SHT_HIP(hipSetDevice(p.device));
    hipStream_t s = nullptr;
    double *spec = nullptr, *grid = nullptr;
    int rc = SHT_OK;
    if (hipStreamCreate(&s) != hipSuccess) rc = fail(SHT_ERR_DEVICE, "autotune: hipStreamCreate");
    const size_t nspec = (size_t)p.K * 2 * nf, ngrid = (size_t)nf * p.jm * p.im;
    if (!rc && (hipMalloc(&spec, sizeof(double) * nspec) != hipSuccess ||
                hipMalloc(&grid, sizeof(double) * ngrid) != hipSuccess))
      rc = fail(SHT_ERR_NOMEM, "autotune: buffer allocation failed");
    if (!rc && (hipMemsetAsync(spec, 0, sizeof(double) * nspec, s) != hipSuccess ||
                hipMemsetAsync(grid, 0, sizeof(double) * ngrid, s) != hipSuccess))
      rc = fail(SHT_ERR_DEVICE, "autotune: hipMemsetAsync");
    // three interleaved rounds; the lowest per-round mean wins, so a drift in
    // clock or background load hits every shape alike
    for (int round = 0; round < 3 && !rc; ++round)
    for (int dir = 0; dir < 2 && !rc; ++dir)
      for (int shape = 0; shape < nshapes(p) && !rc; ++shape) {
        double us = 0.0;
        // a whole direction (Legendre and FFT) per launch, as a caller runs it:
        // the shapes differ in how they share the chip with the FFT's traffic
        const bool f32 = p.prec != SHT_PREC_FP64;
        float *f4 = reinterpret_cast<float *>(p.four);
        rc = time_launches(s, reps, [&] {
          int r;
          if (dir == 0) {
            r = f32 ? launch_leg_inv_mixed_shape(p, shape, spec, f4, nf, s)
                    : launch_leg_inv_f64_shape(p, shape, spec, p.four, p.four2, true, false, nf, s);
            if (!r) r = launch_fft_inv(p, p.four, nullptr, f32, grid, nullptr, nullptr, nf, s);
          } else {
            r = launch_fft_fwd(p, grid, p.four, f32, nf, s);
            if (!r) r = f32 ? launch_leg_fwd_mixed_shape(p, shape, f4, spec, nf, s)
                            : launch_leg_fwd_f64_shape(p, shape, p.four, spec, nf, s);
          }
          return r;
        }, us);
        if (!rc && ((round == 0 && shape == 0) || us < tbest[dir])) {
          tbest[dir] = us;
          best[dir] = shape;
        }
      }
    if (spec) (void)hipFree(spec);
    if (grid) (void)hipFree(grid);
    if (s) (void)hipStreamDestroy(s);
    (void)hipSetDevice(cur);
    if (rc) return rc;
  }
  put_entry(p, nf, best[0], best[1]);
  if (best_us) {
    best_us[0] = tbest[0];
    best_us[1] = tbest[1];
  }
  if (cache_path) {
    cache[key] = {best[0], best[1]};
    return write_cache(cache_path, cache);
  }
  return SHT_OK;
}

}  // extern "C"

// sht_tune.hip -- Legendre kernel shapes and wavenumber classes per field
// count, and the autotuner that picks them by timing on the plan's own device.
//
// The JSON cache is one object whose keys name the problem and whose values
// are the winning choice (host::parse_tune_cache / format_tune_cache):
//   {"T85_256x128_nf97_p0": [3, 4, 0, 19, -1, -1], ...}
// key = truncation, grid, field count and SHT_PREC_*; value = [inverse shape,
// forward shape, inverse m4, inverse m2, forward m4, forward m2], the m4/m2
// pairs being the wavenumber classes of the split shape (-1: built-in).
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "sht_internal.h"

namespace sht {
namespace {

// FP64 plans: every shape but FUSED; mixed-precision plans: KSPLIT (shared
// LDS data tile of one column tile), WAVE (data tile per wave in registers)
// and TILE (shared LDS data tiles of two column tiles), and for the inverse
// STAGED (table in registers, data streamed) and FUSED where the plan has
// them.
int nshapes(const Plan &p) { return p.prec == SHT_PREC_FP64 ? SHT_SHAPE_STAGED + 1 : 3; }

bool shape_ok(const Plan &p, int s, int dir) {
  if (s == SHT_SHAPE_FUSED) return dir == 0 && fused_inv_available(p);
  if (s == SHT_SHAPE_STAGED && p.prec != SHT_PREC_FP64) return dir == 0 && mixed_stream_available(p);
  return s == SHT_SHAPE_AUTO || (s >= 0 && s < nshapes(p));
}

host::Tuning *find_entry(Plan &p, int nf) {
  for (Plan::ShapeEntry &e : p.shapes)
    if (e.nf == nf) return &e.t;
  p.shapes.push_back({nf, host::Tuning{}});
  return &p.shapes.back().t;
}

std::string cache_key(const Plan &p, int nf) {
  std::ostringstream k;
  k << "T" << p.N << "_" << p.im << "x" << p.jm << "_nf" << nf << "_p" << p.prec;
  return k.str();
}

std::map<std::string, host::Tuning> read_cache(const char *path) {
  std::ifstream f(path);
  if (!f) return {};
  std::stringstream ss;
  ss << f.rdbuf();
  return host::parse_tune_cache(ss.str());
}

int write_cache(const char *path, const std::map<std::string, host::Tuning> &c) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return fail(SHT_ERR_INVALID, std::string("sht_plan_autotune: cannot write ") + tmp);
  const std::string text = host::format_tune_cache(c);
  const bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
  if (fclose(f) != 0 || !ok)
    return fail(SHT_ERR_INVALID, std::string("sht_plan_autotune: cannot write ") + tmp);
  if (rename(tmp.c_str(), path) != 0)
    return fail(SHT_ERR_INVALID, std::string("sht_plan_autotune: cannot replace ") + path);
  return SHT_OK;
}

bool classes_ok(const Plan &p, int m4, int m2) {
  return (m4 == -1 && m2 == -1) || (m4 >= 0 && m2 >= m4 && m2 <= p.N + 1);
}

// Mean duration in microseconds of reps launches of one candidate.
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

// A candidate of one direction: a shape, and for the split shape its
// wavenumber classes.
struct Candidate {
  int shape;
  KSplit cls;
};

// Every shape with the built-in classes, and for FP64 plans the split shape
// with other class boundaries: the two-way class ending after 20 % to 70 % of
// the wavenumbers (heaviest first), with no four-way class or one over its
// first quarter or half.  The work per m falls with m, so the boundaries are
// counts of the heaviest wavenumbers.
std::vector<Candidate> candidates(const Plan &p, int dir, bool classes) {
  std::vector<Candidate> c;
  for (int s = 0; s < nshapes(p); ++s) c.push_back({s, KSplit{-1, -1}});
  if (dir == 0 && p.prec != SHT_PREC_FP64 && mixed_stream_available(p)) c.push_back({SHT_SHAPE_STAGED, KSplit{-1, -1}});
  if (dir == 0 && fused_inv_available(p)) c.push_back({SHT_SHAPE_FUSED, KSplit{-1, -1}});
  if (!classes || p.prec != SHT_PREC_FP64 || p.N < 8) return c;
  const KSplit def = default_classes(p, dir);
  for (int pct = 20; pct <= 70; pct += 10) {
    const int m2 = std::max(1, (p.N + 1) * pct / 100);
    for (int q = 0; q < 3; ++q) {
      const int m4 = q == 0 ? 0 : m2 * q / 4;
      if (m4 == def.m4 && m2 == def.m2) continue;
      bool dup = false;
      for (const Candidate &x : c) dup |= x.shape == SHT_SHAPE_TILE_SPLIT && x.cls.m4 == m4 && x.cls.m2 == m2;
      if (!dup) c.push_back({SHT_SHAPE_TILE_SPLIT, KSplit{m4, m2}});
    }
  }
  return c;
}

}  // namespace
}  // namespace sht

using namespace sht;

extern "C" {

int sht_plan_set_shapes(sht_plan *h, int nf, int inv, int fwd) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_set_shapes: plan is NULL");
  Plan &p = h->p;
  if (nf < 0 || nf > p.kmax) return fail(SHT_ERR_INVALID, "sht_plan_set_shapes: nf out of range");
  if (!shape_ok(p, inv, 0) || !shape_ok(p, fwd, 1))
    return fail(SHT_ERR_UNSUPPORTED, "sht_plan_set_shapes: shape not available for this precision");
  if (nf == 0) {
    p.shape_inv = inv;
    p.shape_fwd = fwd;
    return SHT_OK;
  }
  host::Tuning *t = find_entry(p, nf);
  // an AUTO half stays AUTO: at call time it resolves to the plan default,
  // else to the built-in choice for nf fields (resolve_shape)
  t->inv = inv;
  t->fwd = fwd;
  if (inv == SHT_SHAPE_AUTO && fwd == SHT_SHAPE_AUTO && t->inv_m2 < 0 && t->fwd_m2 < 0)
    p.shapes.erase(std::remove_if(p.shapes.begin(), p.shapes.end(),
                                  [nf](const Plan::ShapeEntry &e) { return e.nf == nf; }),
                   p.shapes.end());
  return SHT_OK;
}

int sht_plan_get_shapes(const sht_plan *h, int nf, int *inv, int *fwd) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_get_shapes: plan is NULL");
  if (inv) *inv = resolve_shape(h->p, nf, 0);
  if (fwd) *fwd = resolve_shape(h->p, nf, 1);
  return SHT_OK;
}

int sht_plan_set_classes(sht_plan *h, int nf, int dir, int m4, int m2) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_set_classes: plan is NULL");
  Plan &p = h->p;
  if (nf < 1 || nf > p.kmax) return fail(SHT_ERR_INVALID, "sht_plan_set_classes: nf out of range");
  if (dir != SHT_DIR_INVERSE && dir != SHT_DIR_FORWARD)
    return fail(SHT_ERR_INVALID, "sht_plan_set_classes: dir must be INVERSE or FORWARD");
  if (!classes_ok(p, m4, m2))
    return fail(SHT_ERR_INVALID, "sht_plan_set_classes: need 0 <= m4 <= m2 <= nmax + 1, or -1, -1");
  host::Tuning *t = find_entry(p, nf);
  (dir == SHT_DIR_INVERSE ? t->inv_m4 : t->fwd_m4) = m4;
  (dir == SHT_DIR_INVERSE ? t->inv_m2 : t->fwd_m2) = m2;
  return SHT_OK;
}

int sht_plan_get_classes(const sht_plan *h, int nf, int dir, int *m4, int *m2) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_get_classes: plan is NULL");
  if (dir != SHT_DIR_INVERSE && dir != SHT_DIR_FORWARD)
    return fail(SHT_ERR_INVALID, "sht_plan_get_classes: dir must be INVERSE or FORWARD");
  const KSplit k = resolve_classes(h->p, nf, dir == SHT_DIR_INVERSE ? 0 : 1);
  if (m4) *m4 = k.m4;
  if (m2) *m2 = k.m2;
  return SHT_OK;
}

int sht_plan_autotune(sht_plan *h, int nf, int reps, const char *cache_path, double *best_us) {
  return sht_plan_autotune_flags(h, nf, reps, cache_path, 0, best_us);
}

int sht_plan_autotune_flags(sht_plan *h, int nf, int reps, const char *cache_path, int flags,
                            double *best_us) {
  if (!h) return fail(SHT_ERR_INVALID, "sht_plan_autotune: plan is NULL");
  Plan &p = h->p;
  if (nf < 1 || nf > p.kmax) return fail(SHT_ERR_INVALID, "sht_plan_autotune: nf out of range");
  if (reps < 1 || reps > 10000) return fail(SHT_ERR_INVALID, "sht_plan_autotune: reps out of range");
  if (flags & ~SHT_TUNE_SHAPES_ONLY) return fail(SHT_ERR_INVALID, "sht_plan_autotune: unknown flags");
  const bool classes = !(flags & SHT_TUNE_SHAPES_ONLY);
  if (best_us) best_us[0] = best_us[1] = 0.0;
  // shapes-only choices are cached under a key of their own
  const std::string key = cache_key(p, nf) + (classes ? "" : "_shapes");
  std::map<std::string, host::Tuning> cache;
  if (cache_path) {
    cache = read_cache(cache_path);
    auto it = cache.find(key);
    if (it != cache.end()) {
      const host::Tuning &c = it->second;
      if (shape_ok(p, c.inv, 0) && shape_ok(p, c.fwd, 1) && c.inv >= 0 && c.fwd >= 0 &&
          classes_ok(p, c.inv_m4, c.inv_m2) && classes_ok(p, c.fwd_m4, c.fwd_m2)) {
        *find_entry(p, nf) = c;
        return SHT_OK;
      }
    }
  }
  Candidate best[2] = {{0, KSplit{-1, -1}}, {0, KSplit{-1, -1}}};
  double tbest[2] = {0.0, 0.0};
  const std::vector<Candidate> cand[2] = {candidates(p, 0, classes), candidates(p, 1, classes)};
  if (cand[0].size() > 1 || cand[1].size() > 1) {
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
    // clock or background load hits every candidate alike
    for (int round = 0; round < 3 && !rc; ++round)
      for (int dir = 0; dir < 2 && !rc; ++dir)
        for (size_t ci = 0; ci < cand[dir].size() && !rc; ++ci) {
          const Candidate c = cand[dir][ci];
          double us = 0.0;
          // a whole direction (Legendre and FFT) per launch, as a caller runs
          // it: the shapes differ in how they share the chip with the FFT
          const bool f32 = p.prec != SHT_PREC_FP64;
          float *f4 = reinterpret_cast<float *>(p.four);
          rc = time_launches(s, reps, [&] {
            int r;
            if (dir == 0 && c.shape == SHT_SHAPE_FUSED) {
              r = launch_inv_fused(p, spec, f4, grid, nf, s);
            } else if (dir == 0) {
              r = f32 ? launch_leg_inv_mixed_shape(p, c.shape, spec, f4, nf, s)
                      : launch_leg_inv_f64_shape(p, c.shape, spec, p.four, p.four2, true, false, nf, s, c.cls);
              if (!r) r = launch_fft_inv(p, p.four, nullptr, f32, grid, nullptr, nullptr, nf, s);
            } else {
              r = launch_fft_fwd(p, grid, p.four, f32, nf, s);
              if (!r) r = f32 ? launch_leg_fwd_mixed_shape(p, c.shape, f4, spec, nf, s)
                              : launch_leg_fwd_f64_shape(p, c.shape, p.four, spec, nf, s, c.cls);
            }
            return r;
          }, us);
          if (!rc && ((round == 0 && ci == 0) || us < tbest[dir])) {
            tbest[dir] = us;
            best[dir] = c;
          }
        }
    if (spec) (void)hipFree(spec);
    if (grid) (void)hipFree(grid);
    if (s) (void)hipStreamDestroy(s);
    (void)hipSetDevice(cur);
    if (rc) return rc;
  }
  host::Tuning t;
  t.inv = best[0].shape;
  t.fwd = best[1].shape;
  t.inv_m4 = best[0].cls.m4;
  t.inv_m2 = best[0].cls.m2;
  t.fwd_m4 = best[1].cls.m4;
  t.fwd_m2 = best[1].cls.m2;
  *find_entry(p, nf) = t;
  if (best_us) {
    best_us[0] = tbest[0];
    best_us[1] = tbest[1];
  }
  if (cache_path) {
    cache[key] = t;
    return write_cache(cache_path, cache);
  }
  return SHT_OK;
}

}  // extern "C"

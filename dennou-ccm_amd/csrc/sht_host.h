// sht_host.h -- the host-only logic of libsht_hip.so, free of HIP so that it
// builds with any C++17 compiler: descriptor checks, SPML index helpers, the
// Legendre shape cache format, gmapgen's latitude cells and exchange grid, and
// the request server's admission and batching queue.  The HIP translation
// units call these; tests/host/ builds them alone under AddressSanitizer and
// UndefinedBehaviorSanitizer on the CPU.
#pragma once

#include <cstddef>
#include <deque>
#include <map>
#include <string>
#include <vector>

#pragma GCC visibility push(default)
#include "../../include/sht.h"
#pragma GCC visibility pop

namespace sht {
namespace host {

// ---- the calling thread's last error text (sht_last_error)
void set_error(const std::string &msg);
const char *last_error();

// ---- descriptors: SHT_OK, or an SHT_ERR_* code with the reason in msg
int check_plan_desc(const sht_desc *d, std::string &msg);
int check_zonal_desc(int nmax, int jm, int kmax, std::string &msg);

// ---- SPML spectral index l = n(n+1) + m + 1 (1-based); -1 when invalid
int l_of_nm(int n, int m);
bool nm_of_l(int l, int &n, int &m);

// ---- Legendre kernel choice for one (plan, field count).  Shapes are
// SHT_SHAPE_*; m4/m2 are the wavenumber class boundaries of the split shape
// (K loop split four ways for m < m4, two ways for m4 <= m < m2, not for
// m >= m2); -1 = the built-in classes.
struct Tuning {
  int inv = SHT_SHAPE_AUTO, fwd = SHT_SHAPE_AUTO;
  int inv_m4 = -1, inv_m2 = -1, fwd_m4 = -1, fwd_m2 = -1;
};
// The cache file is one JSON object, {"<key>": [inv, fwd, inv_m4, inv_m2,
// fwd_m4, fwd_m2], ...}; two-element entries (older caches) leave the classes
// at -1.  Anything that does not parse as such an entry is skipped.
std::map<std::string, Tuning> parse_tune_cache(const std::string &text);
std::string format_tune_cache(const std::map<std::string, Tuning> &cache);

// ---- latitude cells of a Gaussian grid (gmapgen_main.f90:364-368):
// edges[0] = -pi/2, edges[j] = asin(w[j-1] + sin(edges[j-1])), edges[jm] = pi/2
void cell_edges(const double *w, int jm, double *edges);
// gmapgen's surface exchange grid (gmapgen_main.f90:336-387); lat and w hold
// jm_a + jm_b - 1 values; returns the count
int exchange_grid(const double *lat_a, const double *w_a, int jm_a, const double *w_b, int jm_b,
                  double *lat, double *w);

// ---- request admission and batching (not thread-safe: the server holds its
// lock around every call)
struct Pending {
  long long ticket = 0;
  int grid = 0, nf = 0, dir = 0;
  const double *in = nullptr;
  double *out = nullptr;
  size_t bytes = 0;
  double t_submit = 0.0;  // caller's clock, passed through
};

class Admission {
 public:
  Admission(size_t budget, int max_fields, int max_requests)
      : budget_(budget), max_fields_(max_fields), max_requests_(max_requests) {}
  // SHT_OK with *ticket set, or SHT_ERR_BUSY when the request's bytes would
  // exceed the budget (counted as refused)
  int submit(int grid, int nf, int dir, const double *in, double *out, size_t bytes, double t_submit,
             long long *ticket);
  bool empty() const { return pending_.empty(); }
  size_t pending() const { return pending_.size(); }
  // The pending requests sharing the oldest one's (grid, direction), in
  // arrival order, up to max_fields fields and max_requests requests; the
  // requests stay admitted (their bytes held) until finish().
  std::vector<Pending> take_batch();
  // A batch ended: releases its bytes; ok = false counts its requests failed.
  void finish(const std::vector<Pending> &batch, bool ok);
  size_t bytes_used() const { return used_; }
  long long refused() const { return refused_; }
  long long batches() const { return batches_; }
  long long completed() const { return completed_; }
  long long completed_fields() const { return fields_; }
  long long failed() const { return failed_; }

 private:
  size_t budget_, used_ = 0;
  int max_fields_, max_requests_;
  long long next_ticket_ = 1, refused_ = 0, batches_ = 0, completed_ = 0, fields_ = 0, failed_ = 0;
  std::deque<Pending> pending_;
};

// ---- restart state (sht_restart_write / sht_restart_read): the ten grid
// fields of the two leapfrog time levels in a netCDF classic file
struct RestartDims {
  int im, jm, km, nq;
};
int restart_write(const char *path, const RestartDims &d, double time, const double *lon, const double *lat,
                  const double *sig, const double *const *fields, std::string &msg);
int restart_read(const char *path, const RestartDims &d, double *time, double *const *fields, std::string &msg);

// ---- copies between a caller's host arrays and pinned staging memory (the
// host drop-in entry points): the calling thread and up to kCopyWorkers
// persistent worker threads take parts of the range from a shared counter
// (a worker the scheduler has not run yet leaves its parts to the others;
// the caller never waits for it).  One such copy runs in parallel at a
// time; a caller that finds the workers busy, or a copy
// below kParallelCopyMin bytes, copies alone, and so does a forked child
// process (the workers are not forked with it).  The workers live until the
// process ends.  Measured on the MI355X host (16 cores, 25.4 MB grid in 8
// chunks, each copied out of pinned memory while the next chunk's PCIe
// transfer runs): 0.93 ms with one thread, 0.62 ms with four; into pinned
// memory the gain depends on the machine (none on one, 1.19 -> 0.80 ms for
// the whole call on another).
constexpr int kCopyWorkers = 3;
constexpr size_t kParallelCopyMin = (size_t)1 << 20;
void parallel_copy(void *dst, const void *src, size_t bytes);

// ---- test hook: the value of environment variable `name` as a count
// (0 when unset or not a positive integer).  SHT_INJECT_PLAN_FAILURE=k makes
// the k-th device allocation of each sht_plan_create fail (SHT_ERR_NOMEM);
// SHT_INJECT_BATCH_FAILURE=k makes the k-th batch of each request server fail
// before it is enqueued.  For tests of the error paths; unset in production.
int injected_count(const char *name);

}  // namespace host
}  // namespace sht

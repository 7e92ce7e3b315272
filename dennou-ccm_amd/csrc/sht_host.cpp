// sht_host.cpp -- host-only logic of libsht_hip.so (see sht_host.h).
#include "sht_host.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

#include <unistd.h>

namespace sht {
namespace host {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
const char *last_error() { return g_error.c_str(); }

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int check_plan_desc(const sht_desc *d, std::string &msg) {
  auto bad = [&](int code, const char *why) {
    msg = std::string("sht_plan_create: ") + why;
    return code;
  };
  if (!d) return bad(SHT_ERR_INVALID, "desc is NULL");
  if (d->nmax < 0 || d->nmax > 255) return bad(SHT_ERR_INVALID, "nmax must be in [0, 255]");
  if (d->jm < 2 || (d->jm & 1) || d->jm > 512) return bad(SHT_ERR_INVALID, "jm must be even and in [2, 512]");
  if (d->jm <= d->nmax) return bad(SHT_ERR_INVALID, "jm must exceed nmax (exact quadrature)");
  if (d->im < 2 * d->nmax + 2) return bad(SHT_ERR_INVALID, "im must be >= 2*nmax + 2");
  if (!is_pow2(d->im) || d->im < 16 || d->im > 2048)
    return bad(SHT_ERR_UNSUPPORTED, "im must be a power of two in [16, 2048]");
  if (d->kmax < 1) return bad(SHT_ERR_INVALID, "kmax must be >= 1");
  if (d->precision < SHT_PREC_FP64 || d->precision > SHT_PREC_SPLIT_BF16X6)
    return bad(SHT_ERR_INVALID, "unknown precision");
  if (d->flags & ~SHT_FLAG_GRAD) return bad(SHT_ERR_INVALID, "unknown flags");
  // the Fourier scratch, [jm/2][nmax+1][2][2 kmax] doubles, must stay below 2^34
  const size_t four = (size_t)(d->jm / 2) * (size_t)(d->nmax + 1) * 4 * (size_t)d->kmax;
  if (four > ((size_t)1 << 34)) return bad(SHT_ERR_INVALID, "kmax too large for the Fourier scratch");
  return SHT_OK;
}

int check_zonal_desc(int nmax, int jm, int kmax, std::string &msg) {
  auto bad = [&](const char *why) {
    msg = std::string("sht_zonal_create: ") + why;
    return SHT_ERR_INVALID;
  };
  if (jm < 2 || (jm & 1) || jm > 8192) return bad("jm must be even and in [2, 8192]");
  if (nmax < 0 || nmax > jm) return bad("nmax must be in [0, jm]");
  const size_t per = (size_t)std::max(jm, nmax + 1);
  if (kmax < 1 || (size_t)kmax * per > ((size_t)1 << 31)) return bad("kmax out of range");
  return SHT_OK;
}

int l_of_nm(int n, int m) {
  if (n < 0 || m < -n || m > n) return -1;
  if (n > 46339) return -1;  // (n + 1)^2 must fit an int
  return n * (n + 1) + m + 1;
}

bool nm_of_l(int l, int &n, int &m) {
  if (l < 1) return false;
  long long nn = (long long)std::sqrt((double)(l - 1));  // floor(sqrt(l - 1)), corrected below
  while (nn > 0 && nn * nn > (long long)l - 1) --nn;
  while ((nn + 1) * (nn + 1) <= (long long)l - 1) ++nn;
  n = (int)nn;
  m = (int)((long long)l - 1 - nn * (nn + 1));
  return true;
}

// ---- shape cache
namespace {

// Parses "[a, b, ...]" of at most six ints starting at t[pos] (== '[');
// returns the count, or -1.
int parse_int_list(const std::string &t, size_t pos, int (&v)[6]) {
  size_t i = pos + 1;
  int n = 0;
  for (;;) {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\t' || t[i] == '\r')) ++i;
    if (i >= t.size()) return -1;
    if (t[i] == ']') return n;
    if (n == 6) return -1;
    const char *s = t.c_str() + i;
    char *end = nullptr;
    errno = 0;
    const long x = std::strtol(s, &end, 10);
    if (end == s || errno == ERANGE || x < INT_MIN || x > INT_MAX) return -1;
    v[n++] = (int)x;
    i += (size_t)(end - s);
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\t' || t[i] == '\r')) ++i;
    if (i < t.size() && t[i] == ',') ++i;
    else if (i < t.size() && t[i] == ']') return n;
    else return -1;
  }
}

}  // namespace

std::map<std::string, Tuning> parse_tune_cache(const std::string &t) {
  std::map<std::string, Tuning> out;
  size_t pos = 0;
  while ((pos = t.find('"', pos)) != std::string::npos) {
    const size_t end = t.find('"', pos + 1);
    if (end == std::string::npos) break;
    const std::string key = t.substr(pos + 1, end - pos - 1);
    size_t c = end + 1;
    while (c < t.size() && (t[c] == ' ' || t[c] == ':' || t[c] == '\n' || t[c] == '\t')) ++c;
    int v[6];
    const int n = (c < t.size() && t[c] == '[') ? parse_int_list(t, c, v) : -1;
    if (n == 2 || n == 6) {
      Tuning r;
      r.inv = v[0];
      r.fwd = v[1];
      if (n == 6) {
        r.inv_m4 = v[2];
        r.inv_m2 = v[3];
        r.fwd_m4 = v[4];
        r.fwd_m2 = v[5];
      }
      out[key] = r;
    }
    pos = end + 1;
  }
  return out;
}

std::string format_tune_cache(const std::map<std::string, Tuning> &c) {
  std::string s = "{";
  bool first = true;
  char buf[160];
  for (const auto &kv : c) {
    const Tuning &r = kv.second;
    std::snprintf(buf, sizeof buf, "[%d, %d, %d, %d, %d, %d]", r.inv, r.fwd, r.inv_m4, r.inv_m2, r.fwd_m4,
                  r.fwd_m2);
    s += first ? "\n  \"" : ",\n  \"";
    s += kv.first;
    s += "\": ";
    s += buf;
    first = false;
  }
  s += "\n}\n";
  return s;
}

// ---- gmapgen cells: the recurrence in the order the generator evaluates
// it, so a rounding that accumulates along the sum is the generator's own
void cell_edges(const double *w, int jm, double *edges) {
  edges[0] = -M_PI / 2.0;
  for (int j = 1; j < jm; ++j) edges[j] = std::asin(w[j - 1] + std::sin(edges[j - 1]));
  edges[jm] = M_PI / 2.0;
}

int exchange_grid(const double *lat_a, const double *w_a, int jm_a, const double *w_b, int jm_b,
                  double *lat, double *w) {
  if (jm_a == jm_b) {  // gmapgen_main.f90:349-355: the exchange grid is grid A
    std::copy(lat_a, lat_a + jm_a, lat);
    std::copy(w_a, w_a + jm_a, w);
    return jm_a;
  }
  std::vector<double> ea(jm_a + 1), eb(jm_b + 1);
  cell_edges(w_a, jm_a, ea.data());
  cell_edges(w_b, jm_b, eb.data());
  // A's edges 0..jm_a-1 and B's 1..jm_b (gmapgen_main.f90:372-373)
  std::vector<double> e(ea.begin(), ea.end() - 1);
  e.insert(e.end(), eb.begin() + 1, eb.end());
  std::sort(e.begin(), e.end());
  int n = 0;
  for (size_t j = 1; j < e.size(); ++j) {
    const double area = std::sin(e[j]) - std::sin(e[j - 1]);
    if (std::fabs(area) > 1e-12) {
      lat[n] = 0.5 * (e[j - 1] + e[j]);
      w[n] = area;
      ++n;
    }
  }
  return n;
}

// ---- admission
int Admission::submit(int grid, int nf, int dir, const double *in, double *out, size_t bytes,
                      double t_submit, long long *ticket) {
  if (bytes > budget_ || used_ > budget_ - bytes) {
    ++refused_;
    return SHT_ERR_BUSY;
  }
  used_ += bytes;
  Pending q;
  q.ticket = next_ticket_++;
  q.grid = grid;
  q.nf = nf;
  q.dir = dir;
  q.in = in;
  q.out = out;
  q.bytes = bytes;
  q.t_submit = t_submit;
  pending_.push_back(q);
  *ticket = q.ticket;
  return SHT_OK;
}

std::vector<Pending> Admission::take_batch() {
  std::vector<Pending> batch;
  if (pending_.empty()) return batch;
  const int g = pending_.front().grid, dir = pending_.front().dir;
  int nf = 0;
  for (auto it = pending_.begin(); it != pending_.end();) {
    // the oldest request always goes (submit keeps nf within max_fields)
    if (batch.empty() || (it->grid == g && it->dir == dir && nf + it->nf <= max_fields_ &&
                          (int)batch.size() < max_requests_)) {
      nf += it->nf;
      batch.push_back(*it);
      it = pending_.erase(it);
    } else {
      ++it;
    }
  }
  if (!batch.empty()) ++batches_;
  return batch;
}

void Admission::finish(const std::vector<Pending> &batch, bool ok) {
  for (const Pending &r : batch) {
    used_ -= r.bytes;
    if (ok) {
      ++completed_;
      fields_ += r.nf;
    } else {
      ++failed_;
    }
  }
}

int injected_count(const char *name) {
  const char *v = std::getenv(name);
  if (!v || !*v) return 0;
  char *end = nullptr;
  const long x = std::strtol(v, &end, 10);
  if (*end != '\0' || x < 1 || x > INT_MAX) return 0;
  return (int)x;
}

// Workers of parallel_copy.  A copy is cut into parts that the caller and
// every worker that joins take from a shared counter, so a worker that is
// slow to be scheduled (the host's cores busy with the caller's own threads)
// only leaves more parts to the others: the caller never waits for a worker
// that has not started, only for parts already taken.  A worker joins a job
// while it is open; the caller closes it once no part is left and waits for
// the workers inside it to leave.
namespace {
struct CopyPool {
  pid_t pid = 0;  // the process that started the workers (a forked child has none)
  std::mutex m;
  std::condition_variable start, done;
  std::mutex busy;  // held by the caller of the running parallel copy
  int nworkers = 0;
  unsigned gen = 0;
  bool open = false;
  int active = 0;  // workers inside the open job
  char *dst = nullptr;
  const char *src = nullptr;
  size_t bytes = 0;
  int nparts = 1;
  std::atomic<int> next{0};

  void run_parts() {
    for (;;) {
      const int i = next.fetch_add(1, std::memory_order_relaxed);
      if (i >= nparts) return;
      const size_t a = bytes * i / nparts, b = bytes * (i + 1) / nparts;
      std::memcpy(dst + a, src + a, b - a);
    }
  }
  void worker() {
    unsigned seen = 0;
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      start.wait(lk, [&] { return gen != seen; });
      seen = gen;
      if (!open) continue;
      ++active;
      lk.unlock();
      run_parts();
      lk.lock();
      if (--active == 0) done.notify_all();
    }
  }
};

CopyPool *copy_pool() {
  // never freed: the detached workers wait on it until the process ends
  static CopyPool *pool = [] {
    auto *q = new CopyPool;
    q->pid = getpid();
    const unsigned hw = std::thread::hardware_concurrency();
    const int want = hw > 1 ? (int)std::min<unsigned>(kCopyWorkers, hw - 1) : 0;
    // a worker that cannot be started (thread limits) leaves fewer workers;
    // none at all leaves plain copies
    for (int i = 0; i < want; ++i) {
      try {
        std::thread(&CopyPool::worker, q).detach();
      } catch (...) {
        break;
      }
      q->nworkers = i + 1;
    }
    return q;
  }();
  return pool;
}
}  // namespace

void parallel_copy(void *dst, const void *src, size_t bytes) {
  CopyPool *p = bytes >= kParallelCopyMin ? copy_pool() : nullptr;
  if (!p || p->nworkers == 0 || p->pid != getpid() || !p->busy.try_lock()) {
    std::memcpy(dst, src, bytes);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(p->m);
    p->dst = static_cast<char *>(dst);
    p->src = static_cast<const char *>(src);
    p->bytes = bytes;
    p->nparts = 4 * (p->nworkers + 1);
    p->next.store(0, std::memory_order_relaxed);
    p->open = true;
    ++p->gen;
  }
  p->start.notify_all();
  p->run_parts();
  {
    std::unique_lock<std::mutex> lk(p->m);
    p->open = false;
    p->done.wait(lk, [&] { return p->active == 0; });
  }
  p->busy.unlock();
}

}  // namespace host
}  // namespace sht

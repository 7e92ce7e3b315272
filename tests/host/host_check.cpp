// CPU checks of libsht_hip.so's host-only logic (dennou-ccm_amd/csrc/sht_host.cpp),
// built alone with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitized.py: descriptor validation over a grid of good and
// bad values, the SPML index helpers at their int limits, the shape-cache
// parser on well-formed, truncated and random text, gmapgen's exchange grid
// on random weight sets, and the request admission queue driven by several
// submitting threads against one consumer that forms batches, fails some of
// them on purpose and finishes the rest.  Prints PASS and exits 0.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include <sys/wait.h>
#include <unistd.h>

#include "sht_host.h"

using namespace sht::host;

static int nbad = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("check failed at line %d: %s\n", __LINE__, #cond); \
      ++nbad;                                                         \
    }                                                                 \
  } while (0)

static void descriptors() {
  std::string msg;
  CHECK(check_plan_desc(nullptr, msg) == SHT_ERR_INVALID && !msg.empty());
  const int vals[] = {-2147483647 - 1, -1, 0, 1, 2, 15, 16, 21, 32, 64, 85, 128, 255, 256, 511, 512,
                      513, 2048, 4096, 2147483647};
  long good = 0;
  for (int n : vals)
    for (int im : vals)
      for (int jm : vals)
        for (int k : {0, 1, 97, 2147483647}) {
          sht_desc d{n, im, jm, k, 0, 0};
          msg.clear();
          const int rc = check_plan_desc(&d, msg);
          CHECK(rc == SHT_OK || !msg.empty());
          if (rc == SHT_OK) {
            ++good;
            CHECK(n >= 0 && n <= 255 && jm > n && jm % 2 == 0 && im >= 2 * n + 2 && k >= 1);
          }
        }
  CHECK(good > 0);
  sht_desc d{21, 64, 32, 49, 7, 0};
  CHECK(check_plan_desc(&d, msg) == SHT_ERR_INVALID);
  d.precision = 0;
  d.flags = 8;
  CHECK(check_plan_desc(&d, msg) == SHT_ERR_INVALID);
  CHECK(check_zonal_desc(1024, 1024, 1, msg) == SHT_OK);
  CHECK(check_zonal_desc(1025, 1024, 1, msg) == SHT_ERR_INVALID);
  CHECK(check_zonal_desc(10, 1024, 2147483647, msg) == SHT_ERR_INVALID);
}

static void indices() {
  for (int n = 0; n < 300; ++n)
    for (int m = -n; m <= n; ++m) {
      const int l = l_of_nm(n, m);
      int n2 = -1, m2 = -1;
      CHECK(l >= 1 && nm_of_l(l, n2, m2) && n2 == n && m2 == m);
    }
  CHECK(l_of_nm(46339, 46339) == 46340 * 46340);
  CHECK(l_of_nm(46340, 0) == -1 && l_of_nm(3, 4) == -1 && l_of_nm(-1, 0) == -1);
  for (int l : {1, 2, 46340 * 46340, 46340 * 46340 + 1, 2147483647}) {
    int n = 0, m = 0;
    CHECK(nm_of_l(l, n, m) && (long long)n * (n + 1) + m + 1 == l && m >= -n && m <= n);
  }
  int n = 0, m = 0;
  CHECK(!nm_of_l(0, n, m) && !nm_of_l(-5, n, m));
}

static void tune_cache() {
  std::map<std::string, Tuning> c;
  Tuning a;
  a.inv = 3; a.fwd = 4; a.inv_m4 = 0; a.inv_m2 = 19; a.fwd_m4 = -1; a.fwd_m2 = -1;
  c["T85_256x128_nf97_p0"] = a;
  Tuning b;
  b.inv = 1; b.fwd = 0;
  c["T21_64x32_nf5_p3"] = b;
  const std::string text = format_tune_cache(c);
  const auto back = parse_tune_cache(text);
  CHECK(back.size() == 2);
  const Tuning &r = back.at("T85_256x128_nf97_p0");
  CHECK(r.inv == 3 && r.fwd == 4 && r.inv_m4 == 0 && r.inv_m2 == 19 && r.fwd_m4 == -1 && r.fwd_m2 == -1);
  // two-element entries from older caches
  const auto old = parse_tune_cache("{\"T42_128x64_nf9_p0\": [1, 4]}");
  CHECK(old.size() == 1 && old.at("T42_128x64_nf9_p0").inv == 1 && old.at("T42_128x64_nf9_p0").inv_m2 == -1);
  // malformed text: never crashes, keeps only well-formed entries
  const char *bad[] = {"", "{", "\"", "\"k\"", "\"k\": [", "\"k\": [1,", "\"k\": [1, 2, 3]",
                       "\"k\": [99999999999999999999, 1]", "\"k\": [1 2]", "\"k\": [1, 2, 3, 4, 5, 6, 7]",
                       "\"a\": [1, 2] \"b\": [x]", "\"\": [0, 0]"};
  for (const char *t : bad) (void)parse_tune_cache(t);
  CHECK(parse_tune_cache("\"k\": [1, 2, 3]").empty());
  CHECK(parse_tune_cache("\"\": [0, 0]").size() == 1);
  // every prefix of a valid file, and random byte strings
  for (size_t i = 0; i <= text.size(); ++i) (void)parse_tune_cache(text.substr(0, i));
  std::mt19937 rng(7);
  const std::string alphabet = "\"[],:{} 0123456789-\nTkx_";
  for (int trial = 0; trial < 20000; ++trial) {
    std::string t(rng() % 64, ' ');
    for (char &ch : t) ch = alphabet[rng() % alphabet.size()];
    (void)parse_tune_cache(t);
  }
}

static void gauss_like_weights(int jm, std::mt19937 &rng, std::vector<double> &lat, std::vector<double> &w) {
  // positive weights summing to 2 and the midpoints of their cells
  std::uniform_real_distribution<double> u(0.2, 1.0);
  w.resize(jm);
  double s = 0.0;
  for (double &x : w) s += (x = u(rng));
  for (double &x : w) x *= 2.0 / s;
  std::vector<double> e(jm + 1);
  cell_edges(w.data(), jm, e.data());
  lat.resize(jm);
  for (int j = 0; j < jm; ++j) lat[j] = 0.5 * (e[j] + e[j + 1]);
}

static void exchange() {
  std::mt19937 rng(11);
  for (int trial = 0; trial < 200; ++trial) {
    const int ja = 1 + rng() % 70, jb = 1 + rng() % 70;
    std::vector<double> la, wa, lb, wb;
    gauss_like_weights(ja, rng, la, wa);
    gauss_like_weights(jb, rng, lb, wb);
    std::vector<double> lat(ja + jb - 1), w(ja + jb - 1);
    const int n = exchange_grid(la.data(), wa.data(), ja, wb.data(), jb, lat.data(), w.data());
    CHECK(n >= 1 && n <= ja + jb - 1);
    double sum = 0.0;
    for (int j = 0; j < n; ++j) {
      sum += w[j];
      CHECK(w[j] > 1e-12);
      if (j) CHECK(lat[j] > lat[j - 1]);
    }
    CHECK(std::fabs(sum - 2.0) < 1e-9);
  }
}

static void admission() {
  const int kThreads = 4, kPerThread = 3000, kMaxFields = 64;
  Admission adm(1 << 20, kMaxFields, 16);
  std::mutex mu;
  std::atomic<int> submitted{0}, refused{0};
  std::atomic<bool> done{false};
  std::vector<std::thread> producers;
  for (int t = 0; t < kThreads; ++t)
    producers.emplace_back([&, t] {
      std::mt19937 rng(100 + t);
      for (int i = 0; i < kPerThread;) {
        const int g = rng() % 3, nf = 1 + rng() % kMaxFields, dir = rng() % 2;
        const size_t bytes = 1 + rng() % 50000;
        long long ticket = 0;
        int rc;
        {
          std::lock_guard<std::mutex> lk(mu);
          rc = adm.submit(g, nf, dir, nullptr, nullptr, bytes, 0.0, &ticket);
        }
        if (rc == SHT_OK) {
          ++submitted;
          ++i;
        } else {
          CHECK(rc == SHT_ERR_BUSY);
          ++refused;
          std::this_thread::yield();
        }
      }
    });
  long long taken = 0, failed_batches = 0;
  std::thread consumer([&] {
    std::mt19937 rng(5);
    std::vector<std::vector<Pending>> inflight;
    for (;;) {
      std::vector<Pending> b;
      {
        std::lock_guard<std::mutex> lk(mu);
        b = adm.take_batch();
        if (!b.empty()) {
          int nf = 0;
          for (const Pending &r : b) {
            nf += r.nf;
            CHECK(r.grid == b[0].grid && r.dir == b[0].dir);
          }
          CHECK(nf <= kMaxFields && (int)b.size() <= 16);
          for (size_t i = 1; i < b.size(); ++i) CHECK(b[i].ticket > b[i - 1].ticket);
          taken += (long long)b.size();
        }
      }
      if (!b.empty()) inflight.push_back(std::move(b));
      // finish a random in-flight batch, failing every seventh
      if (!inflight.empty() && (rng() % 2 || inflight.size() > 4)) {
        const size_t i = rng() % inflight.size();
        const bool ok = rng() % 7 != 0;
        std::lock_guard<std::mutex> lk(mu);
        adm.finish(inflight[i], ok);
        if (!ok) ++failed_batches;
        inflight.erase(inflight.begin() + (long)i);
      }
      if (done && inflight.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        if (adm.empty()) return;
      }
    }
  });
  for (auto &t : producers) t.join();
  done = true;
  consumer.join();
  CHECK(submitted == kThreads * kPerThread);
  CHECK(taken == submitted);
  CHECK(adm.completed() + adm.failed() == submitted);
  CHECK(adm.bytes_used() == 0);
  CHECK(adm.refused() == refused);
  CHECK(failed_batches == 0 || adm.failed() > 0);
  // the budget: exactly-fitting and one-byte-over requests
  Admission tight(100, 8, 4);
  long long t0 = 0;
  CHECK(tight.submit(0, 1, 0, nullptr, nullptr, 100, 0.0, &t0) == SHT_OK);
  CHECK(tight.submit(0, 1, 0, nullptr, nullptr, 1, 0.0, &t0) == SHT_ERR_BUSY);
  Admission huge((size_t)-1, 8, 4);
  CHECK(huge.submit(0, 1, 0, nullptr, nullptr, (size_t)-1, 0.0, &t0) == SHT_OK);
  CHECK(huge.submit(0, 1, 0, nullptr, nullptr, 1, 0.0, &t0) == SHT_ERR_BUSY);  // no wrap-around
}

static void restart() {
  const RestartDims d{8, 4, 3, 2};
  const size_t nxy = 32, nxyz = 96;
  std::vector<std::vector<double>> f(10);
  std::vector<const double *> in(10);
  std::vector<double *> out(10);
  std::vector<std::vector<double>> back(10);
  std::mt19937 rng(3);
  std::normal_distribution<double> nd;
  for (int i = 0; i < 10; ++i) {
    const size_t n = (i % 5 == 4) ? nxy : (i % 5 == 3) ? nxyz * 2 : nxyz;
    f[i].resize(n);
    for (double &x : f[i]) x = nd(rng);
    back[i].assign(n, 0.0);
    in[i] = f[i].data();
    out[i] = back[i].data();
  }
  const double lon[8] = {0, 1, 2, 3, 4, 5, 6, 7}, lat[4] = {-1, -0.3, 0.3, 1}, sig[3] = {0.9, 0.5, 0.1};
  char path[] = "/tmp/sht_rst_XXXXXX";
  const int fd = mkstemp(path);
  CHECK(fd >= 0);
  close(fd);
  std::string msg;
  CHECK(restart_write(path, d, 1234.5, lon, lat, sig, in.data(), msg) == SHT_OK);
  double t = 0.0;
  CHECK(restart_read(path, d, &t, out.data(), msg) == SHT_OK && t == 1234.5);
  for (int i = 0; i < 10; ++i) CHECK(back[i] == f[i]);
  RestartDims other = d;
  other.km = 4;
  CHECK(restart_read(path, other, &t, out.data(), msg) == SHT_ERR_INVALID && !msg.empty());
  // every truncation of the file is refused without reading past its end
  FILE *fp = std::fopen(path, "rb");
  std::vector<char> bytes;
  int c;
  while ((c = std::fgetc(fp)) != EOF) bytes.push_back((char)c);
  std::fclose(fp);
  for (size_t cut = 0; cut < bytes.size(); cut += 1 + cut / 8) {
    FILE *g = std::fopen(path, "wb");
    std::fwrite(bytes.data(), 1, cut, g);
    std::fclose(g);
    CHECK(restart_read(path, d, &t, out.data(), msg) != SHT_OK);
  }
  // random corruption of the header never reads out of bounds
  for (int trial = 0; trial < 300; ++trial) {
    std::vector<char> x = bytes;
    for (int k = 0; k < 4; ++k) x[rng() % 600] = (char)rng();
    FILE *g = std::fopen(path, "wb");
    std::fwrite(x.data(), 1, x.size(), g);
    std::fclose(g);
    (void)restart_read(path, d, &t, out.data(), msg);
  }
  std::remove(path);
  CHECK(restart_write("/nonexistent-dir/x.nc", d, 0.0, lon, lat, sig, in.data(), msg) == SHT_ERR_INVALID);
}

// parallel_copy: every size (below and above the parallel threshold, sizes
// that do not divide into the parts), bytes outside the range untouched, and
// four threads copying at once (one runs in parallel, the others alone)
static void copies() {
  std::mt19937_64 rng(11);
  const size_t sizes[] = {0, 1, 7, kParallelCopyMin - 1, kParallelCopyMin, kParallelCopyMin + 3,
                          (size_t)7 * kParallelCopyMin + 5};
  for (size_t n : sizes) {
    std::vector<unsigned char> src(n + 16), dst(n + 16, 0xAB);
    for (auto &b : src) b = (unsigned char)rng();
    parallel_copy(dst.data() + 8, src.data() + 8, n);
    bool ok = true;
    for (size_t i = 0; i < n + 16; ++i) ok &= dst[i] == ((i >= 8 && i < n + 8) ? src[i] : 0xAB);
    CHECK(ok);
  }
  std::atomic<int> bad{0};
  std::vector<std::thread> th;
  for (int t = 0; t < 4; ++t)
    th.emplace_back([&, t] {
      const size_t n = (size_t)3 * kParallelCopyMin + 17 * t;
      std::vector<unsigned char> a(n), b(n);
      for (int rep = 0; rep < 20; ++rep) {
        for (size_t i = 0; i < n; ++i) a[i] = (unsigned char)(i * 7 + rep + t);
        parallel_copy(b.data(), a.data(), n);
        if (a != b) ++bad;
      }
    });
  for (auto &x : th) x.join();
  CHECK(bad == 0);
  // a forked child (no workers of its own) still copies, alone
  std::fflush(stdout);
  const pid_t pid = fork();
  if (pid == 0) {
    const size_t n = (size_t)2 * kParallelCopyMin + 1;
    std::vector<unsigned char> a(n, 3), b(n, 0);
    parallel_copy(b.data(), a.data(), n);
    _exit(a == b ? 0 : 1);
  }
  int status = -1;
  CHECK(pid > 0 && waitpid(pid, &status, 0) == pid && WIFEXITED(status) && WEXITSTATUS(status) == 0);
}

int main() {
  copies();
  descriptors();
  indices();
  tune_cache();
  exchange();
  admission();
  restart();
  setenv("SHT_TEST_COUNT", "3", 1);
  CHECK(injected_count("SHT_TEST_COUNT") == 3);
  setenv("SHT_TEST_COUNT", "3x", 1);
  CHECK(injected_count("SHT_TEST_COUNT") == 0);
  CHECK(injected_count("SHT_TEST_UNSET_VARIABLE") == 0);
  if (nbad) {
    std::printf("FAIL: %d checks\n", nbad);
    return 1;
  }
  std::printf("PASS host logic under ASan/UBSan\n");
  return 0;
}

// sht_server.hip -- continuous batching of transform requests (include/sht.h,
// "request server").  Requests reference device-resident data; the server
// keeps a plan per grid per slot resident, batches requests on arrival and
// overlaps batches on `slots` HIP streams.
//
// Worker loop (one host thread):
//   1. a slot whose completion event has fired completes its requests
//      (latency = completion seen - submit, host steady clock);
//   2. every free slot takes, in arrival order, the pending requests that share
//      the (grid, direction) of the oldest pending one, up to max_fields fields
//      and kMaxBatchReq requests, and enqueues: gather -> transform -> scatter
//      -> event.  A batch of one request runs in place (no gather/scatter);
//   3. with nothing to do it sleeps on a condition variable (woken by submit)
//      or, while batches run, polls their events every few microseconds.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>

#include "sht_internal.h"

namespace sht {

int run_dir_on(sht_plan *plan, const double *spec_in, double *grid, double *spec_out, int nf,
               int dir, hipStream_t s);

namespace {

constexpr int kMaxBatchReq = 64;
using Clock = std::chrono::steady_clock;

struct ReqEntry {
  const double *src;
  double *dst;
  int f0, nf;
};
struct ReqTable {
  int n;
  ReqEntry r[kMaxBatchReq];
};

__device__ __forceinline__ int find_req(const ReqTable &t, int f) {
  int lo = 0, hi = t.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.r[mid].f0 <= f) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Canonical spectra: batch [K][2 nft] <-> request i [K][2 nf_i] at columns 2 f0_i.
// gather = 1: request -> batch; 0: batch -> request (dst of the entries).
__global__ void spec_move(ReqTable t, double *batch, int K, int nft, int gather) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ncol = 2 * (int64_t)nft;
  if (e >= (int64_t)K * ncol) return;
  const int k = (int)(e / ncol), c = (int)(e % ncol);
  const ReqEntry &q = t.r[find_req(t, c >> 1)];
  const int64_t ri = (int64_t)k * 2 * q.nf + (c - 2 * q.f0);
  if (gather) batch[e] = q.src[ri];
  else q.dst[ri] = batch[e];
}

// Grids: batch [nft][npt] <-> request i [nf_i][npt] at field f0_i.
__global__ void grid_move(ReqTable t, double *batch, int npt, int nft, int gather) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nft * npt) return;
  const int f = (int)(e / npt);
  const ReqEntry &q = t.r[find_req(t, f)];
  const int64_t ri = e - (int64_t)q.f0 * npt;
  if (gather) batch[e] = q.src[ri];
  else q.dst[ri] = batch[e];
}

using Request = host::Pending;

double now_ms() {
  return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count();
}

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  std::vector<sht_plan *> plans;       // per grid
  std::vector<double *> bspec, bgrid;  // per grid, max_fields fields
  std::vector<Request> inflight;
  bool busy = false;
};

}  // namespace
}  // namespace sht

struct sht_server {
  sht_server_desc d{};
  int device = 0;
  std::vector<sht::Slot> slots;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  sht::host::Admission adm{0, 1, 1};                       // pending requests, budget, counters
  std::vector<std::pair<long long, double>> completed;
  int inflight_batches = 0;
  bool stop = false;
  int fail_batch = 0;     // SHT_INJECT_BATCH_FAILURE: this batch (1-based) fails before launch
  long long launched = 0;
  std::string worker_error;   // a device fault: sticky, every later call fails
  std::string batch_error;    // the last failed batch's reason (its requests report -1)
};

namespace sht {
namespace {

size_t grid_doubles(const sht_server_desc &d, int g) { return (size_t)d.im[g] * d.jm[g]; }
size_t spec_doubles(const sht_server_desc &d, int g) {
  return (size_t)(d.nmax[g] + 1) * (d.nmax[g] + 2);  // complex: 2K
}

// Enqueues one batch on slot sl; returns a HIP error string on failure.
std::string launch_batch(sht_server *S, Slot &sl, std::vector<Request> &batch) {
  const sht_server_desc &d = S->d;
  const int g = batch[0].grid, dir = batch[0].dir;
  const int K = (d.nmax[g] + 1) * (d.nmax[g] + 2) / 2;
  const int npt = d.im[g] * d.jm[g];
  sht_plan *plan = sl.plans[g];
  int nft = 0;
  for (auto &r : batch) nft += r.nf;
  if (S->fail_batch && ++S->launched == S->fail_batch) return "injected batch failure (SHT_INJECT_BATCH_FAILURE)";
  int rc;
  if (batch.size() == 1) {  // in place
    const Request &r = batch[0];
    rc = dir == SHT_DIR_INVERSE ? run_dir_on(plan, r.in, r.out, nullptr, r.nf, dir, sl.s)
                                : run_dir_on(plan, nullptr, const_cast<double *>(r.in), r.out, r.nf, dir, sl.s);
    if (rc) return sht_last_error();
  } else {
    ReqTable tin{}, tout{};
    tin.n = tout.n = (int)batch.size();
    int f0 = 0;
    for (size_t i = 0; i < batch.size(); ++i) {
      tin.r[i] = ReqEntry{batch[i].in, nullptr, f0, batch[i].nf};
      tout.r[i] = ReqEntry{nullptr, batch[i].out, f0, batch[i].nf};
      f0 += batch[i].nf;
    }
    const int64_t nspec = (int64_t)K * 2 * nft, ngrid = (int64_t)nft * npt;
    if (dir == SHT_DIR_INVERSE) {
      spec_move<<<(unsigned)((nspec + 255) / 256), 256, 0, sl.s>>>(tin, sl.bspec[g], K, nft, 1);
      rc = run_dir_on(plan, sl.bspec[g], sl.bgrid[g], nullptr, nft, dir, sl.s);
      if (rc) return sht_last_error();
      grid_move<<<(unsigned)((ngrid + 255) / 256), 256, 0, sl.s>>>(tout, sl.bgrid[g], npt, nft, 0);
    } else {
      grid_move<<<(unsigned)((ngrid + 255) / 256), 256, 0, sl.s>>>(tin, sl.bgrid[g], npt, nft, 1);
      rc = run_dir_on(plan, nullptr, sl.bgrid[g], sl.bspec[g], nft, dir, sl.s);
      if (rc) return sht_last_error();
      spec_move<<<(unsigned)((nspec + 255) / 256), 256, 0, sl.s>>>(tout, sl.bspec[g], K, nft, 0);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hipGetErrorString(e);
  }
  hipError_t e = hipEventRecord(sl.done, sl.s);
  if (e != hipSuccess) return hipGetErrorString(e);
  sl.inflight = std::move(batch);
  sl.busy = true;
  return std::string();
}

void worker_loop(sht_server *S) {
  (void)hipSetDevice(S->device);
  const sht_server_desc &d = S->d;
  for (;;) {
    bool progressed = false;
    // 1. completions
    for (auto &sl : S->slots) {
      if (!sl.busy) continue;
      hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      const double now = now_ms();
      std::lock_guard<std::mutex> lk(S->mu);
      if (q != hipSuccess && S->worker_error.empty()) S->worker_error = hipGetErrorString(q);
      for (auto &r : sl.inflight) S->completed.emplace_back(r.ticket, q == hipSuccess ? now - r.t_submit : -1.0);
      S->adm.finish(sl.inflight, q == hipSuccess);
      sl.inflight.clear();
      sl.busy = false;
      S->inflight_batches -= 1;
      progressed = true;
    }
    // 2. launches
    for (auto &sl : S->slots) {
      if (sl.busy) continue;
      std::vector<Request> batch;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->adm.empty()) break;
        batch = S->adm.take_batch();
        S->inflight_batches += 1;
      }
      std::string err = launch_batch(S, sl, batch);
      if (!err.empty()) {
        // this batch's requests fail (latency -1); the others go on, since
        // nothing of it reached the device
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(S->mu);
        S->batch_error = err;
        for (auto &r : batch) S->completed.emplace_back(r.ticket, -1.0);
        S->adm.finish(batch, false);
        S->inflight_batches -= 1;
      }
      progressed = true;
    }
    if (progressed) continue;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->inflight_batches == 0) {
      if (S->stop && S->adm.empty()) return;
      S->cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return S->stop || !S->adm.empty(); });
    } else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
  }
}

void server_release(sht_server *S) {
  for (auto &sl : S->slots) {
    for (auto *p : sl.plans)
      if (p) sht_plan_destroy(p);
    for (auto *q : sl.bspec)
      if (q) (void)hipFree(q);
    for (auto *q : sl.bgrid)
      if (q) (void)hipFree(q);
    if (sl.s) (void)hipStreamDestroy(sl.s);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  delete S;
}

}  // namespace
}  // namespace sht

using namespace sht;

extern "C" {

int sht_server_create(const sht_server_desc *desc, sht_server **out) {
  if (!out) return fail(SHT_ERR_INVALID, "sht_server_create: out is NULL");
  *out = nullptr;
  if (!desc || desc->ngrid < 1 || desc->ngrid > 8 || desc->slots < 1 || desc->slots > 8 ||
      desc->max_fields < 1 || desc->mem_budget == 0)
    return fail(SHT_ERR_INVALID, "sht_server_create: bad descriptor");
  for (int g = 0; g < desc->ngrid; ++g)
    for (int h = 0; h < g; ++h)
      if (desc->nmax[g] == desc->nmax[h])
        return fail(SHT_ERR_INVALID, "sht_server_create: each grid needs a distinct nmax");
  sht_server *S = new (std::nothrow) sht_server;
  if (!S) return fail(SHT_ERR_NOMEM, "sht_server_create: host allocation failed");
  S->d = *desc;
  S->adm = host::Admission(desc->mem_budget, desc->max_fields, kMaxBatchReq);
  S->fail_batch = host::injected_count("SHT_INJECT_BATCH_FAILURE");
  if (hipGetDevice(&S->device) != hipSuccess) {
    delete S;
    return fail(SHT_ERR_DEVICE, "sht_server_create: hipGetDevice failed");
  }
  S->slots.resize(desc->slots);
  for (auto &sl : S->slots) {
    sl.plans.assign(desc->ngrid, nullptr);
    sl.bspec.assign(desc->ngrid, nullptr);
    sl.bgrid.assign(desc->ngrid, nullptr);
    if (hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      server_release(S);
      return fail(SHT_ERR_DEVICE, "sht_server_create: stream/event creation failed");
    }
    for (int g = 0; g < desc->ngrid; ++g) {
      sht_desc pd{desc->nmax[g], desc->im[g], desc->jm[g], desc->max_fields, desc->precision, 0};
      int rc = sht_plan_create(&pd, &sl.plans[g]);
      if (rc) {
        std::string keep = sht_last_error();
        server_release(S);
        return fail(rc, "sht_server_create: " + keep);
      }
      if (hipMalloc(&sl.bspec[g], sizeof(double) * spec_doubles(*desc, g) * desc->max_fields) != hipSuccess ||
          hipMalloc(&sl.bgrid[g], sizeof(double) * grid_doubles(*desc, g) * desc->max_fields) != hipSuccess) {
        server_release(S);
        return fail(SHT_ERR_NOMEM, "sht_server_create: batch buffer allocation failed");
      }
    }
  }
  S->worker = std::thread(worker_loop, S);
  *out = S;
  return SHT_OK;
}

void sht_server_destroy(sht_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  // the worker returns once nothing is pending or in flight; the stream
  // waits only cover a batch whose launch failed part-way
  if (S->worker.joinable()) S->worker.join();
  for (auto &sl : S->slots)
    if (sl.s) (void)hipStreamSynchronize(sl.s);
  server_release(S);
}

int sht_server_submit(sht_server *S, const sht_request *r, long long *ticket) {
  if (!S || !r || !ticket) return fail(SHT_ERR_INVALID, "sht_server_submit: NULL argument");
  int g = -1;
  for (int i = 0; i < S->d.ngrid; ++i)
    if (S->d.nmax[i] == r->nmax) g = i;
  if (g < 0) return fail(SHT_ERR_INVALID, "sht_server_submit: nmax not served");
  if (r->nfield < 1 || r->nfield > S->d.max_fields)
    return fail(SHT_ERR_INVALID, "sht_server_submit: nfield must be in [1, max_fields]");
  if (r->direction != SHT_DIR_INVERSE && r->direction != SHT_DIR_FORWARD)
    return fail(SHT_ERR_INVALID, "sht_server_submit: direction must be INVERSE or FORWARD");
  if (!r->in || !r->out) return fail(SHT_ERR_INVALID, "sht_server_submit: NULL data pointer");
  const size_t bytes = sizeof(double) * (size_t)r->nfield * (grid_doubles(S->d, g) + spec_doubles(S->d, g));
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->worker_error.empty()) return fail(SHT_ERR_DEVICE, "sht_server_submit: " + S->worker_error);
    if (S->adm.submit(g, r->nfield, r->direction, r->in, r->out, bytes, now_ms(), ticket) != SHT_OK)
      return fail(SHT_ERR_BUSY, "sht_server_submit: memory budget exhausted");
  }
  S->cv.notify_one();
  return SHT_OK;
}

int sht_server_poll(sht_server *S, long long *tickets, double *latency_ms, int max, int *n) {
  if (!S || !n) return fail(SHT_ERR_INVALID, "sht_server_poll: NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  int k = 0;
  while (k < max && k < (int)S->completed.size()) {
    if (tickets) tickets[k] = S->completed[k].first;
    if (latency_ms) latency_ms[k] = S->completed[k].second;
    ++k;
  }
  S->completed.erase(S->completed.begin(), S->completed.begin() + k);
  *n = k;
  if (!S->worker_error.empty()) return fail(SHT_ERR_DEVICE, "sht_server: " + S->worker_error);
  return SHT_OK;
}

int sht_server_drain(sht_server *S) {
  if (!S) return fail(SHT_ERR_INVALID, "sht_server_drain: server is NULL");
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (!S->worker_error.empty()) return fail(SHT_ERR_DEVICE, "sht_server: " + S->worker_error);
      if (S->adm.empty() && S->inflight_batches == 0) return SHT_OK;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

int sht_server_stats(const sht_server *S, long long *requests, long long *fields, long long *batches,
                     long long *refused) {
  if (!S) return fail(SHT_ERR_INVALID, "sht_server_stats: server is NULL");
  sht_server *M = const_cast<sht_server *>(S);
  std::lock_guard<std::mutex> lk(M->mu);
  if (requests) *requests = S->adm.completed();
  if (fields) *fields = S->adm.completed_fields();
  if (batches) *batches = S->adm.batches();
  if (refused) *refused = S->adm.refused();
  return SHT_OK;
}

int sht_server_failures(const sht_server *S, long long *failed, char *reason, int reason_len) {
  if (!S) return fail(SHT_ERR_INVALID, "sht_server_failures: server is NULL");
  sht_server *M = const_cast<sht_server *>(S);
  std::lock_guard<std::mutex> lk(M->mu);
  if (failed) *failed = S->adm.failed();
  if (reason && reason_len > 0) {
    const std::string &e = S->worker_error.empty() ? S->batch_error : S->worker_error;
    const size_t n = std::min(e.size(), (size_t)reason_len - 1);
    e.copy(reason, n);
    reason[n] = '\0';
  }
  return SHT_OK;
}

}  // extern "C"

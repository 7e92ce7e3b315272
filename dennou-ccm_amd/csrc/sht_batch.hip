// sht_batch.hip -- ensemble batches: members of uneven truncation and level
// count, grouped by grid, transformed by one hipGraph launch.
//
// Layout: for each group (members with the same nmax, im, jm) the batch owns
// one resident field batch -- canonical spectra in and out [K][2 * nf_group]
// and grid [nf_group][jm][im] -- and member j is the field range
// [f0_j, f0_j + nfield_j) of its group.  Each stage kernel of a group therefore
// covers all of the group's members in a single launch (its tiles are spread
// over the whole chip), and the groups are independent branches of the
// captured graph: the graph forks from the launch stream to one stream per
// group and joins back.
#include <algorithm>
#include <map>
#include <new>
#include <tuple>

#include "sht_internal.h"

namespace sht {

struct BatchGroup {
  int nmax = 0, im = 0, jm = 0, nf = 0;
  sht_plan *plan = nullptr;
  double *spec_in = nullptr, *grid = nullptr, *spec_out = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
};

struct BatchMember {
  int group = 0, f0 = 0, nf = 0;
};

int run_dir_on(sht_plan *plan, const double *spec_in, double *grid, double *spec_out, int nf,
               int dir, hipStream_t s);

}  // namespace sht

struct sht_batch {
  std::vector<sht::BatchGroup> groups;
  std::vector<sht::BatchMember> members;
  hipEvent_t fork = nullptr;
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
};

namespace sht {

static void batch_release(sht_batch *b) {
  for (int d = 0; d < 3; ++d) {
    if (b->exec[d]) (void)hipGraphExecDestroy(b->exec[d]);
    if (b->graph[d]) (void)hipGraphDestroy(b->graph[d]);
  }
  for (auto &g : b->groups) {
    if (g.plan) sht_plan_destroy(g.plan);
    if (g.spec_in) (void)hipFree(g.spec_in);
    if (g.grid) (void)hipFree(g.grid);
    if (g.spec_out) (void)hipFree(g.spec_out);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    if (g.done) (void)hipEventDestroy(g.done);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
}

// Enqueue every group's work; with fork_join the groups go to their own
// streams (for capture into parallel graph branches), else all on s.
static int enqueue_all(sht_batch *b, int dir, hipStream_t s, bool fork_join) {
  if (fork_join) SHT_HIP(hipEventRecord(b->fork, s));
  for (auto &g : b->groups) {
    hipStream_t gs = s;
    if (fork_join) {
      gs = g.stream;
      SHT_HIP(hipStreamWaitEvent(gs, b->fork, 0));
    }
    int rc = run_dir_on(g.plan, g.spec_in, g.grid, g.spec_out, g.nf, dir, gs);
    if (rc) return rc;
    if (fork_join) {
      SHT_HIP(hipEventRecord(g.done, gs));
      SHT_HIP(hipStreamWaitEvent(s, g.done, 0));
    }
  }
  return SHT_OK;
}

}  // namespace sht

using namespace sht;

extern "C" {

int sht_batch_create(const sht_member_desc *members, int nmember, int precision, sht_batch **out) {
  if (!out) return fail(SHT_ERR_INVALID, "sht_batch_create: out is NULL");
  *out = nullptr;
  if (!members || nmember < 1) return fail(SHT_ERR_INVALID, "sht_batch_create: no members");
  sht_batch *b = new (std::nothrow) sht_batch;
  if (!b) return fail(SHT_ERR_NOMEM, "sht_batch_create: host allocation failed");
  std::map<std::tuple<int, int, int>, int> gid;
  for (int j = 0; j < nmember; ++j) {
    const sht_member_desc &d = members[j];
    if (d.nfield < 1) {
      batch_release(b);
      return fail(SHT_ERR_INVALID, "sht_batch_create: member with nfield < 1");
    }
    auto key = std::make_tuple(d.nmax, d.im, d.jm);
    auto it = gid.find(key);
    int g;
    if (it == gid.end()) {
      g = (int)b->groups.size();
      gid[key] = g;
      BatchGroup grp;
      grp.nmax = d.nmax;
      grp.im = d.im;
      grp.jm = d.jm;
      b->groups.push_back(grp);
    } else {
      g = it->second;
    }
    BatchMember m;
    m.group = g;
    m.f0 = b->groups[g].nf;
    m.nf = d.nfield;
    b->groups[g].nf += d.nfield;
    b->members.push_back(m);
  }
  hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    batch_release(b);
    return fail(SHT_ERR_DEVICE, "sht_batch_create: hipEventCreate failed");
  }
  for (auto &g : b->groups) {
    sht_desc d{g.nmax, g.im, g.jm, g.nf, precision, 0};
    int rc = sht_plan_create(&d, &g.plan);
    if (rc) {
      std::string keep = sht_last_error();
      batch_release(b);
      return fail(rc, "sht_batch_create: " + keep);
    }
    const size_t K = (size_t)(g.nmax + 1) * (g.nmax + 2) / 2;
    const size_t nspec = K * 2 * g.nf, ngrid = (size_t)g.nf * g.jm * g.im;
    if (hipMalloc(&g.spec_in, sizeof(double) * nspec) != hipSuccess ||
        hipMalloc(&g.spec_out, sizeof(double) * nspec) != hipSuccess ||
        hipMalloc(&g.grid, sizeof(double) * ngrid) != hipSuccess) {
      batch_release(b);
      return fail(SHT_ERR_NOMEM, "sht_batch_create: member state allocation failed");
    }
    if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g.done, hipEventDisableTiming) != hipSuccess) {
      batch_release(b);
      return fail(SHT_ERR_DEVICE, "sht_batch_create: stream/event creation failed");
    }
    // defined contents before the first run
    if (hipMemset(g.spec_in, 0, sizeof(double) * nspec) != hipSuccess ||
        hipMemset(g.spec_out, 0, sizeof(double) * nspec) != hipSuccess ||
        hipMemset(g.grid, 0, sizeof(double) * ngrid) != hipSuccess) {
      batch_release(b);
      return fail(SHT_ERR_DEVICE, "sht_batch_create: hipMemset failed");
    }
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    batch_release(b);
    return fail(SHT_ERR_DEVICE, "sht_batch_create: device synchronisation failed");
  }
  *out = b;
  return SHT_OK;
}

void sht_batch_destroy(sht_batch *b) {
  if (b) batch_release(b);
}

int sht_batch_member(const sht_batch *b, int j, double **spec_in, double **grid, double **spec_out,
                     int *ld) {
  if (!b || j < 0 || j >= (int)b->members.size())
    return fail(SHT_ERR_INVALID, "sht_batch_member: bad batch or member index");
  const BatchMember &m = b->members[j];
  const BatchGroup &g = b->groups[m.group];
  if (spec_in) *spec_in = g.spec_in + 2 * (size_t)m.f0;
  if (spec_out) *spec_out = g.spec_out + 2 * (size_t)m.f0;
  if (grid) *grid = g.grid + (size_t)m.f0 * g.jm * g.im;
  if (ld) *ld = 2 * g.nf;
  return SHT_OK;
}

int sht_batch_run(sht_batch *b, int dir, void *stream) {
  if (!b) return fail(SHT_ERR_INVALID, "sht_batch_run: batch is NULL");
  if (dir < SHT_DIR_INVERSE || dir > SHT_DIR_ROUNDTRIP)
    return fail(SHT_ERR_INVALID, "sht_batch_run: bad direction");
  hipStream_t s = (hipStream_t)stream;
  if (!b->exec[dir]) {
    // capture on a private stream so the caller's stream may be the legacy one
    hipStream_t cap = nullptr;
    SHT_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    int rc = (e == hipSuccess) ? enqueue_all(b, dir, cap, true)
                               : fail(SHT_ERR_DEVICE, "sht_batch_run: hipStreamBeginCapture failed");
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(cap, &graph);
    (void)hipStreamDestroy(cap);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e2 != hipSuccess) return fail(SHT_ERR_DEVICE, "sht_batch_run: hipStreamEndCapture failed");
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(SHT_ERR_DEVICE, "sht_batch_run: hipGraphInstantiate failed");
    }
    b->graph[dir] = graph;
    b->exec[dir] = exec;
  }
  SHT_HIP(hipGraphLaunch(b->exec[dir], s));
  return SHT_OK;
}

int sht_batch_autotune(sht_batch *b, int reps, const char *cache_path) {
  return sht_batch_autotune_flags(b, reps, cache_path, 0);
}

int sht_batch_autotune_flags(sht_batch *b, int reps, const char *cache_path, int flags) {
  if (!b) return fail(SHT_ERR_INVALID, "sht_batch_autotune: batch is NULL");
  // the tuner runs on a stream of its own through each group plan's Fourier
  // scratch: a graph or eager run still in flight on any stream would share it
  SHT_HIP(hipDeviceSynchronize());
  for (auto &g : b->groups) {
    int rc = sht_plan_autotune_flags(g.plan, g.nf, reps, cache_path, flags, nullptr);
    if (rc) return rc;
  }
  // graphs captured before hold the old kernels: capture again on the next run
  for (int d = 0; d < 3; ++d) {
    if (b->exec[d]) (void)hipGraphExecDestroy(b->exec[d]);
    if (b->graph[d]) (void)hipGraphDestroy(b->graph[d]);
    b->exec[d] = nullptr;
    b->graph[d] = nullptr;
  }
  SHT_HIP(hipDeviceSynchronize());
  return SHT_OK;
}

int sht_batch_run_eager(sht_batch *b, int dir, void *stream) {
  if (!b) return fail(SHT_ERR_INVALID, "sht_batch_run_eager: batch is NULL");
  if (dir < SHT_DIR_INVERSE || dir > SHT_DIR_ROUNDTRIP)
    return fail(SHT_ERR_INVALID, "sht_batch_run_eager: bad direction");
  return enqueue_all(b, dir, (hipStream_t)stream, false);
}

}  // extern "C"

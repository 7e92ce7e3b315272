// sht_tables.hip -- plan-time FP64 kernels: Gauss-Legendre nodes and weights,
// the normalised associated Legendre table Pt_n^m(mu_p), its latitude-derivative
// table, and FFT twiddles.  All of it stays FP64 whatever the precision of the
// Legendre GEMMs (SPML wa_Initial's work, run inside DynamicsHSplVAS83Init at
// atm/dcpam_main_mod.f90:5036).
#include <algorithm>

#include "sht_internal.h"

namespace sht {
namespace {

// P_n(cos theta) and (1 - x^2) dP_n/dx = n (P_{n-1} - x P_n).  Near the pole
// (x > 1/2) the Bonnet recurrence runs in its difference form on
// u = 1 - x = 2 sin^2(theta/2) taken from theta, so the node's colatitude is
// resolved to its own rounding instead of eps/sin(theta):
//   D_k = ((k-1) D_{k-1} - (2k-1) u P_{k-1}) / k,  P_k = P_{k-1} + D_k,
//   P_{n-1} - x P_n = u P_n - D_n.
__device__ void legendre_pn_theta(int n, double th, double *pn, double *dq) {
  double x = cos(th);
  if (x > 0.5) {
    double sh = sin(0.5 * th), u = 2.0 * sh * sh;
    double p = 1.0, d = 0.0;
    for (int k = 1; k <= n; ++k) {
      d = ((k - 1.0) * d - (2.0 * k - 1.0) * u * p) / k;
      p += d;
    }
    *pn = p;
    *dq = n * (u * p - d);
  } else {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    *pn = p1;
    *dq = n * (p0 - x * p1);
  }
}

// Newton iteration in the colatitude, one thread per northern root.
__global__ void gauss_kernel(int jm, double *mu, double *w, double *coslat) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= jm / 2) return;
  double th = M_PI * (4.0 * k + 3.0) / (4.0 * jm + 2.0);
  double pn = 0.0, dq = 0.0;
  for (int it = 0; it < 100; ++it) {
    legendre_pn_theta(jm, th, &pn, &dq);
    double d = -pn * sin(th) / dq;
    th -= d;
    if (fabs(d) < 1e-15 * th) {  // converged: one more step, then stop
      legendre_pn_theta(jm, th, &pn, &dq);
      th -= -pn * sin(th) / dq;
      break;
    }
  }
  legendre_pn_theta(jm, th, &pn, &dq);
  double s = sin(th), x = cos(th);
  double wk = 2.0 * s * s / (dq * dq);
  mu[jm - 1 - k] = x;
  mu[k] = -x;
  w[jm - 1 - k] = wk;
  w[k] = wk;
  coslat[jm - 1 - k] = s;
  coslat[k] = s;
}

__device__ __forceinline__ double eps_nm(int n, int m) {
  double nn = n, mm = m;
  return sqrt((nn * nn - mm * mm) / (4.0 * nn * nn - 1.0));
}

// Extended exponent: a value is x * 2^(960 e), e <= 0 shared by the
// recurrence's live terms.  The sectoral start sin^m(theta) falls below the
// double range at high m near the poles (about 1e-345 at m = 170 on the T170
// grid); the scaled form keeps the recurrence exact until the result is
// representable again.
__device__ __forceinline__ double xout(double x, int e) { return e == 0 ? x : ldexp(x, 960 * e); }

// one thread per (latitude pair p, wavenumber m)
__global__ void legendre_kernel(int N, int P2, int P2p, const double *__restrict__ mu,
                                const double *__restrict__ coslat, const double *__restrict__ w,
                                const int *__restrict__ poff_inv, const int *__restrict__ poff_fwd,
                                double *__restrict__ Pinv, double *__restrict__ Pfwd,
                                double *__restrict__ Hinv, double *__restrict__ Hfwd) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int p = t % P2, m = t / P2;
  if (m > N) return;
  const int jn = P2 + p; // north node of the pair
  const double x = mu[jn], s = coslat[jn], hw = 0.5 * w[jn];
  const int ke = ceil_half(N - m + 1);
  const int KP = roundup(ke, 4), KR = roundup(ke, 16);
  double *Pi = Pinv + poff_inv[m];
  double *Pf = Pfwd + poff_fwd[m];
  double *Hi = Hinv ? Hinv + poff_inv[m] : nullptr;
  double *Hf = Hfwd ? Hfwd + poff_fwd[m] : nullptr;

  double pc = 1.0;
  int e = 0;
  for (int k = 1; k <= m; ++k) {
    pc *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    if (pc != 0.0 && fabs(pc) < 0x1p-480) {
      pc *= 0x1p960;
      e -= 1;
    }
  }
  double pp = 0.0; // Pt_{n-1}
  for (int n = m; n <= N; ++n) {
    // Pt_{n+1} (needed by the derivative of degree n)
    double back = (n > m) ? eps_nm(n, m) * pp : 0.0;
    double pn = (x * pc - back) / eps_nm(n + 1, m);
    if (e < 0 && fabs(pn) >= 0x1p480) {
      pn *= 0x1p-960;
      pc *= 0x1p-960;
      pp *= 0x1p-960;
      e += 1;
    }
    const int d = n - m, par = d & 1, k = d >> 1;
    const double v = xout(pc, e);
    Pi[(size_t)(par * KP + k) * P2p + p] = v;
    Pf[(size_t)(par * P2p + p) * KR + k] = hw * v;
    if (Hi) {
      // (1-mu^2) dPt_n/dmu = (n+1) eps_n Pt_{n-1} - n eps_{n+1} Pt_{n+1}
      double h = (n > m ? (n + 1.0) * eps_nm(n, m) * pp : 0.0) - n * eps_nm(n + 1, m) * pn;
      Hi[(size_t)(par * KP + k) * P2p + p] = xout(h, e);
      // divergence analysis weights -(w/2) H (the lat term integrated by parts)
      Hf[(size_t)(par * P2p + p) * KR + k] = -hw * xout(h, e);
    }
    pp = pc;
    pc = pn;
  }
}

__global__ void twiddle_kernel(int L, double *tw) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L) return;
  double s, c;
  sincospi(2.0 * (double)t / (double)L, &s, &c);
  tw[2 * t] = c;
  tw[2 * t + 1] = s;
}

}  // namespace

int launch_gauss(int jm, double *mu, double *w, double *coslat, hipStream_t s) {
  gauss_kernel<<<(jm / 2 + 63) / 64, 64, 0, s>>>(jm, mu, w, coslat);
  SHT_HIP(hipGetLastError());
  return SHT_OK;
}

int build_tables(Plan &p) {
  const int N = p.N, P2 = p.P2, P2p = p.P2p;
  p.h_poff_inv.resize(N + 1);
  p.h_poff_fwd.resize(N + 1);
  size_t ti = 0, tf = 0;
  for (int m = 0; m <= N; ++m) {
    int ke = ceil_half(N - m + 1);
    p.h_poff_inv[m] = (int)ti;
    p.h_poff_fwd[m] = (int)tf;
    ti += (size_t)2 * roundup(ke, 4) * P2p;
    tf += (size_t)2 * P2p * roundup(ke, 16);
  }
  if (ti > (size_t)INT32_MAX || tf > (size_t)INT32_MAX)
    return fail(SHT_ERR_UNSUPPORTED, "Legendre table exceeds 2^31 entries");
  p.n_inv = ti;
  p.n_fwd = tf;
  SHT_PLAN_ALLOC(p, &p.mu, sizeof(double) * p.jm);
  SHT_PLAN_ALLOC(p, &p.w, sizeof(double) * p.jm);
  SHT_PLAN_ALLOC(p, &p.coslat, sizeof(double) * p.jm);
  SHT_PLAN_ALLOC(p, &p.twid, sizeof(double) * 2 * p.im);
  SHT_PLAN_ALLOC(p, &p.zero, sizeof(double) * 64);
  SHT_HIP(hipMemset(p.zero, 0, sizeof(double) * 64));
  SHT_PLAN_ALLOC(p, &p.Pinv, sizeof(double) * ti);
  SHT_PLAN_ALLOC(p, &p.Pfwd, sizeof(double) * tf);
  if (p.flags & SHT_FLAG_GRAD) SHT_PLAN_ALLOC(p, &p.Hinv, sizeof(double) * ti);
  if (p.flags & SHT_FLAG_GRAD) SHT_PLAN_ALLOC(p, &p.Hfwd, sizeof(double) * tf);
  SHT_PLAN_ALLOC(p, &p.poff_inv, sizeof(int) * (N + 1));
  SHT_PLAN_ALLOC(p, &p.poff_fwd, sizeof(int) * (N + 1));
  SHT_HIP(hipMemcpy(p.poff_inv, p.h_poff_inv.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
  SHT_HIP(hipMemcpy(p.poff_fwd, p.h_poff_fwd.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
  SHT_HIP(hipMemset(p.Pinv, 0, sizeof(double) * ti));
  SHT_HIP(hipMemset(p.Pfwd, 0, sizeof(double) * tf));
  if (p.Hinv) SHT_HIP(hipMemset(p.Hinv, 0, sizeof(double) * ti));
  if (p.Hfwd) SHT_HIP(hipMemset(p.Hfwd, 0, sizeof(double) * tf));
  std::vector<int> units;
  for (int m = 0; m <= N; ++m)
    for (int kt = 0; kt < roundup(ceil_half(N - m + 1), 16) / 16; ++kt) units.push_back(m << 8 | kt);
  p.n_fwd_units = (int)units.size();
  SHT_PLAN_ALLOC(p, &p.fwd_units, sizeof(int) * units.size());
  SHT_HIP(hipMemcpy(p.fwd_units, units.data(), sizeof(int) * units.size(), hipMemcpyHostToDevice));
  for (int dir = 0; dir < 2; ++dir) {
    std::vector<int> work(N + 1), order(N + 1);
    for (int m = 0; m <= N; ++m) {
      const int ke = ceil_half(N - m + 1);
      work[m] = dir == 0 ? roundup(ke, 4) / 4 : roundup(ke, 16) / 16;
      order[m] = m;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    std::vector<std::vector<int>> lists(8);
    std::vector<long> load(8, 0);
    for (int m : order) {
      int x = 0;
      for (int y = 1; y < 8; ++y)
        if (load[y] < load[x]) x = y;
      lists[x].push_back(m);
      load[x] += work[m];
    }
    size_t len = 0;
    for (auto &l : lists) len = std::max(len, l.size());
    std::vector<int> flat(8 * len, -1);
    for (int x = 0; x < 8; ++x)
      for (size_t j = 0; j < lists[x].size(); ++j) flat[x * len + j] = lists[x][j];
    p.xcd_mlen[dir] = (int)len;
    SHT_PLAN_ALLOC(p, &p.xcd_m[dir], sizeof(int) * flat.size());
    SHT_HIP(hipMemcpy(p.xcd_m[dir], flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice));
  }

  int rc = launch_gauss(p.jm, p.mu, p.w, p.coslat, nullptr);
  if (rc) return rc;
  int nthr = P2 * (N + 1);
  legendre_kernel<<<(nthr + 127) / 128, 128>>>(N, P2, P2p, p.mu, p.coslat, p.w, p.poff_inv,
                                                 p.poff_fwd, p.Pinv, p.Pfwd, p.Hinv, p.Hfwd);
  SHT_HIP(hipGetLastError());
  twiddle_kernel<<<(p.im + 127) / 128, 128>>>(p.im, p.twid);
  SHT_HIP(hipGetLastError());
  SHT_HIP(hipDeviceSynchronize());
  p.h_mu.resize(p.jm);
  p.h_w.resize(p.jm);
  p.h_coslat.resize(p.jm);
  SHT_HIP(hipMemcpy(p.h_mu.data(), p.mu, sizeof(double) * p.jm, hipMemcpyDeviceToHost));
  SHT_HIP(hipMemcpy(p.h_w.data(), p.w, sizeof(double) * p.jm, hipMemcpyDeviceToHost));
  SHT_HIP(hipMemcpy(p.h_coslat.data(), p.coslat, sizeof(double) * p.jm, hipMemcpyDeviceToHost));
  return SHT_OK;
}

}  // namespace sht

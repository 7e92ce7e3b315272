/*
 * sht_oracle.c -- CPU FP64 restatement of the Gaussian-grid SHT (see header).
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline, never the product).
 *
 * What the reference fixes and where:
 *   - grid arrays (0:imax-1, 1:jmax, 1:kmax), FP64: atm/dcpam_main_mod.f90:4107-4125
 *   - Gauss latitudes S->N with weights summing to 2 (cell edges built by
 *     cumulative weights from -pi/2 end at +pi/2): tool/gmapgen/gmapgen_main.f90:296-303,
 *     :364-368 and exp/APEI07Couple/ebm/EBM_main.f90:200-204
 *   - zonal normalisation Pt_n = sqrt(2n+1) P_n, l_y = (1/2) sum_j w_j Pt_n f_j:
 *     EBM_main.f90:170-174, :242-246, Bonnet recurrence :97-103
 *   - N/S mirrored latitude pairs as the unit of the Legendre step:
 *     atm/dccm_atm_mod.f90:953-1001
 * The transform itself (SPML wa_module over ISPACK) is not in the reference
 * tree; the algorithm restated here is the textbook one SPML/ISPACK implement:
 * a real FFT per ring, then per zonal wavenumber a Gauss quadrature split into
 * equatorially symmetric and antisymmetric halves.
 */
#include "sht_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre nodes: Newton iteration in the colatitude theta, so that both
 * mu = cos(theta) and cos(lat) = sin(theta) keep full relative precision.     */

/* P_n and the combination n (P_{n-1} - x P_n) = (1 - x^2) dP_n/dx at
 * x = cos(theta).  For x > 1/2 Reinsch's difference form of the Bonnet
 * recurrence (EBM_main.f90:97-103 is the plain form) runs in u = 1 - x =
 * 2 sin^2(theta/2), computed from theta, so no rounded x enters near the pole:
 *   D_k = P_k - P_{k-1} = ((k-1) D_{k-1} - (2k-1) u P_{k-1}) / k,  P_k = P_{k-1} + D_k
 * and P_{n-1} - x P_n = u P_n - D_n.                                          */
static void legendre_pn_theta(int n, double th, double *pn, double *dq) {
  double x = cos(th);
  if (x > 0.5) {
    double sh = sin(0.5 * th), u = 2.0 * sh * sh;
    double p = 1.0, d = 0.0; /* P_0, D_0 (unused) */
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

int orc_gauss(int jm, double *mu, double *w, double *coslat) {
  if (jm < 2 || (jm & 1)) return -1;
  const int h = jm / 2;
  for (int k = 0; k < h; ++k) {
    /* k-th root counted from the north pole: theta ~ pi (4k+3)/(4jm+2). */
    double th = M_PI * (4.0 * k + 3.0) / (4.0 * jm + 2.0);
    double pn = 0, dq = 0;
    for (int it = 0; it < 100; ++it) {
      legendre_pn_theta(jm, th, &pn, &dq);
      /* dP/dtheta = -sin(theta) dP/dx = -dq / sin(theta) */
      double d = -pn * sin(th) / dq;
      th -= d;
      if (fabs(d) < 1e-15 * th) { /* converged: one more step, then stop */
        legendre_pn_theta(jm, th, &pn, &dq);
        th -= -pn * sin(th) / dq;
        break;
      }
    }
    legendre_pn_theta(jm, th, &pn, &dq);
    double s = sin(th);
    double wk = 2.0 * s * s / (dq * dq); /* 2 / ((1-x^2) P_n'(x)^2) */
    /* north node index jm-1-k, mirrored south node k */
    mu[jm - 1 - k] = cos(th);
    mu[k] = -mu[jm - 1 - k];
    w[jm - 1 - k] = w[k] = wk;
    coslat[jm - 1 - k] = coslat[k] = s;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Normalised associated Legendre functions with an extended exponent: values
 * are x * 2^(XEXP*e); e < 0 marks magnitudes below the double range, which
 * lets the sectoral start sin^m(theta) fall below 1e-308 at high m near the
 * poles without flushing the whole column to zero early.                    */
#define XEXP 960
#define XHI 0x1p480
#define XLO 0x1p-480

static double eps_nm(int n, int m) {
  double nn = (double)n, mm = (double)m;
  return sqrt((nn * nn - mm * mm) / (4.0 * nn * nn - 1.0));
}

static double xnorm_out(double x, int e) { return e == 0 ? x : ldexp(x, XEXP * e); }

void orc_legendre_column(int nmax, int m, int npt, const double *mu,
                         const double *coslat, double *out) {
  for (int t = 0; t < npt; ++t) {
    double s = coslat[t], x = mu[t];
    double pm = 1.0;
    int e = 0;
    for (int k = 1; k <= m; ++k) {
      pm *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
      if (pm != 0.0 && fabs(pm) < XLO) { pm *= 0x1p960; e -= 1; }
    }
    double p2 = 0.0, p1 = pm; /* Pt_{n-2}, Pt_{n-1} */
    if (m <= nmax) out[t] = xnorm_out(pm, e);
    for (int n = m + 1; n <= nmax; ++n) {
      double back = (n - 1 > m) ? eps_nm(n - 1, m) * p2 : 0.0; /* eps_m^m = 0 */
      double pn = (x * p1 - back) / eps_nm(n, m);
      if (e < 0 && fabs(pn) >= XHI) { pn *= 0x1p-960; p1 *= 0x1p-960; e += 1; }
      out[(size_t)(n - m) * npt + t] = xnorm_out(pn, e);
      p2 = p1;
      p1 = pn;
    }
  }
}

void orc_dlegendre_column(int nmax, int m, int npt, const double *mu,
                          const double *coslat, double *out) {
  /* (1-mu^2) dPt_n^m/dmu = (n+1) eps_n^m Pt_{n-1}^m - n eps_{n+1}^m Pt_{n+1}^m */
  int len = nmax + 2 - m;
  double *P = (double *)malloc(sizeof(double) * (size_t)len * npt);
  orc_legendre_column(nmax + 1, m, npt, mu, coslat, P);
  for (int n = m; n <= nmax; ++n)
    for (int t = 0; t < npt; ++t) {
      double lo = (n > m) ? (n + 1.0) * eps_nm(n, m) * P[(size_t)(n - 1 - m) * npt + t] : 0.0;
      double hi = n * eps_nm(n + 1, m) * P[(size_t)(n + 1 - m) * npt + t];
      out[(size_t)(n - m) * npt + t] = lo - hi;
    }
  free(P);
}

/* ------------------------------------------------------------------------ */
/* Ring transforms.                                                           */

static void twiddle_table(int L, double *c, double *s) {
  for (int k = 0; k < L; ++k) {
    double a = 2.0 * M_PI * ((double)k / (double)L);
    c[k] = cos(a);
    s[k] = sin(a);
  }
}

void orc_dft_fwd(int im, int mmax, const double *x, double *F) {
  double *c = (double *)malloc(sizeof(double) * im), *s = (double *)malloc(sizeof(double) * im);
  twiddle_table(im, c, s);
  for (int m = 0; m <= mmax; ++m) {
    double re = 0, imv = 0;
    for (int i = 0; i < im; ++i) {
      int k = (int)(((long)m * i) % im);
      re += x[i] * c[k];
      imv -= x[i] * s[k];
    }
    F[2 * m] = re / im;
    F[2 * m + 1] = (m == 0) ? 0.0 : imv / im;
  }
  free(c);
  free(s);
}

void orc_dft_inv(int im, int mmax, const double *F, double *x) {
  double *c = (double *)malloc(sizeof(double) * im), *s = (double *)malloc(sizeof(double) * im);
  twiddle_table(im, c, s);
  for (int i = 0; i < im; ++i) {
    double v = F[0];
    for (int m = 1; m <= mmax; ++m) {
      int k = (int)(((long)m * i) % im);
      v += 2.0 * (F[2 * m] * c[k] - F[2 * m + 1] * s[k]);
    }
    x[i] = v;
  }
  free(c);
  free(s);
}

/* In-place iterative radix-2 complex FFT, sign -1 forward, +1 inverse,
 * unnormalised.  c/s: cos/sin(2 pi k/L) for k < L. */
static void cfft(int L, double *re, double *imv, int sign, const double *c, const double *s) {
  for (int i = 1, j = 0; i < L; ++i) {
    int bit = L >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = imv[i]; imv[i] = imv[j]; imv[j] = t;
    }
  }
  for (int len = 2; len <= L; len <<= 1) {
    int step = L / len;
    for (int i = 0; i < L; i += len)
      for (int k = 0; k < len / 2; ++k) {
        double wr = c[k * step], wi = sign * s[k * step];
        int a = i + k, b = i + k + len / 2;
        double tr = re[b] * wr - imv[b] * wi;
        double ti = re[b] * wi + imv[b] * wr;
        re[b] = re[a] - tr;
        imv[b] = imv[a] - ti;
        re[a] += tr;
        imv[a] += ti;
      }
  }
}

/* Two real rings through one complex FFT: z = xa + i xb. */
static void ring_pair_fwd(int L, int mmax, const double *xa, const double *xb,
                          double *Fa, double *Fb, double *wre, double *wim,
                          const double *c, const double *s) {
  for (int i = 0; i < L; ++i) { wre[i] = xa[i]; wim[i] = xb ? xb[i] : 0.0; }
  cfft(L, wre, wim, -1, c, s);
  for (int m = 0; m <= mmax; ++m) {
    int q = (L - m) % L;
    double zr = wre[m], zi = wim[m], cr = wre[q], ci = -wim[q]; /* conj Z_{L-m} */
    Fa[2 * m] = 0.5 * (zr + cr) / L;
    Fa[2 * m + 1] = 0.5 * (zi + ci) / L;
    if (Fb) {
      /* (Z - conj Z_{L-m}) / (2i) */
      Fb[2 * m] = 0.5 * (zi - ci) / L;
      Fb[2 * m + 1] = -0.5 * (zr - cr) / L;
    }
  }
  Fa[1] = 0.0;
  if (Fb) Fb[1] = 0.0;
}

static void ring_pair_inv(int L, int mmax, const double *Fa, const double *Fb,
                          double *xa, double *xb, double *wre, double *wim,
                          const double *c, const double *s) {
  memset(wre, 0, sizeof(double) * L);
  memset(wim, 0, sizeof(double) * L);
  /* Z_m = Xa_m + i Xb_m ; Z_{L-m} = conj Xa_m + i conj Xb_m ; Im of m=0 ignored */
  wre[0] = Fa[0];
  wim[0] = Fb ? Fb[0] : 0.0;
  for (int m = 1; m <= mmax; ++m) {
    double ar = Fa[2 * m], ai = Fa[2 * m + 1];
    double br = Fb ? Fb[2 * m] : 0.0, bi = Fb ? Fb[2 * m + 1] : 0.0;
    wre[m] = ar - bi;
    wim[m] = ai + br;
    wre[L - m] = ar + bi;
    wim[L - m] = -ai + br;
  }
  cfft(L, wre, wim, +1, c, s);
  for (int i = 0; i < L; ++i) {
    xa[i] = wre[i];
    if (xb) xb[i] = wim[i];
  }
}

static int is_pow2(int L) { return L > 0 && (L & (L - 1)) == 0; }

void orc_fft_fwd(int im, int mmax, const double *x, double *F) {
  double *buf = (double *)malloc(sizeof(double) * 4 * im);
  twiddle_table(im, buf + 2 * im, buf + 3 * im);
  ring_pair_fwd(im, mmax, x, NULL, F, NULL, buf, buf + im, buf + 2 * im, buf + 3 * im);
  free(buf);
}

void orc_fft_inv(int im, int mmax, const double *F, double *x) {
  double *buf = (double *)malloc(sizeof(double) * 4 * im);
  twiddle_table(im, buf + 2 * im, buf + 3 * im);
  ring_pair_inv(im, mmax, F, NULL, x, NULL, buf, buf + im, buf + 2 * im, buf + 3 * im);
  free(buf);
}

/* ------------------------------------------------------------------------ */
/* Plan                                                                        */

struct orc_plan {
  int N, im, jm, P2, K;
  double *mu, *w, *coslat;
  double *P;  /* P[poff[m] + (n-m)*P2 + p], n = m..N+1, at north node of pair p */
  double *H;  /* H[hoff[m] + (n-m)*P2 + p], n = m..N */
  size_t *poff, *hoff;
  double *c, *s; /* twiddles */
};

int orc_plan_ncoef(const orc_plan *p) { return p->K; }

static inline size_t coef_off(int N, int m) { return (size_t)m * (N + 1) - (size_t)m * (m - 1) / 2; }

orc_plan *orc_plan_create(int nmax, int im, int jm) {
  if (nmax < 0 || jm < 2 || (jm & 1) || im < 2 * nmax + 1 || jm <= nmax) return NULL;
  orc_plan *p = (orc_plan *)calloc(1, sizeof(orc_plan));
  p->N = nmax; p->im = im; p->jm = jm; p->P2 = jm / 2;
  p->K = (nmax + 1) * (nmax + 2) / 2;
  p->mu = (double *)malloc(sizeof(double) * jm);
  p->w = (double *)malloc(sizeof(double) * jm);
  p->coslat = (double *)malloc(sizeof(double) * jm);
  orc_gauss(jm, p->mu, p->w, p->coslat);
  p->poff = (size_t *)malloc(sizeof(size_t) * (nmax + 2));
  p->hoff = (size_t *)malloc(sizeof(size_t) * (nmax + 2));
  size_t tp = 0, th = 0;
  for (int m = 0; m <= nmax; ++m) {
    p->poff[m] = tp; tp += (size_t)(nmax + 2 - m) * p->P2;
    p->hoff[m] = th; th += (size_t)(nmax + 1 - m) * p->P2;
  }
  p->P = (double *)malloc(sizeof(double) * tp);
  p->H = (double *)malloc(sizeof(double) * th);
  const double *mun = p->mu + p->P2, *csn = p->coslat + p->P2; /* northern nodes */
  for (int m = 0; m <= nmax; ++m) {
    orc_legendre_column(nmax + 1, m, p->P2, mun, csn, p->P + p->poff[m]);
    orc_dlegendre_column(nmax, m, p->P2, mun, csn, p->H + p->hoff[m]);
  }
  p->c = (double *)malloc(sizeof(double) * im);
  p->s = (double *)malloc(sizeof(double) * im);
  twiddle_table(im, p->c, p->s);
  return p;
}

void orc_plan_destroy(orc_plan *p) {
  if (!p) return;
  free(p->mu); free(p->w); free(p->coslat); free(p->P); free(p->H);
  free(p->poff); free(p->hoff); free(p->c); free(p->s);
  free(p);
}

static void ring_pair_fwd_any(const orc_plan *p, const double *xa, const double *xb,
                              double *Fa, double *Fb, double *wre, double *wim) {
  if (is_pow2(p->im)) {
    ring_pair_fwd(p->im, p->N, xa, xb, Fa, Fb, wre, wim, p->c, p->s);
  } else {
    orc_dft_fwd(p->im, p->N, xa, Fa);
    orc_dft_fwd(p->im, p->N, xb, Fb);
  }
}

static void ring_pair_inv_any(const orc_plan *p, const double *Fa, const double *Fb,
                              double *xa, double *xb, double *wre, double *wim) {
  if (is_pow2(p->im)) {
    ring_pair_inv(p->im, p->N, Fa, Fb, xa, xb, wre, wim, p->c, p->s);
  } else {
    orc_dft_inv(p->im, p->N, Fa, xa);
    orc_dft_inv(p->im, p->N, Fb, xb);
  }
}

/* spectral -> grid.  Fourier buffer four[f][j][m] complex. */
void orc_spec_to_grid(const orc_plan *p, int nf, const double *spec, double *grid,
                      int kind, int nthreads) {
  const int N = p->N, jm = p->jm, P2 = p->P2, im = p->im, M1 = N + 1;
  double *four = (double *)calloc((size_t)nf * jm * M1 * 2, sizeof(double));
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  /* Legendre step: per m, per latitude pair, even/odd partial sums over n. */
#pragma omp parallel num_threads(nthreads)
  {
    double *acc = (double *)malloc(sizeof(double) * 4 * nf);
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m <= N; ++m) {
      const double *s_m = spec + coef_off(N, m) * nf * 2;
      for (int q = 0; q < P2; ++q) {
        memset(acc, 0, sizeof(double) * 4 * nf);
        for (int n = m; n <= N; ++n) {
          double t = (kind == ORC_GRADLAT) ? p->H[p->hoff[m] + (size_t)(n - m) * P2 + q]
                                           : p->P[p->poff[m] + (size_t)(n - m) * P2 + q];
          double *a = acc + ((n - m) & 1) * 2 * nf;
          const double *sv = s_m + (size_t)(n - m) * nf * 2;
          for (int f = 0; f < 2 * nf; ++f) a[f] += t * sv[f];
        }
        int jn = P2 + q, js = P2 - 1 - q;
        double rs = (kind == ORC_VALUE) ? 1.0 : 1.0 / p->coslat[jn];
        for (int f = 0; f < nf; ++f) {
          double er = acc[2 * f], ei = acc[2 * f + 1];
          double orr = acc[2 * nf + 2 * f], oi = acc[2 * nf + 2 * f + 1];
          double nr = er + orr, ni = ei + oi, sr, si;
          if (kind == ORC_GRADLAT) { sr = orr - er; si = oi - ei; } /* H has flipped parity */
          else { sr = er - orr; si = ei - oi; }
          if (kind == ORC_GRADLON) { /* multiply by i m */
            double t;
            t = nr; nr = -m * ni; ni = m * t;
            t = sr; sr = -m * si; si = m * t;
          }
          double *fn = four + (((size_t)f * jm + jn) * M1 + m) * 2;
          double *fs = four + (((size_t)f * jm + js) * M1 + m) * 2;
          fn[0] = nr * rs; fn[1] = ni * rs;
          fs[0] = sr * rs; fs[1] = si * rs;
        }
      }
    }
    free(acc);
  }
  /* Ring synthesis, two rings (the N/S pair) per complex FFT. */
#pragma omp parallel num_threads(nthreads)
  {
    double *wb = (double *)malloc(sizeof(double) * 2 * im);
#pragma omp for schedule(static)
    for (long t = 0; t < (long)nf * P2; ++t) {
      int f = (int)(t / P2), q = (int)(t % P2);
      int jn = P2 + q, js = P2 - 1 - q;
      ring_pair_inv_any(p, four + ((size_t)f * jm + jn) * M1 * 2, four + ((size_t)f * jm + js) * M1 * 2,
                        grid + ((size_t)f * jm + jn) * im, grid + ((size_t)f * jm + js) * im, wb, wb + im);
    }
    free(wb);
  }
  free(four);
}

/* grid -> spectral. */
void orc_grid_to_spec(const orc_plan *p, int nf, const double *grid, double *spec, int nthreads) {
  const int N = p->N, jm = p->jm, P2 = p->P2, im = p->im, M1 = N + 1;
  double *four = (double *)malloc(sizeof(double) * (size_t)nf * jm * M1 * 2);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
#pragma omp parallel num_threads(nthreads)
  {
    double *wb = (double *)malloc(sizeof(double) * 2 * im);
#pragma omp for schedule(static)
    for (long t = 0; t < (long)nf * P2; ++t) {
      int f = (int)(t / P2), q = (int)(t % P2);
      int jn = P2 + q, js = P2 - 1 - q;
      ring_pair_fwd_any(p, grid + ((size_t)f * jm + jn) * im, grid + ((size_t)f * jm + js) * im,
                        four + ((size_t)f * jm + jn) * M1 * 2, four + ((size_t)f * jm + js) * M1 * 2, wb, wb + im);
    }
    free(wb);
  }
  /* Quadrature: s_n^m = sum_q (w_q/2) Pt_n^m(mu_q) [F(mu_q) +- F(-mu_q)] */
#pragma omp parallel num_threads(nthreads)
  {
    double *sa = (double *)malloc(sizeof(double) * (size_t)P2 * 4 * nf);
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m <= N; ++m) {
      for (int q = 0; q < P2; ++q) {
        int jn = P2 + q, js = P2 - 1 - q;
        double *sym = sa + (size_t)q * 4 * nf, *asy = sym + 2 * nf;
        for (int f = 0; f < nf; ++f) {
          const double *fn = four + (((size_t)f * jm + jn) * M1 + m) * 2;
          const double *fs = four + (((size_t)f * jm + js) * M1 + m) * 2;
          sym[2 * f] = fn[0] + fs[0]; sym[2 * f + 1] = fn[1] + fs[1];
          asy[2 * f] = fn[0] - fs[0]; asy[2 * f + 1] = fn[1] - fs[1];
        }
      }
      double *s_m = spec + coef_off(N, m) * nf * 2;
      for (int n = m; n <= N; ++n) {
        double *sv = s_m + (size_t)(n - m) * nf * 2;
        memset(sv, 0, sizeof(double) * 2 * nf);
        int par = (n - m) & 1;
        for (int q = 0; q < P2; ++q) {
          double t = 0.5 * p->w[P2 + q] * p->P[p->poff[m] + (size_t)(n - m) * P2 + q];
          const double *src = sa + (size_t)q * 4 * nf + par * 2 * nf;
          for (int f = 0; f < 2 * nf; ++f) sv[f] += t * src[f];
        }
        if (m == 0)
          for (int f = 0; f < nf; ++f) sv[2 * f + 1] = 0.0;
      }
    }
    free(sa);
  }
  free(four);
}

/* grid -> spectral of the flux divergence
 *   (1/cos phi) dA/dlon + (1/cos phi) d(B cos phi)/dphi
 * (SPML wa_DivLon_xya for A alone, wa_DivLat_xya for B alone).  The lon term
 * is i m A^m / cos phi, analysed like any field; the lat term is integrated by
 * parts over the sphere (the boundary term vanishes with cos phi at the poles):
 *   s_n^m = -sum_q (w_q/2) H_n^m(mu_q) B^m(mu_q) / cos phi_q,
 *   H = (1-mu^2) dPt/dmu,
 * whose integrand is a polynomial for band-limited B, so the Gauss sum is exact.
 * H(-mu) = -(-1)^(n-m) H(mu): an even n-m pairs with the antisymmetric part of
 * B^m over the mirrored latitudes, an odd n-m with the symmetric part. */
void orc_grid_to_div(const orc_plan *p, int nf, const double *ga, const double *gb,
                     double *spec, int nthreads) {
  const int N = p->N, jm = p->jm, P2 = p->P2, im = p->im, M1 = N + 1;
  const double *g[2] = {ga, gb};
  double *four[2] = {NULL, NULL};
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  for (int t = 0; t < 2; ++t) {
    if (!g[t]) continue;
    four[t] = (double *)malloc(sizeof(double) * (size_t)nf * jm * M1 * 2);
    const double *grid = g[t];
    double *fo = four[t];
#pragma omp parallel num_threads(nthreads)
    {
      double *wb = (double *)malloc(sizeof(double) * 2 * im);
#pragma omp for schedule(static)
      for (long u = 0; u < (long)nf * P2; ++u) {
        int f = (int)(u / P2), q = (int)(u % P2);
        int jn = P2 + q, js = P2 - 1 - q;
        ring_pair_fwd_any(p, grid + ((size_t)f * jm + jn) * im, grid + ((size_t)f * jm + js) * im,
                          fo + ((size_t)f * jm + jn) * M1 * 2, fo + ((size_t)f * jm + js) * M1 * 2,
                          wb, wb + im);
      }
      free(wb);
    }
  }
#pragma omp parallel num_threads(nthreads)
  {
    /* per pair q: [A sym, A asy, B sym, B asy] x 2 nf */
    double *sa = (double *)malloc(sizeof(double) * (size_t)P2 * 8 * nf);
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m <= N; ++m) {
      for (int q = 0; q < P2; ++q) {
        int jn = P2 + q, js = P2 - 1 - q;
        double rc = 1.0 / p->coslat[jn];
        double *blk = sa + (size_t)q * 8 * nf;
        for (int t = 0; t < 2; ++t) {
          double *sym = blk + (size_t)t * 4 * nf, *asy = sym + 2 * nf;
          for (int f = 0; f < nf; ++f) {
            double nr = 0, ni = 0, sr = 0, si = 0;
            if (four[t]) {
              const double *fn = four[t] + (((size_t)f * jm + jn) * M1 + m) * 2;
              const double *fs = four[t] + (((size_t)f * jm + js) * M1 + m) * 2;
              nr = fn[0] * rc; ni = fn[1] * rc; sr = fs[0] * rc; si = fs[1] * rc;
              if (t == 0) { /* times i m */
                double x;
                x = nr; nr = -m * ni; ni = m * x;
                x = sr; sr = -m * si; si = m * x;
              }
            }
            sym[2 * f] = nr + sr; sym[2 * f + 1] = ni + si;
            asy[2 * f] = nr - sr; asy[2 * f + 1] = ni - si;
          }
        }
      }
      double *s_m = spec + coef_off(N, m) * nf * 2;
      for (int n = m; n <= N; ++n) {
        double *sv = s_m + (size_t)(n - m) * nf * 2;
        memset(sv, 0, sizeof(double) * 2 * nf);
        int par = (n - m) & 1;
        for (int q = 0; q < P2; ++q) {
          double hw = 0.5 * p->w[P2 + q];
          double tp = hw * p->P[p->poff[m] + (size_t)(n - m) * P2 + q];
          double th = -hw * p->H[p->hoff[m] + (size_t)(n - m) * P2 + q];
          const double *blk = sa + (size_t)q * 8 * nf;
          const double *a = blk + par * 2 * nf;             /* A: same parity      */
          const double *b = blk + 4 * nf + (1 - par) * 2 * nf; /* B: opposite parity */
          for (int f = 0; f < 2 * nf; ++f) sv[f] += tp * a[f] + th * b[f];
        }
        if (m == 0)
          for (int f = 0; f < nf; ++f) sv[2 * f + 1] = 0.0;
      }
    }
    free(sa);
  }
  free(four[0]);
  free(four[1]);
}

/* Powers of the unit-sphere Laplacian on canonical spectra: out = (-n(n+1))^power in
 * (SPML wa_Lapla_wa: power 1; wa_LaplaInv_wa: power -1).  Negative powers set
 * n = 0 to zero.  in and out may alias. */
void orc_spec_lapla(int nmax, int nf, const double *in, double *out, int power) {
  for (int m = 0; m <= nmax; ++m)
    for (int n = m; n <= nmax; ++n) {
      const double e = -(double)n * (n + 1);
      double fac = 1.0;
      if (power < 0 && n == 0) fac = 0.0;
      else
        for (int k = 0; k < (power < 0 ? -power : power); ++k) fac = power < 0 ? fac / e : fac * e;
      const size_t o = (size_t)(coef_off(nmax, m) + n - m) * nf * 2;
      for (int f = 0; f < 2 * nf; ++f) out[o + f] = fac * in[o + f];
    }
}

/* Coupler remap, the loop of interpolate_data_latlon
 * (common/interpolation_data_latlon_mod.f90:274-306) with 0-based indices:
 * recv(:, d) = 0; then for each d and each triplet i in order,
 * recv(r_i, d) = recv(r_i, d) + send(s_i, d) * coef_i. */
void orc_remap(const int *send_index, const int *recv_index, const double *coef, int nnz,
               int nsend, int nrecv, int ndata, const double *send, double *recv) {
  for (long t = 0; t < (long)nrecv * ndata; ++t) recv[t] = 0.0;
  for (int d = 0; d < ndata; ++d)
    for (int i = 0; i < nnz; ++i) {
      double *r = recv + (size_t)d * nrecv + recv_index[i];
      *r = *r + send[(size_t)d * nsend + send_index[i]] * coef[i];
    }
}

/* ------------------------------------------------------------------------ */

void orc_spml_pack(int nmax, int nf, const double *spec, double *wa) {
  const int L = (nmax + 1) * (nmax + 1);
  for (int f = 0; f < nf; ++f)
    for (int m = 0; m <= nmax; ++m)
      for (int n = m; n <= nmax; ++n) {
        const double *s = spec + ((coef_off(nmax, m) + n - m) * nf + f) * 2;
        double *w = wa + (size_t)f * L;
        if (m == 0) {
          w[n * (n + 1)] = s[0];
        } else {
          w[n * (n + 1) + m] = M_SQRT2 * s[0];
          w[n * (n + 1) - m] = -M_SQRT2 * s[1];
        }
      }
}

void orc_spml_unpack(int nmax, int nf, const double *wa, double *spec) {
  const int L = (nmax + 1) * (nmax + 1);
  for (int f = 0; f < nf; ++f)
    for (int m = 0; m <= nmax; ++m)
      for (int n = m; n <= nmax; ++n) {
        double *s = spec + ((coef_off(nmax, m) + n - m) * nf + f) * 2;
        const double *w = wa + (size_t)f * L;
        if (m == 0) {
          s[0] = w[n * (n + 1)];
          s[1] = 0.0;
        } else {
          s[0] = w[n * (n + 1) + m] / M_SQRT2;
          s[1] = -w[n * (n + 1) - m] / M_SQRT2;
        }
      }
}

void orc_avr_lonlat(const orc_plan *p, int nf, const double *grid, double *avg) {
  for (int f = 0; f < nf; ++f) {
    double acc = 0;
    for (int j = 0; j < p->jm; ++j) {
      double r = 0;
      for (int i = 0; i < p->im; ++i) r += grid[((size_t)f * p->jm + j) * p->im + i];
      acc += p->w[j] * r;
    }
    avg[f] = acc / (2.0 * p->im); /* (1/4pi) * (2pi/im) * sum w f */
  }
}

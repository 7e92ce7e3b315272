/*
 * sht_oracle.h -- CPU FP64 restatement of the Gaussian-grid spherical-harmonic
 * transform (SHT) that SPML's wa_module provides to DCPAM's dynamical core.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product path never links or calls anything here.
 *
 * Conventions (shared with the HIP library, include/sht.h):
 *   grid   g[f][j][i]     f = field (level), j = 0..jm-1 Gauss latitudes S->N,
 *                         i = 0..im-1 longitudes lambda_i = 2*pi*i/im
 *                         (Fortran xya(0:im-1,1:jm,1:k), dcpam_main_mod.f90:4107-4125)
 *   spectra (canonical)   complex s_n^m, m = 0..N, n = m..N, stored
 *                         s[((off(m) + n - m) * nf + f) * 2 + {0 re, 1 im}],
 *                         off(m) = m*(N+1) - m*(m-1)/2
 *   synthesis  f(lambda, mu) = sum_n sum_{m=-n..n} s_n^m Pt_n^|m|(mu) e^{i m lambda},
 *              s_n^{-m} = conj(s_n^m), Im s_n^0 ignored (produced as 0)
 *   Pt_n^m     (1/2) int_{-1}^{1} Pt^2 dmu = 1, no Condon-Shortley phase; the
 *              zonal case Pt_n = sqrt(2n+1) P_n is the normalisation
 *              EBM_main.f90:170-174 and :242-246 rely on.
 */
#ifndef SHT_ORACLE_H
#define SHT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Legendre nodes mu_j (ascending, S->N), weights w_j (sum 2) and
 * cos(latitude) = sqrt(1-mu^2) computed from the colatitude directly. */
int orc_gauss(int jm, double *mu, double *w, double *coslat);

/* Pt_n^m(mu_t) for n = m..nmax at npt arbitrary points with cos(lat) given
 * (coslat = sqrt(1-mu^2) supplied separately for polar accuracy).
 * out[(n-m)*npt + t].  Extended-exponent recurrence: no spurious underflow. */
void orc_legendre_column(int nmax, int m, int npt, const double *mu,
                         const double *coslat, double *out);

/* (1 - mu^2) dPt_n^m/dmu  = cos(lat) dPt_n^m/dlat, same layout. */
void orc_dlegendre_column(int nmax, int m, int npt, const double *mu,
                          const double *coslat, double *out);

/* Real DFTs along one ring.  F[m] complex (2*(mmax+1) doubles),
 * F_m = (1/im) sum_i x_i e^{-2 pi i m i / im}.  Direct O(im*mmax) form. */
void orc_dft_fwd(int im, int mmax, const double *x, double *F);
/* x_i = sum_{m=-mmax..mmax} F_m e^{+2 pi i m i/im} with F_{-m} = conj F_m. */
void orc_dft_inv(int im, int mmax, const double *F, double *x);
/* Same two maps through an iterative radix-2 FFT (im a power of two). */
void orc_fft_fwd(int im, int mmax, const double *x, double *F);
void orc_fft_inv(int im, int mmax, const double *F, double *x);

/* Transform plan: Gauss grid, Legendre tables and FFT twiddles. */
typedef struct orc_plan orc_plan;
orc_plan *orc_plan_create(int nmax, int im, int jm);
void orc_plan_destroy(orc_plan *p);
int orc_plan_ncoef(const orc_plan *p); /* K = (N+1)(N+2)/2 */

/* kind for orc_spec_to_grid */
enum { ORC_VALUE = 0, ORC_GRADLON = 1, ORC_GRADLAT = 2 };

/* spectral (canonical) -> grid, nf fields; kind selects f, (1/cos)df/dlon
 * or df/dlat.  nthreads <= 0 means "OpenMP default". */
void orc_spec_to_grid(const orc_plan *p, int nf, const double *spec,
                      double *grid, int kind, int nthreads);
/* grid -> spectral (canonical), nf fields */
void orc_grid_to_spec(const orc_plan *p, int nf, const double *grid,
                      double *spec, int nthreads);

/* grid -> spectral of (1/cos phi) dA/dlon + (1/cos phi) d(B cos phi)/dphi
 * (SPML wa_DivLon_xya, wa_DivLat_xya and their sum); A or B may be NULL. */
void orc_grid_to_div(const orc_plan *p, int nf, const double *a, const double *b,
                     double *spec, int nthreads);
/* out = (-n(n+1))^power in on canonical spectra (negative powers: n = 0 -> 0) */
void orc_spec_lapla(int nmax, int nf, const double *in, double *out, int power);

/* Coupler remap (interpolate_data_latlon), 0-based triplets, fields
 * contiguous per point: send[ndata][nsend] -> recv[ndata][nrecv]. */
void orc_remap(const int *send_index, const int *recv_index, const double *coef, int nnz,
               int nsend, int nrecv, int ndata, const double *send, double *recv);

/* Canonical complex <-> real (N+1)^2 packing wa[f][l], l = n(n+1)+m,
 * m = -n..n; wa(n,0) = Re s, wa(n,m>0) = sqrt2 Re s, wa(n,-m) = -sqrt2 Im s
 * (orthonormal real harmonics with unit mean square). */
void orc_spml_pack(int nmax, int nf, const double *spec, double *wa);
void orc_spml_unpack(int nmax, int nf, const double *wa, double *spec);

/* Global mean (1/4pi) sum_ij (2pi/im) w_j f_ij for each field. */
void orc_avr_lonlat(const orc_plan *p, int nf, const double *grid, double *avg);

#ifdef __cplusplus
}
#endif
#endif

/*
 * f64 CPU oracle for the spectral IMEX step — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * entry point and the cpu_baseline leg of bench.py may load it.  The GPU
 * library (libtsphere) never links or calls it.
 *
 * It restates, in plain C and double precision, the reference's own CPU
 * algorithm for the path:
 *   src/sht/layout.ts   (spectral packing, config validation, grid rule)
 *   src/sht/gauss.ts    (Gauss-Legendre nodes/weights by Newton on P_n)
 *   src/sht/coeffs.ts   (orthonormal Condon-Shortley recurrence coefficients)
 *   src/sht/reference.ts(direct-summation synth/analys, xorshift32 spectrum)
 *   src/mgpu/noise.ts   (mulberry32 uniform deviates)
 *   src/mgpu/model.ts   (Laplace-Beltrami eigenvalues l(l+1))
 *   models/ + bench/shtns/spec.h:146-169 (reaction and IMEX arithmetic)
 * Each function cites the reference lines it follows.
 */
#ifndef TSPHERE_SHT_ORACLE_H
#define TSPHERE_SHT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* layout.ts:24-27 */
long long orc_nlm_calc(int lmax, int mmax);
/* layout.ts:30-32 */
long long orc_lm_index(int lmax, int l, int m);
/* layout.ts:34-43: 0 if valid, else 1..4 naming the failed rule
 * (1 lmax, 2 mmax, 3 nlat, 4 nphi). */
int orc_validate_config(int lmax, int mmax, int nlat, int nphi);
/* layout.ts:53-59 */
void orc_grid_for_lmax(int lmax, int pdeg, int *nlat, int *nphi);

/* gauss.ts:11-51: x decreasing (north first), w for integration over x. */
void orc_gauss_nodes_weights(int n, double *x, double *w);

/* coeffs.ts:28-56: amm[mmax+1], ab[2*nlm] = (a_l^m, b_l^m). */
void orc_legendre_coeffs(int lmax, int mmax, double *amm, double *ab);
/* coeffs.ts:64-85: out[l-m] = ytilde_l^m(theta) for l=m..lmax. */
void orc_legendre_row(const double *amm, const double *ab, int lmax, int m, double ct,
                      double st, double *out);

/* reference.ts:37-57: fm[2*(m*nlat+i)+{0,1}] = sum_l Q_lm ytilde_l^m(theta_i). */
void orc_legendre_synth(int lmax, int mmax, int nlat, const double *qlm, double *fm);
/* reference.ts:60-78 (direct Fourier sum). rows [row0,row1) only, for bounded
 * timing samples; pass 0,nlat for the full transform. */
void orc_synth_rows(int lmax, int mmax, int nlat, int nphi, const double *qlm, double *spat,
                    int row0, int row1);
void orc_synth(int lmax, int mmax, int nlat, int nphi, const double *qlm, double *spat);
/* reference.ts:85-101: gm[2*(m*nlat+i)+{0,1}] = (2pi/nphi) sum_j f_ij e^{-i m phi_j}. */
void orc_fourier_analys_rows(int mmax, int nlat, int nphi, const double *spat, double *gm,
                             int row0, int row1);
/* reference.ts:103-116: Gauss-weighted Legendre accumulation of gm into qlm
 * over latitudes [row0,row1) (qlm is accumulated into, not cleared). */
void orc_legendre_analys_rows(int lmax, int mmax, int nlat, const double *gm, double *qlm,
                              int row0, int row1);
/* reference.ts:81-118 */
void orc_analys(int lmax, int mmax, int nlat, int nphi, const double *spat, double *qlm);

/* reference.ts:122-136 (JS xorshift32 semantics, result rounded to f32). */
void orc_random_spectrum(int lmax, int mmax, uint32_t seed, float *q);
/* noise.ts:15-24: advances *state, returns a uniform deviate in [0,1). */
double orc_mulberry32(uint32_t *state);
/* n successive mulberry32 deviates from `seed` into out (f64). */
void orc_mulberry32_fill(uint32_t seed, long long n, double *out);
/* model.ts:46-56 */
void orc_eigenvalues(int lmax, int mmax, float *lam);

/* Reaction (the models/ reaction lines, spec.h:146-161), one grid point, f64.
 * model 0 schnakenberg (p0=a, p1=b), 1 brusselator (p0=A, p1=B),
 * 2 allencahn (r2 untouched). */
void orc_react(int model, double p0, double p1, double u, double v, double *r1, double *r2);
/* IMEX diagonal solve (spec.h:166-169): (U + dt*R) / (1 + dtD*lam). */
double orc_imex(double dt, double dtD, double U, double R, double lam);

/* Threads used by the *_rows loops (OpenMP); 0 = library default. */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif

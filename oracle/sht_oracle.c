/*
 * f64 CPU oracle — TEST INFRASTRUCTURE ONLY (see sht_oracle.h).
 *
 * Plain restatement of the reference's CPU algorithm, loop for loop, in
 * double precision.  OpenMP only splits the outer latitude / m loops; the
 * arithmetic inside each (latitude, m) is the reference's.
 */
#include "sht_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* layout.ts:24-27 */
long long orc_nlm_calc(int lmax, int mmax) {
  return (long long)(mmax + 1) * (lmax + 1) - ((long long)mmax * (mmax + 1)) / 2;
}

/* layout.ts:30-32 */
long long orc_lm_index(int lmax, int l, int m) {
  return (long long)m * (lmax + 1) - ((long long)m * (m - 1)) / 2 + (l - m);
}

/* layout.ts:34-43 */
int orc_validate_config(int lmax, int mmax, int nlat, int nphi) {
  if (lmax < 1) return 1;
  if (mmax < 0 || mmax > lmax) return 2;
  if (nlat <= lmax) return 3;
  if (nphi < 2 * mmax + 1) return 4;
  return 0;
}

/* layout.ts:53-59 */
void orc_grid_for_lmax(int lmax, int pdeg, int *nlat, int *nphi) {
  double min_lat = ((double)(pdeg + 1) * lmax + 1) / 2.0;
  if (min_lat < lmax + 1) min_lat = lmax + 1;
  *nlat = 2 * (int)ceil(min_lat / 2.0);
  int n = 1;
  while (n < (pdeg + 1) * lmax + 1) n *= 2;
  *nphi = n;
}

/* gauss.ts:11-51: Newton on P_n from the Tricomi-like guess, tolerance 1e-15,
 * then one polishing iteration. */
void orc_gauss_nodes_weights(int n, double *x, double *w) {
  const int half = (n + 1) >> 1;
  for (int i = 0; i < half; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int iter = 0; iter < 100; iter++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double dz = p1 / pp;
      z -= dz;
      if (fabs(dz) < 1e-15 * fabs(z) + 1e-300) {
        double q1 = 1.0, q2 = 0.0;
        for (int j = 1; j <= n; j++) {
          double q3 = q2;
          q2 = q1;
          q1 = ((2.0 * j - 1.0) * z * q2 - (j - 1.0) * q3) / j;
        }
        pp = n * (z * q1 - q2) / (z * z - 1.0);
        z -= q1 / pp;
        break;
      }
    }
    x[i] = z;
    x[n - 1 - i] = -z;
    double wi = 2.0 / ((1.0 - z * z) * pp * pp);
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n & 1) x[half - 1] = 0.0;
}

/* coeffs.ts:28-56 */
void orc_legendre_coeffs(int lmax, int mmax, double *amm, double *ab) {
  const long long nlm = orc_nlm_calc(lmax, mmax);
  memset(ab, 0, sizeof(double) * 2 * (size_t)nlm);
  double t = 1.0 / (4.0 * M_PI);
  amm[0] = sqrt(t);
  for (int m = 1; m <= mmax; m++) {
    t *= (2.0 * m + 1.0) / (2.0 * m);
    amm[m] = -sqrt(t);
    if (m % 2 == 0) amm[m] = -amm[m];
  }
  for (int m = 0; m <= mmax; m++) {
    if (m + 1 <= lmax) {
      long long lm = orc_lm_index(lmax, m + 1, m);
      ab[2 * lm] = sqrt(2.0 * m + 3.0);
      ab[2 * lm + 1] = 0.0;
    }
    for (int l = m + 2; l <= lmax; l++) {
      long long lm = orc_lm_index(lmax, l, m);
      double t1 = (double)(l + m) * (l - m);
      double t2 = (double)(l - 1 + m) * (l - 1 - m);
      ab[2 * lm] = sqrt(((2.0 * l + 1.0) * (2.0 * l - 1.0)) / t1);
      ab[2 * lm + 1] = -sqrt(((2.0 * l + 1.0) / (2.0 * l - 3.0)) * (t2 / t1));
    }
  }
}

/* coeffs.ts:64-85 */
void orc_legendre_row(const double *amm, const double *ab, int lmax, int m, double ct,
                      double st, double *out) {
  double y0 = amm[m] * pow(st, m);
  out[0] = y0;
  if (m == lmax) return;
  const long long base = orc_lm_index(lmax, m, m);
  double y1 = ab[2 * (base + 1)] * ct * y0;
  out[1] = y1;
  for (int l = m + 2; l <= lmax; l++) {
    long long lm = base + (l - m);
    double y2 = ab[2 * lm] * ct * y1 + ab[2 * lm + 1] * y0;
    y0 = y1;
    y1 = y2;
    out[l - m] = y2;
  }
}

/* The ShtReference constructor's tables (reference.ts:21-31). */
struct orc_tables {
  double *ct, *st, *wg, *amm, *ab;
};

static void tables_make(struct orc_tables *t, int lmax, int mmax, int nlat) {
  const long long nlm = orc_nlm_calc(lmax, mmax);
  t->ct = (double *)malloc(sizeof(double) * nlat);
  t->st = (double *)malloc(sizeof(double) * nlat);
  t->wg = (double *)malloc(sizeof(double) * nlat);
  t->amm = (double *)malloc(sizeof(double) * (mmax + 1));
  t->ab = (double *)malloc(sizeof(double) * 2 * (size_t)nlm);
  orc_gauss_nodes_weights(nlat, t->ct, t->wg);
  for (int i = 0; i < nlat; i++) t->st[i] = sqrt(1.0 - t->ct[i] * t->ct[i]);
  orc_legendre_coeffs(lmax, mmax, t->amm, t->ab);
}

static void tables_free(struct orc_tables *t) {
  free(t->ct);
  free(t->st);
  free(t->wg);
  free(t->amm);
  free(t->ab);
}

#ifdef _OPENMP
#define ORC_NT (g_threads > 0 ? g_threads : omp_get_max_threads())
#endif

/* reference.ts:37-57 */
static void legendre_synth_t(const struct orc_tables *t, int lmax, int mmax, int nlat,
                             const double *qlm, double *fm, int row0, int row1) {
#ifdef _OPENMP
#pragma omp parallel num_threads(ORC_NT)
#endif
  {
    double *row = (double *)malloc(sizeof(double) * (lmax + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int i = row0; i < row1; i++) {
      for (int m = 0; m <= mmax; m++) {
        orc_legendre_row(t->amm, t->ab, lmax, m, t->ct[i], t->st[i], row);
        double re = 0.0, im = 0.0;
        const long long base = orc_lm_index(lmax, m, m);
        for (int l = m; l <= lmax; l++) {
          double y = row[l - m];
          re += y * qlm[2 * (base + l - m)];
          im += y * qlm[2 * (base + l - m) + 1];
        }
        long long o = 2LL * ((long long)m * nlat + i);
        fm[o] = re;
        fm[o + 1] = im;
      }
    }
    free(row);
  }
}

void orc_legendre_synth(int lmax, int mmax, int nlat, const double *qlm, double *fm) {
  struct orc_tables t;
  tables_make(&t, lmax, mmax, nlat);
  legendre_synth_t(&t, lmax, mmax, nlat, qlm, fm, 0, nlat);
  tables_free(&t);
}

/* reference.ts:60-78: direct Fourier sum, cos/sin of m*phi per term. */
void orc_synth_rows(int lmax, int mmax, int nlat, int nphi, const double *qlm, double *spat,
                    int row0, int row1) {
  struct orc_tables t;
  tables_make(&t, lmax, mmax, nlat);
  double *fm = (double *)calloc(2 * (size_t)(mmax + 1) * nlat, sizeof(double));
  legendre_synth_t(&t, lmax, mmax, nlat, qlm, fm, row0, row1);
#ifdef _OPENMP
#pragma omp parallel for num_threads(ORC_NT) schedule(dynamic, 1)
#endif
  for (int i = row0; i < row1; i++) {
    for (int j = 0; j < nphi; j++) {
      const double phi = (2.0 * M_PI * j) / nphi;
      double v = fm[2LL * i]; /* m = 0: real part */
      for (int m = 1; m <= mmax; m++) {
        long long o = 2LL * ((long long)m * nlat + i);
        double c = cos(m * phi);
        double s = sin(m * phi);
        v += 2.0 * (fm[o] * c - fm[o + 1] * s);
      }
      spat[(long long)i * nphi + j] = v;
    }
  }
  free(fm);
  tables_free(&t);
}

void orc_synth(int lmax, int mmax, int nlat, int nphi, const double *qlm, double *spat) {
  orc_synth_rows(lmax, mmax, nlat, nphi, qlm, spat, 0, nlat);
}

/* reference.ts:85-101 */
void orc_fourier_analys_rows(int mmax, int nlat, int nphi, const double *spat, double *gm,
                             int row0, int row1) {
  const double norm = (2.0 * M_PI) / nphi;
#ifdef _OPENMP
#pragma omp parallel for num_threads(ORC_NT) schedule(dynamic, 1)
#endif
  for (int i = row0; i < row1; i++) {
    for (int m = 0; m <= mmax; m++) {
      double re = 0.0, im = 0.0;
      for (int j = 0; j < nphi; j++) {
        const double phi = (2.0 * M_PI * j) / nphi;
        const double f = spat[(long long)i * nphi + j];
        re += f * cos(m * phi);
        im -= f * sin(m * phi);
      }
      long long o = 2LL * ((long long)m * nlat + i);
      gm[o] = re * norm;
      gm[o + 1] = im * norm;
    }
  }
}

static void legendre_analys_t(const struct orc_tables *t, int lmax, int mmax, int nlat,
                              const double *gm, double *qlm, int row0, int row1) {
#ifdef _OPENMP
#pragma omp parallel num_threads(ORC_NT)
#endif
  {
    double *row = (double *)malloc(sizeof(double) * (lmax + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int m = 0; m <= mmax; m++) {
      const long long base = orc_lm_index(lmax, m, m);
      for (int i = row0; i < row1; i++) {
        orc_legendre_row(t->amm, t->ab, lmax, m, t->ct[i], t->st[i], row);
        long long o = 2LL * ((long long)m * nlat + i);
        const double wr = t->wg[i] * gm[o];
        const double wi = t->wg[i] * gm[o + 1];
        for (int l = m; l <= lmax; l++) {
          const double y = row[l - m];
          qlm[2 * (base + l - m)] += y * wr;
          qlm[2 * (base + l - m) + 1] += y * wi;
        }
      }
    }
    free(row);
  }
}

/* reference.ts:103-116 */
void orc_legendre_analys_rows(int lmax, int mmax, int nlat, const double *gm, double *qlm,
                              int row0, int row1) {
  struct orc_tables t;
  tables_make(&t, lmax, mmax, nlat);
  legendre_analys_t(&t, lmax, mmax, nlat, gm, qlm, row0, row1);
  tables_free(&t);
}

/* reference.ts:81-118 */
void orc_analys(int lmax, int mmax, int nlat, int nphi, const double *spat, double *qlm) {
  const long long nlm = orc_nlm_calc(lmax, mmax);
  double *gm = (double *)calloc(2 * (size_t)(mmax + 1) * nlat, sizeof(double));
  memset(qlm, 0, sizeof(double) * 2 * (size_t)nlm);
  orc_fourier_analys_rows(mmax, nlat, nphi, spat, gm, 0, nlat);
  orc_legendre_analys_rows(lmax, mmax, nlat, gm, qlm, 0, nlat);
  free(gm);
}

/* reference.ts:122-136.  The JS source shifts a uint32 `s` with `>>`, which
 * first reinterprets it as int32: the middle xorshift step is therefore an
 * ARITHMETIC right shift, reproduced here on an int32 view. */
void orc_random_spectrum(int lmax, int mmax, uint32_t seed, float *q) {
  const long long nlm = orc_nlm_calc(lmax, mmax);
  uint32_t s = seed;
  for (long long k = 0; k < 2 * nlm; k++) {
    s ^= s << 13;
    int32_t t = (int32_t)s;
    t ^= (t >> 17);
    s = (uint32_t)t;
    s ^= s << 5;
    q[k] = (float)(((double)s / 4294967296.0) * 2.0 - 1.0);
  }
  for (int l = 0; l <= lmax; l++) q[2 * orc_lm_index(lmax, l, 0) + 1] = 0.0f;
}

/* noise.ts:15-24 (uint32 wrap-around equals the JS imul/>>>0 sequence) */
double orc_mulberry32(uint32_t *state) {
  *state = *state + 0x6d2b79f5u;
  uint32_t t = *state;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}

void orc_mulberry32_fill(uint32_t seed, long long n, double *out) {
  uint32_t s = seed;
  for (long long i = 0; i < n; i++) out[i] = orc_mulberry32(&s);
}

/* model.ts:46-56 */
void orc_eigenvalues(int lmax, int mmax, float *lam) {
  for (int m = 0; m <= mmax; m++) {
    for (int l = m; l <= lmax; l++) {
      long long i = orc_lm_index(lmax, l, m);
      lam[2 * i] = (float)(l * (l + 1));
      lam[2 * i + 1] = (float)(l * (l + 1));
    }
  }
}

/* models/schnakenberg.m:26-28, brusselator.m:25-27, allencahn.m:21;
 * operand order as spec.h:146-161 writes it. */
void orc_react(int model, double p0, double p1, double u, double v, double *r1, double *r2) {
  if (model == 2) {
    *r1 = u - u * u * u;
    return;
  }
  const double uuv = u * u * v;
  if (model == 0) {
    *r1 = p0 - u + uuv;
    *r2 = p1 - uuv;
  } else {
    *r1 = p0 - (p1 + 1.0) * u + uuv;
    *r2 = p1 * u - uuv;
  }
}

/* spec.h:166-169 */
double orc_imex(double dt, double dtD, double U, double R, double lam) {
  return (U + dt * R) / (1.0 + dtD * lam);
}

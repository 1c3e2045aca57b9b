/*
 * Thin C-ABI wrapper around the reference's own native spec header,
 * bench/shtns/spec.h, compiled where it lies (the Makefile in this folder
 * adds its directory to the include path; nothing of it is copied here).
 *
 * TEST INFRASTRUCTURE ONLY: the resulting oracle/_ref/libref_spec.so pins the
 * oracle's restatement of the seeded RNG (spec.h:236-241, the mulberry32 of
 * src/mgpu/noise.ts), the seeded spectrum (spec.h:277-288), the grid rule
 * (spec.h:201-208) and the per-point reaction / IMEX arithmetic
 * (spec.h:146-169, both in f64 and in f32, the latter being what a generated
 * fp32 kernel must reproduce bit for bit).
 */
#include "spec.h"

extern "C" {

double ref_rand(uint32_t *state) {
  struct shtb_rng r;
  shtb_rng_init(&r, *state);
  double v = shtb_rand(&r);
  *state = r.s;
  return v;
}

void ref_seeded_spectrum(int lmax, int mmax, uint32_t seed, float *qlm) {
  shtb_seeded_spectrum(lmax, mmax, seed, qlm);
}

void ref_grid_for_lmax(int lmax, int pdeg, int *nlat, int *nphi) {
  shtb_grid_for_lmax(lmax, pdeg, nlat, nphi);
}

long ref_nlm_calc(int lmax, int mmax) { return shtb_nlm_calc(lmax, mmax); }

void ref_react_f64(int model, double p0, double p1, double u, double v, double *r1,
                   double *r2) {
  shtb_step_const<double> c;
  c.model = model;
  c.p0 = p0;
  c.p1 = p1;
  c.dt = 0;
  c.dtD[0] = c.dtD[1] = 0;
  double a = 0, b = 0;
  shtb_react(c, u, v, &a, &b);
  *r1 = a;
  *r2 = b;
}

void ref_react_f32(int model, float p0, float p1, const float *u, const float *v, float *r1,
                   float *r2, long n) {
  shtb_step_const<float> c;
  c.model = model;
  c.p0 = p0;
  c.p1 = p1;
  c.dt = 0;
  c.dtD[0] = c.dtD[1] = 0;
  for (long i = 0; i < n; i++) {
    float a = 0, b = 0;
    shtb_react(c, u[i], v ? v[i] : 0.0f, &a, &b);
    r1[i] = a;
    if (r2) r2[i] = b;
  }
}

/* dtD is passed as the product the caller formed, as spec.h:402-406 does. */
void ref_imex_f32(float dt, float dtD, const float *U, const float *R, const float *lam,
                  float *out, long n) {
  shtb_step_const<float> c;
  c.model = 0;
  c.p0 = c.p1 = 0;
  c.dt = dt;
  c.dtD[0] = dtD;
  c.dtD[1] = 0;
  for (long i = 0; i < n; i++) out[i] = shtb_imex(c, 0, U[i], R[i], lam[i]);
}

double ref_imex_f64(double dt, double dtD, double U, double R, double lam) {
  shtb_step_const<double> c;
  c.model = 0;
  c.p0 = c.p1 = 0;
  c.dt = dt;
  c.dtD[0] = dtD;
  c.dtD[1] = 0;
  return shtb_imex(c, 0, U, R, lam);
}

}  // extern "C"

/*
 * gradient.c -- FP64 reference gradient of the Yoshida-4 propagation, by complex-step
 * differentiation.  TEST INFRASTRUCTURE ONLY: the checker for the engine's hand-written
 * adjoint.
 *
 * The force model and the object step of force_model.inc are evaluated in double complex
 * with one input perturbed by i*h (Squire & Trapp, SIAM Rev. 40 (1998) 110; Martins,
 * Sturdza & Alonso, ACM TOMS 29 (2003) 245).  Im(f(x + i h))/h equals f'(x) up to
 * O(h^2) with no subtractive cancellation, so with h = 1e-20 the derivative is exact to
 * FP64 rounding of the same arithmetic the oracle's real propagation performs.  It shares
 * no derivative formula with the adjoint it checks.  Branch decisions (atmosphere row,
 * shadow) are taken on the real part, i.e. the derivative is that of the branch the real
 * trajectory takes.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_internal.h"

#define OR_T double complex
#define OR_SQRT csqrt
#define OR_EXP cexp
#define OR_RE(x) creal(x)
#define OR_FN(name) cs_##name
#include "force_model.inc"

void or_yoshida_vjp(size_t n, const double *rx, const double *ry, const double *rz, const double *vx,
                    const double *vy, const double *vz, const double *aom, const double *bc_inv, const double *cr,
                    const int32_t *activity, double t0, double dt, unsigned mask, int nsteps, const double *g_rx,
                    const double *g_ry, const double *g_rz, const double *g_vx, const double *g_vy,
                    const double *g_vz, double *o_rx, double *o_ry, double *o_rz, double *o_vx, double *o_vy,
                    double *o_vz, double *o_bc, int nthreads) {
  double c[4], d[3];
  or_yoshida_coefficients(c, d);
  const double tau[3] = {c[0], c[0] + c[1], c[0] + c[1] + c[2]};
  or_env *env = (or_env *)malloc(sizeof(or_env) * 3 * (size_t)(nsteps > 0 ? nsteps : 1));
  for (int s = 0; s < nsteps; ++s) {
    const double t = t0 + (double)s * dt;
    for (int k = 0; k < 3; ++k) or_env_at(t + tau[k] * dt, &env[3 * s + k]);
  }
  const double h = 1e-20;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (long long i = 0; i < (long long)n; ++i) {
    const double x0[7] = {rx[i], ry[i], rz[i], vx[i], vy[i], vz[i], bc_inv ? bc_inv[i] : 0.0};
    const double g[6] = {g_rx[i], g_ry[i], g_rz[i], g_vx[i], g_vy[i], g_vz[i]};
    double out[7];
    for (int j = 0; j < 7; ++j) {
      double complex x[7];
      for (int k = 0; k < 7; ++k) x[k] = x0[k];
      x[j] += h * I;
      const cs_params p = {aom ? aom[i] : 0.0, x[6], cr ? cr[i] : 1.0, activity ? activity[i] : 0};
      double complex r[3] = {x[0], x[1], x[2]}, v[3] = {x[3], x[4], x[5]};
      cs_yoshida_object(r, v, &p, env, c, d, dt, mask, nsteps);
      const double complex y[6] = {r[0], r[1], r[2], v[0], v[1], v[2]};
      double acc = 0.0;
      for (int k = 0; k < 6; ++k) acc += g[k] * (cimag(y[k]) / h);
      out[j] = acc;
    }
    o_rx[i] = out[0];
    o_ry[i] = out[1];
    o_rz[i] = out[2];
    o_vx[i] = out[3];
    o_vy[i] = out[4];
    o_vz[i] = out[5];
    if (o_bc) o_bc[i] = out[6];
  }
  free(env);
}

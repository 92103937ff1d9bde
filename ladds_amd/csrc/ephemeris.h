// ephemeris.h -- uniform per-sub-step environment in FP64 (host and device).
//
// The Sun/Moon positions and the Earth rotation angle are identical for every
// object of a sub-step, so they are evaluated once per sub-step and passed to the
// propagation kernel (the reference evaluates them inside OrbitPropagator's
// SOL/LUN/C22/S22 components; its accumulator starts at t0 = 0.0, Simulation.cpp:143).
//   Sun:  Montenbruck & Gill (2000) eq. 3.46-3.47 (low-precision solar coordinates)
//   Moon: Montenbruck & Gill (2000) eq. 3.48-3.50 (low-precision lunar series)
//   Earth rotation: GMST(J2000.0) + omega_E t.
// t is seconds since J2000.0; TT and UT1 are not distinguished.  The simulation loop
// evaluates the environments on the device (k_substep_env, one thread per sub-step);
// the adjoint's host driver evaluates the same functions on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "constants.h"
#include "engine.h"

namespace ladds {

namespace eph {
constexpr double kDeg = kPi / 180.0;
constexpr double kArcsec = kDeg / 3600.0;
}  // namespace eph

__host__ __device__ inline void sun_position(double t, double out[3]) {
  using eph::kArcsec;
  using eph::kDeg;
  const double T = t / kSecPerCentury;
  const double M = (357.5256 + 35999.049 * T) * kDeg;
  const double L = 282.9400 * kDeg + M + (6892.0 * sin(M) + 72.0 * sin(2.0 * M)) * kArcsec;
  const double r = (149.619 - 2.499 * cos(M) - 0.021 * cos(2.0 * M)) * 1.0e6;
  const double sl = sin(L);
  out[0] = r * cos(L);
  out[1] = r * sl * cos(kObliquity);
  out[2] = r * sl * sin(kObliquity);
}

__host__ __device__ inline void moon_position(double t, double out[3]) {
  using eph::kArcsec;
  using eph::kDeg;
  const double T = t / kSecPerCentury;
  const double L0 = (218.31617 + 481267.88088 * T - 1.3972 * T) * kDeg;
  const double l = (134.96292 + 477198.86753 * T) * kDeg;
  const double lp = (357.52543 + 35999.04944 * T) * kDeg;
  const double F = (93.27283 + 483202.01873 * T) * kDeg;
  const double D = (297.85027 + 445267.11135 * T) * kDeg;
  const double dlam = 22640.0 * sin(l) + 769.0 * sin(2.0 * l) - 4586.0 * sin(l - 2.0 * D) + 2370.0 * sin(2.0 * D) -
                      668.0 * sin(lp) - 412.0 * sin(2.0 * F) - 212.0 * sin(2.0 * l - 2.0 * D) -
                      206.0 * sin(l + lp - 2.0 * D) + 192.0 * sin(l + 2.0 * D) - 165.0 * sin(lp - 2.0 * D) +
                      148.0 * sin(l - lp) - 125.0 * sin(D) - 110.0 * sin(l + lp) - 55.0 * sin(2.0 * F - 2.0 * D);
  const double lam = L0 + dlam * kArcsec;
  const double beta = (18520.0 * sin(F + lam - L0 + (412.0 * sin(2.0 * F) + 541.0 * sin(lp)) * kArcsec) -
                       526.0 * sin(F - 2.0 * D) + 44.0 * sin(l + F - 2.0 * D) - 31.0 * sin(-l + F - 2.0 * D) -
                       25.0 * sin(-2.0 * l + F) - 23.0 * sin(lp + F - 2.0 * D) + 21.0 * sin(-l + F) +
                       11.0 * sin(-lp + F - 2.0 * D)) *
                      kArcsec;
  const double r = 385000.0 - 20905.0 * cos(l) - 3699.0 * cos(2.0 * D - l) - 2956.0 * cos(2.0 * D) -
                   570.0 * cos(2.0 * l) + 246.0 * cos(2.0 * l - 2.0 * D) - 205.0 * cos(lp - 2.0 * D) -
                   171.0 * cos(l + 2.0 * D) - 152.0 * cos(l + lp - 2.0 * D);
  const double xe = r * cos(lam) * cos(beta);
  const double ye = r * sin(lam) * cos(beta);
  const double ze = r * sin(beta);
  const double ce = cos(kObliquity), se = sin(kObliquity);
  out[0] = xe;
  out[1] = ye * ce - ze * se;
  out[2] = ye * se + ze * ce;
}

__host__ __device__ inline double earth_rotation_angle(double t) {
  double th = fmod(kThetaJ2000 + kOmegaEarth * t, 2.0 * kPi);
  if (th < 0.0) th += 2.0 * kPi;
  return th;
}

__host__ __device__ inline void substep_env(double t, SubstepEnv<double> &e64, SubstepEnv<float> &e32) {
  const double th = earth_rotation_angle(t);
  e64.ct = cos(th);
  e64.st = sin(th);
  sun_position(t, e64.sun);
  moon_position(t, e64.moon);
  const double s2 = e64.sun[0] * e64.sun[0] + e64.sun[1] * e64.sun[1] + e64.sun[2] * e64.sun[2];
  const double m2 = e64.moon[0] * e64.moon[0] + e64.moon[1] * e64.moon[1] + e64.moon[2] * e64.moon[2];
  e64.sun_inv_s2 = 1.0 / s2;
  e64.sun_k = kGmSun / (s2 * sqrt(s2));
  e64.moon_inv_s2 = 1.0 / m2;
  e64.moon_k = kGmMoon / (m2 * sqrt(m2));
  for (int k = 0; k < 3; ++k) {
    e64.sun_acc[k] = e64.sun_k * e64.sun[k];
    e64.moon_acc[k] = e64.moon_k * e64.moon[k];
  }
  e64.pad[0] = e64.pad[1] = 0.0;
  e32.ct = (float)e64.ct;
  e32.st = (float)e64.st;
  for (int k = 0; k < 3; ++k) {
    e32.sun[k] = (float)e64.sun[k];
    e32.moon[k] = (float)e64.moon[k];
  }
  e32.sun_inv_s2 = (float)e64.sun_inv_s2;
  e32.sun_k = (float)e64.sun_k;
  e32.moon_inv_s2 = (float)e64.moon_inv_s2;
  e32.moon_k = (float)e64.moon_k;
  for (int k = 0; k < 3; ++k) {
    e32.sun_acc[k] = (float)e64.sun_acc[k];
    e32.moon_acc[k] = (float)e64.moon_acc[k];
  }
  e32.pad[0] = e32.pad[1] = 0.f;
}

void yoshida_coefficients(double c[4], double d[3]);

// Environments of steps [first, first + count): step s starts at t0 + (base + s) dt and
// its three kicks sit at the drift times t + tau[k] dt; entry 3 s + k of env64 / env32.
hipError_t launch_substep_env(double t0, double dt, uint64_t base, const double tau[3], int first, int count,
                              SubstepEnv<double> *env64, SubstepEnv<float> *env32, hipStream_t stream);

}  // namespace ladds

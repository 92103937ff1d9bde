// ephemeris.cpp -- uniform per-sub-step environment computed on the host in FP64.
//
// The Sun/Moon positions and the Earth rotation angle are identical for every
// object of a sub-step, so they are evaluated once per sub-step here and passed
// to the propagation kernel (the reference evaluates them inside OrbitPropagator's
// SOL/LUN/C22/S22 components; its accumulator starts at t0 = 0.0, Simulation.cpp:143).
//   Sun:  Montenbruck & Gill (2000) eq. 3.46-3.47 (low-precision solar coordinates)
//   Moon: Montenbruck & Gill (2000) eq. 3.48-3.50 (low-precision lunar series)
//   Earth rotation: GMST(J2000.0) + omega_E t.
// t is seconds since J2000.0; TT and UT1 are not distinguished.
#include <cmath>

#include "constants.h"
#include "ephemeris.h"

namespace ladds {

namespace {
constexpr double kDeg = kPi / 180.0;
constexpr double kArcsec = kDeg / 3600.0;
}  // namespace

void sun_position(double t, double out[3]) {
  const double T = t / kSecPerCentury;
  const double M = (357.5256 + 35999.049 * T) * kDeg;
  const double L = 282.9400 * kDeg + M + (6892.0 * std::sin(M) + 72.0 * std::sin(2.0 * M)) * kArcsec;
  const double r = (149.619 - 2.499 * std::cos(M) - 0.021 * std::cos(2.0 * M)) * 1.0e6;
  const double sl = std::sin(L);
  out[0] = r * std::cos(L);
  out[1] = r * sl * std::cos(kObliquity);
  out[2] = r * sl * std::sin(kObliquity);
}

void moon_position(double t, double out[3]) {
  const double T = t / kSecPerCentury;
  const double L0 = (218.31617 + 481267.88088 * T - 1.3972 * T) * kDeg;
  const double l = (134.96292 + 477198.86753 * T) * kDeg;
  const double lp = (357.52543 + 35999.04944 * T) * kDeg;
  const double F = (93.27283 + 483202.01873 * T) * kDeg;
  const double D = (297.85027 + 445267.11135 * T) * kDeg;
  using std::cos;
  using std::sin;
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

double earth_rotation_angle(double t) {
  double th = std::fmod(kThetaJ2000 + kOmegaEarth * t, 2.0 * kPi);
  if (th < 0.0) th += 2.0 * kPi;
  return th;
}

void substep_env(double t, SubstepEnv<double> &e64, SubstepEnv<float> &e32) {
  const double th = earth_rotation_angle(t);
  e64.ct = std::cos(th);
  e64.st = std::sin(th);
  sun_position(t, e64.sun);
  moon_position(t, e64.moon);
  const double s2 = e64.sun[0] * e64.sun[0] + e64.sun[1] * e64.sun[1] + e64.sun[2] * e64.sun[2];
  const double m2 = e64.moon[0] * e64.moon[0] + e64.moon[1] * e64.moon[1] + e64.moon[2] * e64.moon[2];
  e64.sun_inv_s2 = 1.0 / s2;
  e64.sun_k = kGmSun / (s2 * std::sqrt(s2));
  e64.moon_inv_s2 = 1.0 / m2;
  e64.moon_k = kGmMoon / (m2 * std::sqrt(m2));
  for (int k = 0; k < 3; ++k) {
    e64.sun_acc[k] = e64.sun_k * e64.sun[k];
    e64.moon_acc[k] = e64.moon_k * e64.moon[k];
  }
  for (double &x : e64.pad) x = 0.0;
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
  for (float &x : e32.pad) x = 0.f;
}

void yoshida_coefficients(double c[4], double d[3]) {
  const double cb = std::cbrt(2.0);
  const double w1 = 1.0 / (2.0 - cb);
  const double w0 = -cb / (2.0 - cb);
  c[0] = c[3] = w1 / 2.0;
  c[1] = c[2] = (w0 + w1) / 2.0;
  d[0] = d[2] = w1;
  d[1] = w0;
}

}  // namespace ladds

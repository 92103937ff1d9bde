// physics.h -- device force model of the Yoshida propagator.
//
// Replaces OrbitPropagator's AccelerationAccumulator and its eight components
// (upstream FG-TUM/OrbitPropagator @ b90f3bae, selected by the flag array built at
// src/ladds/Simulation.cpp:127-135).  The upstream source is not part of the
// LADDS tree, so each component is restated from its published form:
//   KEP, J2, C22, S22  Montenbruck & Gill, Satellite Orbits (2000) Sec. 3.2
//   SOL, LUN           third-body term, M&G eq. 3.37
//   SRP                cannonball model, M&G eq. 3.75, cylindrical shadow (Sec. 3.4.2)
//   DRAG               M&G eq. 3.97 with a co-rotating atmosphere and the exponential
//                      density table of Vallado (2013) Table 8-4
// evasivePreserving objects receive the Keplerian term only (Particle.h:49-52).
//
// Precision forms:
//   kFormFp64      every term in FP64.  Third-body and SRP use the direct form
//                  GM (d/|d|^3 - s/|s|^3): its cancellation costs ~1e-12 relative in FP64.
//   kFormSplit     Kepler, J2, C22/S22 and drag in FP64; Sun, Moon and SRP in FP32,
//                  accumulated into the FP64 kick.  J2's acceleration needs ~29
//                  significant bits to hold 1 cm/day (FP32 has 24), and drag moves a
//                  low object hundreds of km along-track per day, so FP32 rounding of
//                  either costs centimetres (measured: up to 1 m for J2, 25 cm for drag).
//                  The tesseral terms in FP32 (Earth-fixed rotation and field) leave
//                  0.5 cm at the 99.9th percentile and over 1 cm in the tail of a
//                  1M-object day, so they stay FP64 too; the FP32 terms then contribute
//                  ~0.01 cm.  The FP32 third-body terms use Battin's cancellation-free
//                  form a = -GM/|s|^3 (1+q)^-3/2 (r + f(q) s) (Battin 1987, eq. 8.61).
//                  The SRP shadow decision is taken on the FP64 state.
//   kFormSplitAll32  J2, C22/S22, Sun, Moon and SRP in FP32 (Kepler and drag FP64).
#pragma once

#include <hip/hip_runtime.h>

#include "constants.h"
#include "engine.h"

namespace ladds {

enum : unsigned {
  kKepBit = 1u << 0,
  kJ2Bit = 1u << 1,
  kC22Bit = 1u << 2,
  kS22Bit = 1u << 3,
  kSolBit = 1u << 4,
  kLunBit = 1u << 5,
  kSrpBit = 1u << 6,
  kDragBit = 1u << 7
};

// Exponential atmosphere staged in LDS by the propagation kernel: per-lane table reads
// then cost one LDS access instead of a chain of dependent global loads.  h0 is padded
// to 32 rows with +inf so the row search is a fixed 5-step binary search.
constexpr int kAtmPadded = 32;
struct AtmTable {
  double h0[kAtmPadded];
  double rho[kAtmRows];
  double inv_scale[kAtmRows];
};

// Host-side fill of the table from the constants (uploaded to __constant__ memory by the
// kernels' translation units).
inline void fill_atm_table(AtmTable &t) {
  for (int i = 0; i < kAtmPadded; ++i) t.h0[i] = i < kAtmRows ? kAtmH0[i] : __builtin_inf();
  for (int i = 0; i < kAtmRows; ++i) {
    t.rho[i] = kAtmRho[i];
    t.inv_scale[i] = 1.0 / kAtmScale[i];
  }
}

// Row of the table: largest i with h0[i] <= h (0 below the table and for NaN).
__device__ __forceinline__ int atm_row(double h, const AtmTable &t) {
  int i = 0;
#pragma unroll
  for (int step = kAtmPadded / 2; step > 0; step >>= 1)
    if (t.h0[i + step] <= h) i += step;
  return i;
}

// rho0[i] * exp(-(h - h0[i]) / H[i]) in FP64.
__device__ __forceinline__ double atm_density(double h, const AtmTable &t) {
  const int i = atm_row(h, t);
  return t.rho[i] * exp(-(h - t.h0[i]) * t.inv_scale[i]);
}

// Tesseral degree-2 order-2 terms in the Earth-fixed frame, rotated back to ECI.
template <typename T>
__device__ __forceinline__ void tesseral(T x, T y, T z, T inv_r2, T inv_r5, unsigned m, T ct, T st, T &ax, T &ay,
                                         T &az) {
  const T xe = ct * x + st * y;
  const T ye = -st * x + ct * y;
  const T inv_r7 = inv_r5 * inv_r2;
  T ex = 0, ey = 0, ez = 0;
  if (m & kC22Bit) {
    const T k = (T)kC22Coef;
    const T q = (xe * xe - ye * ye) * inv_r7 * (T)5;
    ex += k * (T(2) * xe * inv_r5 - xe * q);
    ey += k * (T(-2) * ye * inv_r5 - ye * q);
    ez += k * (-z * q);
  }
  if (m & kS22Bit) {
    const T k = (T)kS22Coef;
    const T q = (xe * ye) * inv_r7 * (T)5;
    ex += k * (ye * inv_r5 - xe * q);
    ey += k * (xe * inv_r5 - ye * q);
    ez += k * (-z * q);
  }
  ax += ct * ex - st * ey;
  ay += st * ex + ct * ey;
  az += ez;
}

// Cylindrical shadow test (M&G Sec. 3.4.2), always in FP64: it is a discontinuous
// decision, so it is taken on the FP64 state in every precision form.
__device__ __forceinline__ bool in_shadow(double x, double y, double z, double r2, const SubstepEnv<double> &e) {
  const double rs = x * e.sun[0] + y * e.sun[1] + z * e.sun[2];
  return (rs < 0.0) && (r2 - rs * rs * e.sun_inv_s2 < kREarth * kREarth);
}

// FP64 third body, direct form: GM d/|d|^3 - GM s/|s|^3 with d = s - r.
__device__ __forceinline__ void third_body64(double x, double y, double z, const double *s, double gm,
                                             const double *acc, double &ax, double &ay, double &az) {
  const double dx = s[0] - x, dy = s[1] - y, dz = s[2] - z;
  const double inv_d = rsqrt(dx * dx + dy * dy + dz * dz);
  const double f = gm * inv_d * inv_d * inv_d;
  ax += f * dx - acc[0];
  ay += f * dy - acc[1];
  az += f * dz - acc[2];
}

// FP64 SRP: srp (r - s)/|r - s|^3, srp = kSrpCoef * cr * aom.
__device__ __forceinline__ void srp64(double x, double y, double z, const double *s, double srp, double &ax,
                                      double &ay, double &az) {
  const double dx = x - s[0], dy = y - s[1], dz = z - s[2];
  const double inv_d = rsqrt(dx * dx + dy * dy + dz * dz);
  const double f = srp * inv_d * inv_d * inv_d;
  ax += f * dx;
  ay += f * dy;
  az += f * dz;
}

// FP32 third body (Battin's form) and, sharing its geometry, SRP for the Sun.
// k = GM/|s|^3; inv_p32 = (1+q)^{-3/2} = |s|^3 / |r - s|^3.
__device__ __forceinline__ void third_body32(float x, float y, float z, float r2, const float *s, float inv_s2,
                                             float k, bool gravity, float srp, float &ax, float &ay, float &az) {
  const float rs = x * s[0] + y * s[1] + z * s[2];
  const float q = (r2 - 2.f * rs) * inv_s2;
  const float onepq = 1.f + q;
  const float rq = __builtin_amdgcn_rsqf(onepq);
  const float inv_p32 = rq * rq * rq;
  if (gravity) {
    const float p32 = onepq * onepq * rq;
    const float f = q * (3.f + q * (3.f + q)) * __builtin_amdgcn_rcpf(1.f + p32);
    const float kk = k * inv_p32;
    ax -= kk * (x + f * s[0]);
    ay -= kk * (y + f * s[1]);
    az -= kk * (z + f * s[2]);
  }
  if (srp != 0.f) {
    // 1/|r - s|^3 = (k / GM_sun) (1+q)^{-3/2}
    const float g = srp * (k * (float)(1.0 / kGmSun)) * inv_p32;
    ax += g * (x - s[0]);
    ay += g * (y - s[1]);
    az += g * (z - s[2]);
  }
}

// Drag with a co-rotating atmosphere; drag64 = kDragCoef * bc_inv.
__device__ __forceinline__ void drag_term64(double x, double y, double vx, double vy, double vz, double h,
                                            double drag64, const AtmTable &t, double &ax, double &ay, double &az) {
  const double om = kOmegaEarth;
  const double wx = vx + om * y, wy = vy - om * x, wz = vz;
  const double vn = sqrt(wx * wx + wy * wy + wz * wz);
  const double f = drag64 * atm_density(h, t) * vn;
  ax += f * wx;
  ay += f * wy;
  az += f * wz;
}

struct ObjParams {
  double srp64, drag64;
  float srp32;
  int activity;
};

// Acceleration [km/s^2] of one object in the given precision form.
template <int FORM>
__device__ __forceinline__ void acceleration(double x, double y, double z, double vx, double vy, double vz,
                                             const ObjParams &p, unsigned mask, const SubstepEnv<double> &e64,
                                             const SubstepEnv<float> &e32, const AtmTable &atm, double &ax,
                                             double &ay, double &az) {
  const unsigned m = (p.activity == 2) ? (mask & kKepBit) : mask;
  const double r2 = x * x + y * y + z * z;
  const double inv_r = rsqrt(r2);
  const double inv_r2 = inv_r * inv_r;
  const double inv_r3 = inv_r2 * inv_r;
  ax = ay = az = 0.0;
  if (m & kKepBit) {
    const double f = -kMuEarth * inv_r3;
    ax = f * x;
    ay = f * y;
    az = f * z;
  }
  constexpr bool j2_64 = (FORM != kFormSplitAll32);
  if (j2_64 && (m & kJ2Bit)) {
    const double z2 = z * z * inv_r2 * 5.0;
    const double f = -kJ2Coef * inv_r3 * inv_r2;
    ax += f * x * (1.0 - z2);
    ay += f * y * (1.0 - z2);
    az += f * z * (3.0 - z2);
  }
  // drag in FP64 in every form
  if (m & kDragBit) drag_term64(x, y, vx, vy, vz, r2 * inv_r - kREarth, p.drag64, atm, ax, ay, az);
  if constexpr (FORM == kFormFp64) {
    if (m & (kC22Bit | kS22Bit)) tesseral<double>(x, y, z, inv_r2, inv_r3 * inv_r2, m, e64.ct, e64.st, ax, ay, az);
    if (m & kSolBit) third_body64(x, y, z, e64.sun, kGmSun, e64.sun_acc, ax, ay, az);
    if (m & kLunBit) third_body64(x, y, z, e64.moon, kGmMoon, e64.moon_acc, ax, ay, az);
    if ((m & kSrpBit) && !in_shadow(x, y, z, r2, e64)) srp64(x, y, z, e64.sun, p.srp64, ax, ay, az);
  } else {
    constexpr bool tess_64 = (FORM == kFormSplit);
    if (tess_64 && (m & (kC22Bit | kS22Bit)))
      tesseral<double>(x, y, z, inv_r2, inv_r3 * inv_r2, m, e64.ct, e64.st, ax, ay, az);
    constexpr unsigned kFp64Bits = kKepBit | kDragBit | (j2_64 ? kJ2Bit : 0u) | (tess_64 ? (kC22Bit | kS22Bit) : 0u);
    if (m & ~kFp64Bits) {
      const float xf = (float)x, yf = (float)y, zf = (float)z;
      const float r2f = (float)r2;
      const float inv_r2f = (float)inv_r2;
      const float inv_r3f = (float)inv_r3;
      float fx = 0.f, fy = 0.f, fz = 0.f;
      if (!j2_64 && (m & kJ2Bit)) {
        const float z2 = zf * zf * inv_r2f * 5.f;
        const float f = -(float)kJ2Coef * inv_r3f * inv_r2f;
        fx += f * xf * (1.f - z2);
        fy += f * yf * (1.f - z2);
        fz += f * zf * (3.f - z2);
      }
      if (!tess_64 && (m & (kC22Bit | kS22Bit)))
        tesseral<float>(xf, yf, zf, inv_r2f, inv_r3f * inv_r2f, m, e32.ct, e32.st, fx, fy, fz);
      const bool sun_grav = (m & kSolBit) != 0;
      const float srp = ((m & kSrpBit) && !in_shadow(x, y, z, r2, e64)) ? p.srp32 : 0.f;
      if (sun_grav || srp != 0.f)
        third_body32(xf, yf, zf, r2f, e32.sun, e32.sun_inv_s2, e32.sun_k, sun_grav, srp, fx, fy, fz);
      if (m & kLunBit) third_body32(xf, yf, zf, r2f, e32.moon, e32.moon_inv_s2, e32.moon_k, true, 0.f, fx, fy, fz);
      ax += (double)fx;
      ay += (double)fy;
      az += (double)fz;
    }
  }
}

}  // namespace ladds

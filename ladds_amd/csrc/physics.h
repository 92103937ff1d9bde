// physics.h -- device force model of the Yoshida propagator.
//
// Replaces OrbitPropagator's AccelerationAccumulator and its eight components
// (upstream FG-TUM/OrbitPropagator @ b90f3bae, selected by the flag array built at
// src/ladds/Simulation.cpp:127-135).  The upstream source is not part of the
// LADDS tree, so each component is restated from its published form:
//   KEP, J2, C22, S22  Montenbruck & Gill, Satellite Orbits (2000) Sec. 3.2
//   SOL, LUN           third-body term, M&G eq. 3.37, evaluated in Battin's
//                      cancellation-free form (Battin 1987, eq. 8.61)
//   SRP                cannonball model, M&G eq. 3.75, cylindrical shadow (Sec. 3.4.2)
//   DRAG               M&G eq. 3.97 with a co-rotating atmosphere and the exponential
//                      density table of Vallado (2013) Table 8-4
// evasivePreserving objects receive the Keplerian term only (Particle.h:49-52).
//
// Precision forms:
//   kFormFp64      every term in FP64.
//   kFormSplit     Kepler and J2 in FP64 (J2's acceleration needs ~29 significant
//                  bits to hold 1 cm/day, FP32 has 24); C22, S22, Sun, Moon, SRP and
//                  drag in FP32, accumulated into the FP64 kick.
//   kFormSplitAll32  as kFormSplit with J2 in FP32.
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

__constant__ extern double c_atm_h0[kAtmRows];
__constant__ extern double c_atm_rho[kAtmRows];
__constant__ extern double c_atm_inv_scale[kAtmRows];

// Row of the exponential atmosphere: largest i with h0[i] <= h (0 below the table).
__device__ __forceinline__ int atm_row(double h) {
  int lo = 0, hi = kAtmRows - 1;
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    const int mid = (lo + hi + 1) >> 1;
    if (c_atm_h0[mid] <= h)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <typename T>
__device__ __forceinline__ T atm_density(double h) {
  const int i = atm_row(h);
  if constexpr (sizeof(T) == 8) {
    return c_atm_rho[i] * exp(-(h - c_atm_h0[i]) * c_atm_inv_scale[i]);
  } else {
    const float dh = (float)(h - c_atm_h0[i]);
    return (float)c_atm_rho[i] * expf(-dh * (float)c_atm_inv_scale[i]);
  }
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

// Third body in Battin's form; also returns (1+q) and (1+q)^{3/2} for reuse by SRP.
template <typename T>
__device__ __forceinline__ void third_body(T x, T y, T z, T r2, const T *s, T inv_s2, T k, T &ax, T &ay, T &az) {
  const T rs = x * s[0] + y * s[1] + z * s[2];
  const T q = (r2 - T(2) * rs) * inv_s2;
  const T onepq = T(1) + q;
  const T p32 = onepq * sqrt(onepq);
  const T f = q * (T(3) + q * (T(3) + q)) / (T(1) + p32);
  const T kk = k / p32;
  ax -= kk * (x + f * s[0]);
  ay -= kk * (y + f * s[1]);
  az -= kk * (z + f * s[2]);
}

// Solar radiation pressure (cannonball) with cylindrical shadow.  srp = kSrpCoef * cr * aom.
template <typename T>
__device__ __forceinline__ void srp_term(T x, T y, T z, T r2, const T *s, T inv_s2, T sun_k, T srp, T &ax, T &ay,
                                         T &az) {
  const T rs = x * s[0] + y * s[1] + z * s[2];
  const bool shadow = (rs < T(0)) && (r2 - rs * rs * inv_s2 < (T)(kREarth * kREarth));
  const T q = (r2 - T(2) * rs) * inv_s2;
  const T onepq = T(1) + q;
  // 1/|r-s|^3 = (1/|s|^3) (1+q)^{-3/2};  sun_k = GM_sun / |s|^3
  const T inv_d3 = sun_k * (T)(1.0 / kGmSun) / (onepq * sqrt(onepq));
  const T f = shadow ? T(0) : srp * inv_d3;
  ax += f * (x - s[0]);
  ay += f * (y - s[1]);
  az += f * (z - s[2]);
}

template <typename T>
__device__ __forceinline__ void drag_term(T x, T y, T vx, T vy, T vz, double h, T drag, T &ax, T &ay, T &az) {
  const T om = (T)kOmegaEarth;
  const T wx = vx + om * y;
  const T wy = vy - om * x;
  const T wz = vz;
  const T vn = sqrt(wx * wx + wy * wy + wz * wz);
  const T f = drag * atm_density<T>(h) * vn;  // drag = kDragCoef * bc_inv
  ax += f * wx;
  ay += f * wy;
  az += f * wz;
}

struct ObjParams {
  double srp64, drag64;
  float srp32, drag32;
  int activity;
};

// Acceleration [km/s^2] of one object in the given precision form.
template <int FORM>
__device__ __forceinline__ void acceleration(double x, double y, double z, double vx, double vy, double vz,
                                             const ObjParams &p, unsigned mask, const SubstepEnv<double> &e64,
                                             const SubstepEnv<float> &e32, double &ax, double &ay, double &az) {
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
  if constexpr (FORM == kFormFp64) {
    if (m & (kC22Bit | kS22Bit)) tesseral<double>(x, y, z, inv_r2, inv_r3 * inv_r2, m, e64.ct, e64.st, ax, ay, az);
    if (m & kSolBit) third_body<double>(x, y, z, r2, e64.sun, e64.sun_inv_s2, e64.sun_k, ax, ay, az);
    if (m & kLunBit) third_body<double>(x, y, z, r2, e64.moon, e64.moon_inv_s2, e64.moon_k, ax, ay, az);
    if (m & kSrpBit) srp_term<double>(x, y, z, r2, e64.sun, e64.sun_inv_s2, e64.sun_k, p.srp64, ax, ay, az);
    if (m & kDragBit) drag_term<double>(x, y, vx, vy, vz, r2 * inv_r - kREarth, p.drag64, ax, ay, az);
  } else {
    if (m & ~(kKepBit | (j2_64 ? kJ2Bit : 0u))) {
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
      if (m & (kC22Bit | kS22Bit)) tesseral<float>(xf, yf, zf, inv_r2f, inv_r3f * inv_r2f, m, e32.ct, e32.st, fx, fy, fz);
      if (m & kSolBit) third_body<float>(xf, yf, zf, r2f, e32.sun, e32.sun_inv_s2, e32.sun_k, fx, fy, fz);
      if (m & kLunBit) third_body<float>(xf, yf, zf, r2f, e32.moon, e32.moon_inv_s2, e32.moon_k, fx, fy, fz);
      if (m & kSrpBit) srp_term<float>(xf, yf, zf, r2f, e32.sun, e32.sun_inv_s2, e32.sun_k, p.srp32, fx, fy, fz);
      if (m & kDragBit)
        drag_term<float>(xf, yf, (float)vx, (float)vy, (float)vz, r2 * inv_r - kREarth, p.drag32, fx, fy, fz);
      ax += (double)fx;
      ay += (double)fy;
      az += (double)fz;
    }
  }
}

}  // namespace ladds

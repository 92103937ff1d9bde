// adjoint.hip -- hand-written adjoint of the FP64 Yoshida-4 propagation.
//
// Forward: the FP64 step of propagate.hip (the same acceleration<kFormFp64>), one thread
// per object, every step of the trajectory in one launch, the state stored every K steps
// (checkpoints).  Backward: one launch; each thread walks its checkpoint segments from the
// last to the first, recomputes the K step-start states of a segment from its checkpoint
// into a workspace, then runs the reverse sweep through each step:
//
//   step (v_k: velocity before kick k, r_k: position after drift k, h = dt)
//     r_0 = r + c_0 h v_0;    v_1 = v_0 + d_0 h a(r_0, v_0, t_0)
//     r_1 = r_0 + c_1 h v_1;  v_2 = v_1 + d_1 h a(r_1, v_1, t_1)
//     r_2 = r_1 + c_2 h v_2;  v_3 = v_2 + d_2 h a(r_2, v_2, t_2)
//     r'  = r_2 + c_3 h v_3;  v' = v_3
//   reverse, with w = d_k h lambda_v:
//     lambda_v += c_3 h lambda_r
//     for k = 2, 1, 0:  lambda_r += J_r(r_k, v_k)^T w;  lambda_v += J_v(r_k, v_k)^T w;
//                       lambda_bc += (da/dbc_inv) . w;  lambda_v += c_k h lambda_r
//
// The Jacobian-transpose products are derived by hand per component.  Every term except
// drag is the gradient of a potential (Kepler, J2, C22/S22 rotated with the Earth, the
// third bodies, cannonball SRP on the lit branch), so its Jacobian is a symmetric Hessian
// and J^T w is the directional derivative along w.  Drag a = D rho(h) |W| W with
// W = v - omega x r, D = -1/2 1e3 bc_inv:
//   J_v^T w = D rho (|W| w + W (W.w)/|W|)
//   J_r^T w = D |W| (W.w) rho'(h) r/|r| + B^T J_v^T w,  B^T u = (-omega u_y, omega u_x, 0)
//   da/dbc_inv . w = -1/2 1e3 rho |W| (W.w)
// Branch decisions (shadow, atmosphere row) are those of the forward trajectory: the
// result is the derivative of the branch taken, as for any piecewise-smooth model.
//
// No LADDS counterpart: the reference has no differentiable propagation.  The forward
// matches YoshidaIntegrator::integrate (OrbitPropagator @ b90f3bae, Simulation.cpp:297).
#include <cstring>
#include <string>
#include <vector>

#include "ephemeris.h"
#include "ladds_amd.h"
#include "physics.h"

namespace ladds {

__constant__ AtmTable c_atm_adj;

struct AdjointArgs {
  size_t n;
  int nsteps, every, nck;
  unsigned mask;
  double cdt[4], ddt[3];
  const SubstepEnv<double> *env;
  const double *aom, *bc, *cr;
  const int32_t *activity;
  const double *r0, *v0;  // [3][n]
  double *rT, *vT;        // [3][n]
  double *ck;             // [nck][6][n]
  const double *g_rT, *g_vT;
  double *g_r0, *g_v0, *g_bc;
  double *work;  // [every][6][n]
};

__device__ __forceinline__ void stage_atm(AtmTable &atm) {
  const double *src = reinterpret_cast<const double *>(&c_atm_adj);
  double *dst = reinterpret_cast<double *>(&atm);
  for (int k = threadIdx.x; k < (int)(sizeof(AtmTable) / sizeof(double)); k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

__device__ __forceinline__ ObjParams object_params(const AdjointArgs &a, size_t i) {
  ObjParams p;
  const double aom = a.aom ? a.aom[i] : 0.0;
  const double cr = a.cr ? a.cr[i] : 1.0;
  const double bc = a.bc ? a.bc[i] : 0.0;
  p.srp64 = kSrpCoef * cr * aom;
  p.drag64 = kDragCoef * bc;
  p.srp32 = 0.f;
  p.activity = a.activity ? a.activity[i] : 0;
  return p;
}

struct State {
  double x, y, z, vx, vy, vz;
};

__device__ __forceinline__ void yoshida_step(State &s, const ObjParams &p, unsigned mask, const double *cdt,
                                             const double *ddt, const SubstepEnv<double> *env,
                                             const AtmTable &atm) {
  const SubstepEnv<float> unused{};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.x += cdt[k] * s.vx;
    s.y += cdt[k] * s.vy;
    s.z += cdt[k] * s.vz;
    double ax, ay, az;
    acceleration<kFormFp64>(s.x, s.y, s.z, s.vx, s.vy, s.vz, p, mask, env[k], unused, atm, ax, ay, az);
    s.vx += ddt[k] * ax;
    s.vy += ddt[k] * ay;
    s.vz += ddt[k] * az;
  }
  s.x += cdt[3] * s.vx;
  s.y += cdt[3] * s.vy;
  s.z += cdt[3] * s.vz;
}

// Hessian-vector product of a central term f d, f = k/|d|^3, with d = sign (r - s):
// d(d/|d|^3)/dd = (I - 3 d d^T/|d|^2)/|d|^3 and dd/dr = sign I, so
// d/dr [f d] . w = sign f (w - 3 tau d),  tau = (d.w) / |d|^2.
__device__ __forceinline__ void central_hvp(double dx, double dy, double dz, double f, double inv_d2, double sign,
                                            double wx, double wy, double wz, double &gx, double &gy, double &gz) {
  const double tau = (dx * wx + dy * wy + dz * wz) * inv_d2;
  const double sf = sign * f;
  gx += sf * (wx - 3.0 * tau * dx);
  gy += sf * (wy - 3.0 * tau * dy);
  gz += sf * (wz - 3.0 * tau * dz);
}

// J_r^T w, J_v^T w and (da/dbc_inv).w of acceleration<kFormFp64> at (r, v).
__device__ __forceinline__ void acceleration_vjp(double x, double y, double z, double vx, double vy, double vz,
                                                 const ObjParams &p, unsigned mask, const SubstepEnv<double> &e,
                                                 const AtmTable &atm, double wx, double wy, double wz, double &gx,
                                                 double &gy, double &gz, double &gvx, double &gvy, double &gvz,
                                                 double &gbc) {
  const unsigned m = (p.activity == 2) ? (mask & kKepBit) : mask;
  const double r2 = x * x + y * y + z * z;
  const double inv_r = rsqrt(r2);
  const double inv_r2 = inv_r * inv_r;
  const double inv_r3 = inv_r2 * inv_r;
  const double sigma = (x * wx + y * wy + z * wz) * inv_r2;  // (r.w)/r^2
  gx = gy = gz = gvx = gvy = gvz = gbc = 0.0;
  if (m & kKepBit) {
    // a = -mu r / r^3
    const double f = -kMuEarth * inv_r3;
    gx += f * (wx - 3.0 * sigma * x);
    gy += f * (wy - 3.0 * sigma * y);
    gz += f * (wz - 3.0 * sigma * z);
  }
  if (m & kJ2Bit) {
    // a = f (x (1 - z2), y (1 - z2), z (3 - z2)),  f = -kJ2Coef / r^5,  z2 = 5 z^2 / r^2
    const double z2 = z * z * inv_r2 * 5.0;
    const double f = -kJ2Coef * inv_r3 * inv_r2;
    const double df = -5.0 * f * sigma;
    const double dz2 = 10.0 * inv_r2 * z * (wz - z * sigma);
    gx += df * x * (1.0 - z2) + f * (wx * (1.0 - z2) - x * dz2);
    gy += df * y * (1.0 - z2) + f * (wy * (1.0 - z2) - y * dz2);
    gz += df * z * (3.0 - z2) + f * (wz * (3.0 - z2) - z * dz2);
  }
  if (m & (kC22Bit | kS22Bit)) {
    // Earth-fixed coordinates of r and w; tesseral() differentiated along w
    const double ct = e.ct, st = e.st;
    const double xe = ct * x + st * y, ye = -st * x + ct * y;
    const double wxe = ct * wx + st * wy, wye = -st * wx + ct * wy;
    const double inv_r5 = inv_r3 * inv_r2;
    const double inv_r7 = inv_r5 * inv_r2;
    double dex = 0.0, dey = 0.0, dez = 0.0;
    if (m & kC22Bit) {
      const double k = kC22Coef;
      const double q = (xe * xe - ye * ye) * inv_r7 * 5.0;
      const double dq = 10.0 * (xe * wxe - ye * wye) * inv_r7 - 7.0 * q * sigma;
      dex += k * (2.0 * wxe * inv_r5 - 10.0 * xe * inv_r5 * sigma - wxe * q - xe * dq);
      dey += k * (-2.0 * wye * inv_r5 + 10.0 * ye * inv_r5 * sigma - wye * q - ye * dq);
      dez += k * (-wz * q - z * dq);
    }
    if (m & kS22Bit) {
      const double k = kS22Coef;
      const double q = (xe * ye) * inv_r7 * 5.0;
      const double dq = 5.0 * (wxe * ye + xe * wye) * inv_r7 - 7.0 * q * sigma;
      dex += k * (wye * inv_r5 - 5.0 * ye * inv_r5 * sigma - wxe * q - xe * dq);
      dey += k * (wxe * inv_r5 - 5.0 * xe * inv_r5 * sigma - wye * q - ye * dq);
      dez += k * (-wz * q - z * dq);
    }
    gx += ct * dex - st * dey;
    gy += st * dex + ct * dey;
    gz += dez;
  }
  if (m & kSolBit) {
    // a += GM d/|d|^3 with d = s - r (the indirect term does not depend on r)
    const double dx = e.sun[0] - x, dy = e.sun[1] - y, dz = e.sun[2] - z;
    const double inv_d2 = 1.0 / (dx * dx + dy * dy + dz * dz);
    const double inv_d = sqrt(inv_d2);
    central_hvp(dx, dy, dz, kGmSun * inv_d2 * inv_d, inv_d2, -1.0, wx, wy, wz, gx, gy, gz);
  }
  if (m & kLunBit) {
    const double dx = e.moon[0] - x, dy = e.moon[1] - y, dz = e.moon[2] - z;
    const double inv_d2 = 1.0 / (dx * dx + dy * dy + dz * dz);
    const double inv_d = sqrt(inv_d2);
    central_hvp(dx, dy, dz, kGmMoon * inv_d2 * inv_d, inv_d2, -1.0, wx, wy, wz, gx, gy, gz);
  }
  if ((m & kSrpBit) && !in_shadow(x, y, z, r2, e)) {
    // a += srp d/|d|^3 with d = r - s
    const double dx = x - e.sun[0], dy = y - e.sun[1], dz = z - e.sun[2];
    const double inv_d2 = 1.0 / (dx * dx + dy * dy + dz * dz);
    const double inv_d = sqrt(inv_d2);
    central_hvp(dx, dy, dz, p.srp64 * inv_d2 * inv_d, inv_d2, 1.0, wx, wy, wz, gx, gy, gz);
  }
  if (m & kDragBit) {
    const double om = kOmegaEarth;
    const double Wx = vx + om * y, Wy = vy - om * x, Wz = vz;
    const double vn = sqrt(Wx * Wx + Wy * Wy + Wz * Wz);
    const double h = r2 * inv_r - kREarth;
    const int row = atm_row(h, atm);
    const double rho = atm.rho[row] * exp(-(h - atm.h0[row]) * atm.inv_scale[row]);
    const double drho = -rho * atm.inv_scale[row];
    const double Ww = Wx * wx + Wy * wy + Wz * wz;
    const double D = p.drag64;
    // J_v^T w
    const double inv_vn = vn > 0.0 ? 1.0 / vn : 0.0;
    const double jvx = D * rho * (vn * wx + Wx * Ww * inv_vn);
    const double jvy = D * rho * (vn * wy + Wy * Ww * inv_vn);
    const double jvz = D * rho * (vn * wz + Wz * Ww * inv_vn);
    gvx += jvx;
    gvy += jvy;
    gvz += jvz;
    // J_r^T w: density gradient along r/|r| plus the co-rotation term B^T J_v^T w
    const double fr = D * vn * Ww * drho * inv_r;
    gx += fr * x - om * jvy;
    gy += fr * y + om * jvx;
    gz += fr * z;
    gbc += kDragCoef * rho * vn * Ww;
  }
}

__device__ __forceinline__ void load_state(State &s, const double *base, size_t n, size_t i) {
  s.x = base[i];
  s.y = base[n + i];
  s.z = base[2 * n + i];
  s.vx = base[3 * n + i];
  s.vy = base[4 * n + i];
  s.vz = base[5 * n + i];
}

__device__ __forceinline__ void store_state(const State &s, double *base, size_t n, size_t i) {
  base[i] = s.x;
  base[n + i] = s.y;
  base[2 * n + i] = s.z;
  base[3 * n + i] = s.vx;
  base[4 * n + i] = s.vy;
  base[5 * n + i] = s.vz;
}

// Four wavefronts per workgroup share 64 objects: wavefront w evaluates the component
// group kGroupMask[w] of every acceleration (and of its adjoint) for the 64 objects, the
// four partial results are summed through LDS in a fixed order, and every wavefront then
// carries the same integrator state.  One object's force evaluation is thus spread over
// the four SIMDs of a CU (the sweep over steps is sequential, so this is where the
// parallelism is).  Groups are balanced by cost.
constexpr int kAdjWaves = 4;
constexpr int kAdjLanes = 64;
__device__ constexpr unsigned kGroupMask[kAdjWaves] = {kKepBit | kJ2Bit | kSrpBit, kDragBit, kC22Bit | kS22Bit,
                                                       kSolBit | kLunBit};

// Partial sums of the four wavefronts, double-buffered so that one barrier per exchange
// suffices: a buffer is rewritten only after the next exchange's barrier, which every
// wavefront reaches after reading it.
struct GroupExchange {
  double buf[2][kAdjWaves][8][kAdjLanes];
  int parity;
};

template <int NV>
__device__ __forceinline__ void group_sum(GroupExchange &x, int &parity, double (&v)[NV]) {
  const int w = threadIdx.x / kAdjLanes, l = threadIdx.x % kAdjLanes;
#pragma unroll
  for (int k = 0; k < NV; ++k) x.buf[parity][w][k][l] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = ((x.buf[parity][0][k][l] + x.buf[parity][1][k][l]) + x.buf[parity][2][k][l]) + x.buf[parity][3][k][l];
  parity ^= 1;
}

__device__ __forceinline__ void group_acceleration(const State &s, const ObjParams &p, unsigned gmask,
                                                   const SubstepEnv<double> &env, const AtmTable &atm,
                                                   GroupExchange &x, int &parity, double &ax, double &ay,
                                                   double &az) {
  const SubstepEnv<float> unused{};
  double v[3];
  acceleration<kFormFp64>(s.x, s.y, s.z, s.vx, s.vy, s.vz, p, gmask, env, unused, atm, v[0], v[1], v[2]);
  group_sum<3>(x, parity, v);
  ax = v[0];
  ay = v[1];
  az = v[2];
}

// One Yoshida step (the composition of propagate.hip) with group-split accelerations;
// `moving` false keeps the state (dummies) while the wavefront still takes part in the
// exchanges.
__device__ __forceinline__ void group_step(State &s, const ObjParams &p, unsigned gmask, bool moving,
                                           const double *cdt, const double *ddt, const SubstepEnv<double> *env,
                                           const AtmTable &atm, GroupExchange &x, int &parity) {
  const double go = moving ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.x += go * cdt[k] * s.vx;
    s.y += go * cdt[k] * s.vy;
    s.z += go * cdt[k] * s.vz;
    double ax, ay, az;
    group_acceleration(s, p, gmask, env[k], atm, x, parity, ax, ay, az);
    s.vx += go * ddt[k] * ax;
    s.vy += go * ddt[k] * ay;
    s.vz += go * ddt[k] * az;
  }
  s.x += go * cdt[3] * s.vx;
  s.y += go * cdt[3] * s.vy;
  s.z += go * cdt[3] * s.vz;
}

__global__ __launch_bounds__(kAdjWaves * kAdjLanes) void k_adjoint_forward(AdjointArgs a) {
  __shared__ AtmTable atm;
  __shared__ GroupExchange x;
  stage_atm(atm);
  const int w = threadIdx.x / kAdjLanes, l = threadIdx.x % kAdjLanes;
  const size_t n = a.n, i = (size_t)blockIdx.x * kAdjLanes + l;
  const bool live = i < n;
  const size_t ii = live ? i : 0;  // lanes past the end compute on object 0 and write nothing
  State s{a.r0[ii], a.r0[n + ii], a.r0[2 * n + ii], a.v0[ii], a.v0[n + ii], a.v0[2 * n + ii]};
  const ObjParams p = object_params(a, ii);
  const bool moving = p.activity >= 0;  // dummies keep their state
  const unsigned gmask = a.mask & kGroupMask[w];
  int parity = 0;
  for (int step = 0; step < a.nsteps; ++step) {
    if (w == 0 && live && step % a.every == 0) store_state(s, a.ck + (size_t)(step / a.every) * 6 * n, n, i);
    group_step(s, p, gmask, moving, a.cdt, a.ddt, a.env + 3 * step, atm, x, parity);
  }
  if (w == 0 && live) {
    a.rT[i] = s.x;
    a.rT[n + i] = s.y;
    a.rT[2 * n + i] = s.z;
    a.vT[i] = s.vx;
    a.vT[n + i] = s.vy;
    a.vT[2 * n + i] = s.vz;
  }
}

__global__ __launch_bounds__(kAdjWaves * kAdjLanes) void k_adjoint_backward(AdjointArgs a) {
  __shared__ AtmTable atm;
  __shared__ GroupExchange x;
  stage_atm(atm);
  const int w = threadIdx.x / kAdjLanes, l = threadIdx.x % kAdjLanes;
  const size_t n = a.n, i = (size_t)blockIdx.x * kAdjLanes + l;
  const bool live = i < n;
  const size_t ii = live ? i : 0;
  double lx = a.g_rT ? a.g_rT[ii] : 0.0, ly = a.g_rT ? a.g_rT[n + ii] : 0.0, lz = a.g_rT ? a.g_rT[2 * n + ii] : 0.0;
  double lvx = a.g_vT ? a.g_vT[ii] : 0.0, lvy = a.g_vT ? a.g_vT[n + ii] : 0.0,
         lvz = a.g_vT ? a.g_vT[2 * n + ii] : 0.0;
  double lbc = 0.0;
  const ObjParams p = object_params(a, ii);
  const bool moving = p.activity >= 0;
  const unsigned gmask = a.mask & kGroupMask[w];
  int parity = 0;
  for (int seg = a.nck - 1; seg >= 0; --seg) {
    const int s0 = seg * a.every;
    const int s1 = min(s0 + a.every, a.nsteps);
    // Recompute the step-start states of the segment.  Every wavefront holds the same
    // values (same operations on the same inputs) and stores them to the same addresses,
    // so each thread reads back exactly what it wrote.
    State s;
    load_state(s, a.ck + (size_t)seg * 6 * n, n, ii);
    for (int step = s0; step < s1; ++step) {
      if (live) store_state(s, a.work + (size_t)(step - s0) * 6 * n, n, i);
      if (step + 1 < s1) group_step(s, p, gmask, moving, a.cdt, a.ddt, a.env + 3 * step, atm, x, parity);
    }
    for (int step = s1 - 1; step >= s0; --step) {
      if (live) load_state(s, a.work + (size_t)(step - s0) * 6 * n, n, i);
      const SubstepEnv<double> *env = a.env + 3 * step;
      // forward through the step, keeping the kick inputs (r_k, v_k)
      double rx[3], ry[3], rz[3], ux[3], uy[3], uz[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        s.x += a.cdt[k] * s.vx;
        s.y += a.cdt[k] * s.vy;
        s.z += a.cdt[k] * s.vz;
        rx[k] = s.x;
        ry[k] = s.y;
        rz[k] = s.z;
        ux[k] = s.vx;
        uy[k] = s.vy;
        uz[k] = s.vz;
        if (k < 2) {
          double ax, ay, az;
          group_acceleration(s, p, gmask, env[k], atm, x, parity, ax, ay, az);
          s.vx += a.ddt[k] * ax;
          s.vy += a.ddt[k] * ay;
          s.vz += a.ddt[k] * az;
        }
      }
      // reverse sweep
      lvx += a.cdt[3] * lx;
      lvy += a.cdt[3] * ly;
      lvz += a.cdt[3] * lz;
#pragma unroll
      for (int k = 2; k >= 0; --k) {
        const double wx = a.ddt[k] * lvx, wy = a.ddt[k] * lvy, wz = a.ddt[k] * lvz;
        double g[7];
        acceleration_vjp(rx[k], ry[k], rz[k], ux[k], uy[k], uz[k], p, gmask, env[k], atm, wx, wy, wz, g[0], g[1],
                         g[2], g[3], g[4], g[5], g[6]);
        group_sum<7>(x, parity, g);
        lx += g[0];
        ly += g[1];
        lz += g[2];
        lvx += g[3];
        lvy += g[4];
        lvz += g[5];
        lbc += g[6];
        lvx += a.cdt[k] * lx;
        lvy += a.cdt[k] * ly;
        lvz += a.cdt[k] * lz;
      }
    }
  }
  if (w == 0 && live) {
    // dummies: no gradient flows through a deleted object
    const double keep = moving ? 1.0 : 0.0;
    a.g_r0[i] = moving ? lx : (a.g_rT ? a.g_rT[i] : 0.0);
    a.g_r0[n + i] = moving ? ly : (a.g_rT ? a.g_rT[n + i] : 0.0);
    a.g_r0[2 * n + i] = moving ? lz : (a.g_rT ? a.g_rT[2 * n + i] : 0.0);
    a.g_v0[i] = moving ? lvx : (a.g_vT ? a.g_vT[i] : 0.0);
    a.g_v0[n + i] = moving ? lvy : (a.g_vT ? a.g_vT[n + i] : 0.0);
    a.g_v0[2 * n + i] = moving ? lvz : (a.g_vT ? a.g_vT[2 * n + i] : 0.0);
    if (a.g_bc) a.g_bc[i] = keep * lbc;
  }
}

}  // namespace ladds

using namespace ladds;

struct ladds_adjoint {
  ladds_adjoint_config cfg{};
  int every = 64, nck = 0;
  std::string err;
  SubstepEnv<double> *env = nullptr;
  double *work = nullptr;
  double cdt[4], ddt[3];
  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  int hipfail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? LADDS_E_NOMEM : LADDS_E_HIP;
  }
};

#define ADJCHK(p, call)                                   \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return (p)->hipfail(e_, #call); \
  } while (0)

extern "C" {

int ladds_adjoint_create(const ladds_adjoint_config *cfg, ladds_adjoint **out) {
  if (!cfg || !out || cfg->n == 0 || cfg->nsteps < 1 || cfg->checkpoint_every < 0 || !(cfg->dt > 0.0))
    return LADDS_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) return LADDS_E_NODEVICE;
  auto *p = new ladds_adjoint;
  p->cfg = *cfg;
  p->every = cfg->checkpoint_every > 0 ? cfg->checkpoint_every : 64;
  if (p->every > cfg->nsteps) p->every = cfg->nsteps;
  p->nck = (cfg->nsteps + p->every - 1) / p->every;
  double c[4], d[3];
  yoshida_coefficients(c, d);
  for (int k = 0; k < 4; ++k) p->cdt[k] = c[k] * cfg->dt;
  for (int k = 0; k < 3; ++k) p->ddt[k] = d[k] * cfg->dt;
  const double tau[3] = {c[0], c[0] + c[1], c[0] + c[1] + c[2]};
  std::vector<SubstepEnv<double>> env(3 * (size_t)cfg->nsteps);
  for (int s = 0; s < cfg->nsteps; ++s) {
    const double t = cfg->t0 + (double)s * cfg->dt;
    for (int k = 0; k < 3; ++k) {
      SubstepEnv<float> e32;
      substep_env(t + tau[k] * cfg->dt, env[3 * s + k], e32);
    }
  }
  hipError_t e = hipSetDevice(cfg->device);
  AtmTable t;
  fill_atm_table(t);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_atm_adj), &t, sizeof(t));
  if (e == hipSuccess) e = hipMalloc(&p->env, sizeof(SubstepEnv<double>) * env.size());
  if (e == hipSuccess) e = hipMemcpy(p->env, env.data(), sizeof(SubstepEnv<double>) * env.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&p->work, sizeof(double) * 6 * cfg->n * (size_t)p->every);
  if (e != hipSuccess) {
    hipFree(p->env);
    hipFree(p->work);
    delete p;
    return e == hipErrorOutOfMemory ? LADDS_E_NOMEM : LADDS_E_HIP;
  }
  *out = p;
  return LADDS_OK;
}

int ladds_adjoint_destroy(ladds_adjoint *p) {
  if (!p) return LADDS_E_INVALID;
  (void)hipSetDevice(p->cfg.device);
  hipFree(p->env);
  hipFree(p->work);
  delete p;
  return LADDS_OK;
}

const char *ladds_adjoint_last_error(const ladds_adjoint *p) { return p ? p->err.c_str() : "null plan"; }

uint64_t ladds_adjoint_checkpoint_doubles(const ladds_adjoint *p) {
  return p ? (uint64_t)p->nck * 6u * p->cfg.n : 0u;
}

static AdjointArgs make_args(const ladds_adjoint *p) {
  AdjointArgs a{};
  a.n = p->cfg.n;
  a.nsteps = p->cfg.nsteps;
  a.every = p->every;
  a.nck = p->nck;
  a.mask = p->cfg.components;
  for (int k = 0; k < 4; ++k) a.cdt[k] = p->cdt[k];
  for (int k = 0; k < 3; ++k) a.ddt[k] = p->ddt[k];
  a.env = p->env;
  a.work = p->work;
  return a;
}

int ladds_adjoint_forward(ladds_adjoint *p, const double *r0, const double *v0, const double *aom,
                          const double *bc_inv, const double *cr, const int32_t *activity, double *rT, double *vT,
                          double *checkpoints, void *stream) {
  if (!p) return LADDS_E_INVALID;
  if (!r0 || !v0 || !rT || !vT || !checkpoints) return p->fail(LADDS_E_INVALID, "null state or checkpoint pointer");
  ADJCHK(p, hipSetDevice(p->cfg.device));
  AdjointArgs a = make_args(p);
  a.aom = aom;
  a.bc = bc_inv;
  a.cr = cr;
  a.activity = activity;
  a.r0 = r0;
  a.v0 = v0;
  a.rT = rT;
  a.vT = vT;
  a.ck = checkpoints;
  const unsigned blocks = (unsigned)((a.n + kAdjLanes - 1) / kAdjLanes);
  hipLaunchKernelGGL(k_adjoint_forward, dim3(blocks), dim3(kAdjWaves * kAdjLanes), 0, (hipStream_t)stream, a);
  ADJCHK(p, hipGetLastError());
  return LADDS_OK;
}

int ladds_adjoint_backward(ladds_adjoint *p, const double *aom, const double *bc_inv, const double *cr,
                           const int32_t *activity, const double *checkpoints, const double *g_rT,
                           const double *g_vT, double *g_r0, double *g_v0, double *g_bc, void *stream) {
  if (!p) return LADDS_E_INVALID;
  if (!checkpoints || !g_r0 || !g_v0) return p->fail(LADDS_E_INVALID, "null checkpoint or gradient pointer");
  ADJCHK(p, hipSetDevice(p->cfg.device));
  AdjointArgs a = make_args(p);
  a.aom = aom;
  a.bc = bc_inv;
  a.cr = cr;
  a.activity = activity;
  a.ck = const_cast<double *>(checkpoints);
  a.g_rT = g_rT;
  a.g_vT = g_vT;
  a.g_r0 = g_r0;
  a.g_v0 = g_v0;
  a.g_bc = g_bc;
  const unsigned blocks = (unsigned)((a.n + kAdjLanes - 1) / kAdjLanes);
  hipLaunchKernelGGL(k_adjoint_backward, dim3(blocks), dim3(kAdjWaves * kAdjLanes), 0, (hipStream_t)stream, a);
  ADJCHK(p, hipGetLastError());
  return LADDS_OK;
}

}  // extern "C"

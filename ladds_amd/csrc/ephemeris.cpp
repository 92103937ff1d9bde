// ephemeris.cpp -- Yoshida coefficients and the device evaluation of the per-sub-step
// environments (the functions themselves are in ephemeris.h, shared with the host).
#include "ephemeris.h"

namespace ladds {

void yoshida_coefficients(double c[4], double d[3]) {
  const double cb = std::cbrt(2.0);
  const double w1 = 1.0 / (2.0 - cb);
  const double w0 = -cb / (2.0 - cb);
  c[0] = c[3] = w1 / 2.0;
  c[1] = c[2] = (w0 + w1) / 2.0;
  d[0] = d[2] = w1;
  d[1] = w0;
}

__global__ __launch_bounds__(64) void k_substep_env(double t0, double dt, uint64_t base, double tau0, double tau1,
                                                    double tau2, int first, int count, SubstepEnv<double> *env64,
                                                    SubstepEnv<float> *env32) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * count) return;
  const int s = first + k / 3, sub = k % 3;
  const double t = t0 + (double)(base + (uint64_t)s) * dt;
  const double tau = sub == 0 ? tau0 : (sub == 1 ? tau1 : tau2);
  SubstepEnv<double> e64;
  SubstepEnv<float> e32;
  substep_env(t + tau * dt, e64, e32);
  env64[3 * s + sub] = e64;
  env32[3 * s + sub] = e32;
}

hipError_t launch_substep_env(double t0, double dt, uint64_t base, const double tau[3], int first, int count,
                              SubstepEnv<double> *env64, SubstepEnv<float> *env32, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((3 * count + 63) / 64);
  hipLaunchKernelGGL(k_substep_env, dim3(blocks), dim3(64), 0, stream, t0, dt, base, tau[0], tau[1], tau[2], first,
                     count, env64, env32);
  return hipGetLastError();
}

}  // namespace ladds

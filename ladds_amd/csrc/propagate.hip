// propagate.hip -- Yoshida-4 propagation kernel (one thread per object).
//
// Replaces YoshidaIntegrator<AutoPas_t>::integrate (OrbitPropagator @ b90f3bae, called at
// src/ladds/Simulation.cpp:297).  One step is the 4th-order composition of Yoshida
// (Phys. Lett. A 150 (1990) 262):
//   r += c1 dt v; v += d1 dt a(r, v, t + c1 dt)
//   r += c2 dt v; v += d2 dt a(r, v, t + (c1+c2) dt)
//   r += c3 dt v; v += d3 dt a(r, v, t + (c1+c2+c3) dt)
//   r += c4 dt v
// with w1 = 1/(2 - 2^(1/3)), w0 = -2^(1/3)/(2 - 2^(1/3)), c1 = c4 = w1/2,
// c2 = c3 = (w0+w1)/2, d1 = d3 = w1, d2 = w0.  Velocity-dependent drag uses the
// latest kicked velocity; time-dependent terms use the sub-step time.
//
// Layout: SoA FP64 state in HBM, one coalesced 8-byte load/store per component.
// Several steps per launch keep the state in registers between steps (used when
// no conjunction scan is due in between).  When `count` is set, the kernel also
// performs the counting pass of the cell binning for the final positions.
#include "binning.h"
#include "launch.h"
#include "physics.h"

namespace ladds {

__constant__ AtmTable c_atm;

struct PropagateArgs {
  DeviceObjects o;
  const SubstepEnv<double> *env64;  // 3 per step
  const SubstepEnv<float> *env32;
  int nsteps;
  unsigned mask;
  double cdt[4], ddt[3];
  GridParams grid;
  uint32_t *keys;   // non-null: write the bucket key of the binning point
  uint32_t *count;  // non-null: counting pass of the atomic counting sort
  uint32_t *slot;
  double burnup_r2;  // > 0: objects with |r|^2 below become dummies (deleteBurnUps)
  unsigned long long *burnups;  // non-null: count of objects flagged here
  const StepState *step_state;  // non-null (graph replay): environments start at step_state->step + step_offset
  uint32_t step_offset;
  // output state (the input arrays of `o` when propagating in place)
  double *orx, *ory, *orz, *ovx, *ovy, *ovz;
  int32_t *oact;
};

template <int FORM>
__global__ __launch_bounds__(256) void k_propagate(PropagateArgs a) {
  __shared__ AtmTable atm;
  {
    const double *src = reinterpret_cast<const double *>(&c_atm);
    double *dst = reinterpret_cast<double *>(&atm);
    for (int k = threadIdx.x; k < (int)(sizeof(AtmTable) / sizeof(double)); k += blockDim.x) dst[k] = src[k];
    __syncthreads();
  }
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.o.n) return;
  const size_t env0 = a.step_state ? 3u * ((size_t)a.step_state->step + a.step_offset) : 0u;
  double x = a.o.rx[i], y = a.o.ry[i], z = a.o.rz[i];
  double vx = a.o.vx[i], vy = a.o.vy[i], vz = a.o.vz[i];
  ObjParams p;
  p.activity = a.o.activity[i];
  // dummies (burnt-up objects awaiting compaction) keep their state
  const int nsteps = p.activity < 0 ? 0 : a.nsteps;
  p.drag64 = a.o.drag64[i];
  if constexpr (FORM == kFormFp64) {
    p.srp64 = a.o.srp64[i];
    p.srp32 = 0.f;
  } else {
    p.srp32 = a.o.srp32[i];
    p.srp64 = 0.0;
  }
  bool burnt = false;
  for (int s = 0; s < nsteps; ++s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double cdt = a.cdt[k], ddt = a.ddt[k];
      x += cdt * vx;
      y += cdt * vy;
      z += cdt * vz;
      double ax, ay, az;
      acceleration<FORM>(x, y, z, vx, vy, vz, p, a.mask, a.env64[env0 + 3 * s + k], a.env32[env0 + 3 * s + k],
                         atm, ax, ay, az);
      vx += ddt * ax;
      vy += ddt * ay;
      vz += ddt * az;
    }
    x += a.cdt[3] * vx;
    y += a.cdt[3] * vy;
    z += a.cdt[3] * vz;
    // Simulation::deleteBurnUps after every integrate (Simulation.cpp:301-303, 589-590)
    if (a.burnup_r2 > 0.0 && exact_dot3(x, y, z, x, y, z) < a.burnup_r2) {
      burnt = true;
      break;
    }
  }
  if (burnt) {
    a.oact[i] = -1;
    if (a.burnups) atomicAdd(a.burnups, 1ull);  // rare: one atomic per burnt-up object
  }
  else if (a.oact != a.o.activity)
    a.oact[i] = p.activity;
  a.orx[i] = x;
  a.ory[i] = y;
  a.orz[i] = z;
  a.ovx[i] = vx;
  a.ovy[i] = vy;
  a.ovz[i] = vz;
  if (a.keys || a.count) {
    binning_point<FORM != kFormFp64>(x, y, z);
    const CellInfo c = cell_of(x, y, z, a.grid);
    if (a.keys)
      a.keys[i] = c.bucket;
    else
      a.slot[i] = atomicAdd(&a.count[c.bucket], 1u);
  }
}

hipError_t upload_atmosphere() {
  AtmTable t;
  fill_atm_table(t);
  return hipMemcpyToSymbol(HIP_SYMBOL(c_atm), &t, sizeof(t));
}

hipError_t launch_propagate(int form, const DeviceObjects &o, const SubstepEnv<double> *env64,
                            const SubstepEnv<float> *env32, int nsteps, unsigned mask, const double cdt[4],
                            const double ddt[3], const GridParams *grid, uint32_t *keys, uint32_t *count,
                            uint32_t *slot, double burnup_r2, const StepState *step_state, hipStream_t stream,
                            const DeviceObjects *out, uint32_t step_offset, unsigned long long *burnups) {
  if (o.n == 0 || nsteps <= 0) return hipSuccess;
  PropagateArgs a;
  a.o = o;
  const DeviceObjects &w = out ? *out : o;
  a.orx = w.rx;
  a.ory = w.ry;
  a.orz = w.rz;
  a.ovx = w.vx;
  a.ovy = w.vy;
  a.ovz = w.vz;
  a.oact = w.activity;
  a.env64 = env64;
  a.env32 = env32;
  a.nsteps = nsteps;
  a.mask = mask;
  for (int k = 0; k < 4; ++k) a.cdt[k] = cdt[k];
  for (int k = 0; k < 3; ++k) a.ddt[k] = ddt[k];
  if (grid) a.grid = *grid;
  a.keys = keys;
  a.count = count;
  a.slot = slot;
  a.burnup_r2 = burnup_r2;
  a.burnups = burnups;
  a.step_state = step_state;
  a.step_offset = step_offset;
  const unsigned blocks = (unsigned)((o.n + 255) / 256);
  switch (form) {
    case kFormFp64:
      hipLaunchKernelGGL(k_propagate<kFormFp64>, dim3(blocks), dim3(256), 0, stream, a);
      break;
    case kFormSplit:
      hipLaunchKernelGGL(k_propagate<kFormSplit>, dim3(blocks), dim3(256), 0, stream, a);
      break;
    default:
      hipLaunchKernelGGL(k_propagate<kFormSplitAll32>, dim3(blocks), dim3(256), 0, stream, a);
      break;
  }
  return hipGetLastError();
}

__global__ void k_set_step(StepState *st, uint32_t step, uint64_t iter_base) {
  if (threadIdx.x == 0) {
    st->step = step;
    st->pad = 0;
    st->iter_base = iter_base;
  }
}

__global__ void k_advance_step(StepState *st) {
  if (threadIdx.x == 0) st->step += 1u;
}

hipError_t launch_set_step(StepState *st, uint32_t step, uint64_t iter_base, hipStream_t stream) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, stream, st, step, iter_base);
  return hipGetLastError();
}

hipError_t launch_advance_step(StepState *st, hipStream_t stream) {
  hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, stream, st);
  return hipGetLastError();
}

// Counters to the host: a kernel storing into coherent mapped host memory finishes a
// synchronised readback in ~11 us on MI355X, an 80-byte hipMemcpyAsync in ~21 us.
__global__ void k_publish(const unsigned long long *src, unsigned long long *dst, int words) {
  for (int k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
}

hipError_t launch_publish(const unsigned long long *src, unsigned long long *dst, int words, hipStream_t stream) {
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, src, dst, words);
  return hipGetLastError();
}

}  // namespace ladds

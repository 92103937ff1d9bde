// export.hip -- device-side packing of the output records LADDS's HDF5 writer stores.
//
// HDF5Writer::writeParticles (src/ladds/io/hdf5/HDF5Writer.cpp:89-149) stores per output
// iteration float32 Vec3 positions and velocities and uint32 ids
// (HDF5Definitions.h:25-37); writeConjunctions / writeEvasions (HDF5Writer.cpp:151-187)
// store one {uint32 idA, uint32 idB, float32 d^2} record per pair of one iteration
// (HDF5Definitions.h:41-64).  Both conversions run here on the resident data, so only
// the narrow records cross PCIe.  Casts: double -> float rounds to nearest, uint64 ->
// uint32 keeps the low 32 bits (the reference's static_cast).
#include <hip/hip_runtime.h>

#include "launch.h"

namespace ladds {

namespace {
constexpr int kThreads = 256;
}

__global__ __launch_bounds__(kThreads) void k_particle_records(DeviceObjects o, size_t n, float *pos, float *vel,
                                                               uint32_t *ids) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  pos[3 * i + 0] = (float)o.rx[i];
  pos[3 * i + 1] = (float)o.ry[i];
  pos[3 * i + 2] = (float)o.rz[i];
  vel[3 * i + 0] = (float)o.vx[i];
  vel[3 * i + 1] = (float)o.vy[i];
  vel[3 * i + 2] = (float)o.vz[i];
  ids[i] = (uint32_t)o.id[i];
}

// Pairs of one iteration from the device log, split into collisions and evasions and
// compacted per wavefront (one atomic per wavefront and kind).  counts[0..1] must be zero
// on entry and receive the full counts; records past a capacity are not written.
__global__ __launch_bounds__(kThreads) void k_log_records(const LoggedPair *log, const unsigned long long *log_count,
                                                          uint64_t log_cap, uint64_t iteration, PairRecord32 *hits,
                                                          uint64_t hcap, PairRecord32 *ev, uint64_t ecap,
                                                          unsigned long long *counts) {
  const unsigned long long total = *log_count;
  const uint64_t n = total < log_cap ? total : log_cap;
  const int lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  // the loop bound is uniform across the block, so every wavefront reaches each ballot
  for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < n; base += (uint64_t)gridDim.x * kThreads) {
    const uint64_t i = base + threadIdx.x;
    bool is_hit = false, is_ev = false;
    PairRecord32 r{};
    if (i < n && log[i].iteration == iteration) {
      const LoggedPair &e = log[i];
      r.id_a = (uint32_t)e.pair.id_a;
      r.id_b = (uint32_t)e.pair.id_b;
      r.d2 = (float)e.pair.d2;
      is_ev = e.evaded != 0;
      is_hit = !is_ev;
    }
    const uint64_t bh = __ballot(is_hit), be = __ballot(is_ev);
    unsigned long long oh = 0, oe = 0;
    if (lane == 0) {
      if (bh) oh = atomicAdd(&counts[0], (unsigned long long)__popcll(bh));
      if (be) oe = atomicAdd(&counts[1], (unsigned long long)__popcll(be));
    }
    oh = __shfl(oh, 0, 64);
    oe = __shfl(oe, 0, 64);
    if (is_hit) {
      const uint64_t p = oh + (uint64_t)__popcll(bh & below);
      if (p < hcap) hits[p] = r;
    }
    if (is_ev) {
      const uint64_t p = oe + (uint64_t)__popcll(be & below);
      if (p < ecap) ev[p] = r;
    }
  }
}

// Gather of the objects a conjunction list names (the drop-in adapter refreshes only the
// Particle entries a detection returns, instead of downloading every object).
__global__ __launch_bounds__(kThreads) void k_gather_state(DeviceObjects o, const uint32_t *idx, size_t count,
                                                           double *out) {
  const size_t k = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const uint32_t i = idx[k];
  double *d = out + 6 * k;
  d[0] = o.rx[i];
  d[1] = o.ry[i];
  d[2] = o.rz[i];
  d[3] = o.vx[i];
  d[4] = o.vy[i];
  d[5] = o.vz[i];
}

hipError_t launch_gather_state(const DeviceObjects &o, const uint32_t *idx, size_t count, double *out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((count + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_gather_state, dim3(blocks), dim3(kThreads), 0, s, o, idx, count, out);
  return hipGetLastError();
}

hipError_t launch_particle_records(const DeviceObjects &o, size_t n, float *pos, float *vel, uint32_t *ids,
                                   hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_particle_records, dim3(blocks), dim3(kThreads), 0, s, o, n, pos, vel, ids);
  return hipGetLastError();
}

hipError_t launch_log_records(const LoggedPair *log, const unsigned long long *log_count, uint64_t log_cap,
                              uint64_t iteration, PairRecord32 *hits, uint64_t hcap, PairRecord32 *ev, uint64_t ecap,
                              unsigned long long *counts, hipStream_t s) {
  const uint64_t want = (log_cap + kThreads - 1) / kThreads;
  const unsigned blocks = (unsigned)(want < 2048 ? (want ? want : 1) : 2048);
  hipLaunchKernelGGL(k_log_records, dim3(blocks), dim3(kThreads), 0, s, log, log_count, log_cap, iteration, hits,
                     hcap, ev, ecap, counts);
  return hipGetLastError();
}

}  // namespace ladds

// compact.hip -- stable on-device removal of burnt-up objects.
//
// Replaces the deletion half of Simulation::deleteBurnUps (src/ladds/Simulation.cpp:561-596)
// and AutoPas::updateContainer's removal of dummies: objects flagged as dummies
// (activity < 0) are dropped and the survivors are packed in their original order.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace ladds {

namespace {
constexpr int kThreads = 256;
constexpr int kItems = kCompactTile / kThreads;  // 8

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace

__global__ __launch_bounds__(kThreads) void k_count_alive(const int32_t *activity, size_t n, uint32_t *tile_sum) {
  __shared__ uint32_t lds[kThreads / 64];
  const size_t base = (size_t)blockIdx.x * kCompactTile + (size_t)threadIdx.x * kItems;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) c += (base + k < n && activity[base + k] >= 0) ? 1u : 0u;
  c = wave_sum(c);
  if (__lane_id() == 0) lds[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < kThreads / 64; ++w) s += lds[w];
    tile_sum[blockIdx.x] = s;
  }
}

template <typename T>
__device__ __forceinline__ void move(const T *src, T *dst, size_t from, size_t to) {
  dst[to] = src[from];
}

__global__ __launch_bounds__(kThreads) void k_compact(DeviceObjects src, DeviceObjects dst, const uint32_t *tile_sum,
                                                      uint32_t ntiles, uint32_t *n_alive) {
  __shared__ uint32_t lds[kThreads / 64];
  __shared__ uint32_t wsum[kThreads / 64];
  uint32_t part = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kThreads) part += tile_sum[j];
  part = wave_sum(part);
  if (__lane_id() == 0) lds[threadIdx.x >> 6] = part;
  __syncthreads();
  uint32_t offset = 0;
  for (int w = 0; w < kThreads / 64; ++w) offset += lds[w];

  const size_t base = (size_t)blockIdx.x * kCompactTile + (size_t)threadIdx.x * kItems;
  bool alive[kItems];
  uint32_t tot = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    alive[k] = base + k < src.n && src.activity[base + k] >= 0;
    tot += alive[k] ? 1u : 0u;
  }
  uint32_t inc = tot;
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += t;
  }
  const uint32_t w = threadIdx.x >> 6;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t wb = 0;
  for (uint32_t k = 0; k < w; ++k) wb += wsum[k];
  uint32_t out = offset + wb + inc - tot;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (!alive[k]) continue;
    const size_t f = base + k;
    move(src.rx, dst.rx, f, out);
    move(src.ry, dst.ry, f, out);
    move(src.rz, dst.rz, f, out);
    move(src.vx, dst.vx, f, out);
    move(src.vy, dst.vy, f, out);
    move(src.vz, dst.vz, f, out);
    move(src.radius_m, dst.radius_m, f, out);
    move(src.aom, dst.aom, f, out);
    move(src.bc_inv, dst.bc_inv, f, out);
    move(src.cr, dst.cr, f, out);
    move(src.srp64, dst.srp64, f, out);
    move(src.drag64, dst.drag64, f, out);
    move(src.srp32, dst.srp32, f, out);
    move(src.drag32, dst.drag32, f, out);
    move(src.activity, dst.activity, f, out);
    move(src.id, dst.id, f, out);
    move(src.parent, dst.parent, f, out);
    ++out;
  }
  if (blockIdx.x == ntiles - 1 && threadIdx.x == kThreads - 1) *n_alive = out;
}

hipError_t launch_count_alive(const DeviceObjects &o, uint32_t *tile_sum, uint32_t ntiles, hipStream_t s) {
  hipLaunchKernelGGL(k_count_alive, dim3(ntiles), dim3(kThreads), 0, s, o.activity, o.n, tile_sum);
  return hipGetLastError();
}

hipError_t launch_compact(const DeviceObjects &src, const DeviceObjects &dst, const uint32_t *tile_sum,
                          uint32_t ntiles, uint32_t *n_alive, hipStream_t s) {
  hipLaunchKernelGGL(k_compact, dim3(ntiles), dim3(kThreads), 0, s, src, dst, tile_sum, ntiles, n_alive);
  return hipGetLastError();
}

}  // namespace ladds

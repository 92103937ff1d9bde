// compact.hip -- stable on-device removal of burnt-up objects, and the reset of parent
// ids that ends the spawn protection of breakup fragments.
//
// Compaction replaces the deletion half of Simulation::deleteBurnUps
// (src/ladds/Simulation.cpp:561-596) and AutoPas::updateContainer's removal of dummies:
// objects flagged as dummies (activity < 0) are dropped and the survivors are packed in
// their original order.  A tile of kCompactTile objects is read as 8 rounds of 256
// consecutive objects (coalesced loads); the survivors' positions come from wavefront
// ballots, so the stores of a round are contiguous as well.
//
// Spawn protection: Simulation::simulationLoop resets every parent id when 100 iterations
// have passed without a collision (Simulation.cpp:287-293, 399-402, 438, 470-473).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "launch.h"

namespace ladds {

namespace {
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = kCompactTile / kThreads;  // 8

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace

__global__ __launch_bounds__(kThreads) void k_count_alive(const int32_t *activity, size_t n, uint32_t *tile_sum) {
  __shared__ uint32_t lds[kWaves];
  const size_t base = (size_t)blockIdx.x * kCompactTile + threadIdx.x;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kRounds; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    c += (i < n && activity[i] >= 0) ? 1u : 0u;
  }
  c = wave_sum(c);
  if (__lane_id() == 0) lds[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < kWaves; ++w) s += lds[w];
    tile_sum[blockIdx.x] = s;
  }
}

template <typename T>
__device__ __forceinline__ void move(const T *src, T *dst, size_t from, size_t to) {
  dst[to] = src[from];
}

__global__ __launch_bounds__(kThreads) void k_compact(DeviceObjects src, DeviceObjects dst, const uint32_t *tile_sum,
                                                      uint32_t ntiles, uint32_t *n_alive) {
  __shared__ uint32_t part[kWaves];
  __shared__ uint32_t cnt[kRounds * kWaves];  // survivors per (round, wave), then their exclusive prefix
  // survivors of the tiles before this one
  uint32_t before = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kThreads) before += tile_sum[j];
  before = wave_sum(before);
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  if (lane == 0) part[w] = before;
  const size_t base = (size_t)blockIdx.x * kCompactTile + threadIdx.x;
  unsigned long long mask[kRounds];
#pragma unroll
  for (int k = 0; k < kRounds; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    mask[k] = __ballot(i < src.n && src.activity[i] >= 0);
    if (lane == 0) cnt[k * kWaves + w] = (uint32_t)__popcll(mask[k]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int q = 0; q < kWaves; ++q) run += part[q];
    for (int q = 0; q < kRounds * kWaves; ++q) {
      const uint32_t c = cnt[q];
      cnt[q] = run;
      run += c;
    }
    if (blockIdx.x == ntiles - 1) *n_alive = run;
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < kRounds; ++k) {
    if (!((mask[k] >> lane) & 1ull)) continue;
    const size_t f = base + (size_t)k * kThreads;
    const size_t out = cnt[k * kWaves + w] + (uint32_t)__popcll(mask[k] & below);
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
  }
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

// ---------------------------------------------------------------------------
// spawn protection
// ---------------------------------------------------------------------------
constexpr int kSpawnBlocks = 256;

// Every parent id reset (Simulation::removeParticleSpawnProtection, Simulation.cpp:470-473).
// With `st`, only when the count of quiet iterations reaches 100 at the start of one of the
// `span` iterations being accounted: their starts have counts quiet .. quiet + span - 1.
__global__ __launch_bounds__(kThreads) void k_spawn_pre(SpawnState *st, uint64_t span, uint64_t *parent, size_t n) {
  if (st) {
    const unsigned long long q = st->quiet;
    if (!(q <= 100ull && q + span - 1ull >= 100ull)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->cleared = 1ull;
  }
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
    parent[i] = ~0ull;
}

// After the last of `span` iterations (with a detection when `detect`): the count at its
// start is quiet + span - 1; a detection that found collisions (ScanCounters::hits grew)
// sets it to 0 (Simulation.cpp:399-402); the end of the iteration adds one (:438).
__global__ void k_spawn_post(SpawnState *st, uint64_t span, const ScanCounters *counters, int detect) {
  if (threadIdx.x != 0) return;
  unsigned long long c = st->quiet + span - 1ull;
  const unsigned long long hits = counters->hits;
  if (detect && hits != st->seen) c = 0ull;
  st->seen = hits;
  st->quiet = c + 1ull;
}

hipError_t launch_clear_parents(SpawnState *st, uint64_t span, uint64_t *parent, size_t n, hipStream_t s) {
  if (n == 0 || span == 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>(kSpawnBlocks, (n + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_spawn_pre, dim3(blocks), dim3(kThreads), 0, s, st, span, parent, n);
  return hipGetLastError();
}

hipError_t launch_spawn_post(SpawnState *st, uint64_t span, const ScanCounters *counters, bool detect,
                             hipStream_t s) {
  if (span == 0) return hipSuccess;
  hipLaunchKernelGGL(k_spawn_post, dim3(1), dim3(64), 0, s, st, span, counters, detect ? 1 : 0);
  return hipGetLastError();
}

}  // namespace ladds

// conjunction.hip -- device binning and the pair scan (the CollisionFunctor traversal).
//
// Replaces AutoPas::iteratePairwise(&CollisionFunctor) with the LinkedCells /
// Newton3 traversal configured at src/ladds/Simulation.cpp:89-92, and
// CollisionFunctor::AoSFunctor (src/ladds/CollisionFunctor.cpp:40-131) with
// utils::distanceByLinearInterpolation (src/ladds/utils/DistanceApproximation.cpp:13-58).
//
// Pipeline per detection:
//   count    bucket of every object, slot = atomicAdd(count[bucket])   (fused into
//            the propagation kernel when a scan follows a propagation)
//   reduce   per-tile sums of the bucket counts
//   scan     exclusive prefix of the counts -> bucket starts; counts re-zeroed
//   scatter  sorted record (position, search-block corner) at start[bucket] + slot
//   pairs    per sorted record: its 2x2x2 search block = 4 contiguous z-runs of
//            buckets; every later record there passes the cutoff prefilter (FP64
//            record, or FP32 record with a conservatively inflated radius), then the
//            exact FP64 predicate; hits are compacted per wavefront with a ballot
//            and one atomic per wavefront.
#include <hip/hip_runtime.h>

#include "binning.h"
#include "launch.h"
#include "pairs.h"

namespace ladds {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 2048 buckets per tile

__device__ __forceinline__ uint32_t block_reduce_sum(uint32_t v, uint32_t *lds) {
  v = wave_reduce_sum(v);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  uint32_t s = 0;
  for (uint32_t k = 0; k < blockDim.x / 64; ++k) s += lds[k];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
template <bool F32>
__global__ __launch_bounds__(256) void k_bin_count(const double *rx, const double *ry, const double *rz, size_t n,
                                                   GridParams g, uint32_t *count, uint32_t *slot) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = rx[i], y = ry[i], z = rz[i];
  binning_point<F32>(x, y, z);
  const CellInfo c = cell_of(x, y, z, g);
  slot[i] = atomicAdd(&count[c.bucket], 1u);
}

__global__ __launch_bounds__(kScanThreads) void k_bin_reduce(const uint32_t *count, uint32_t *tile_sum) {
  __shared__ uint32_t lds[kScanThreads / 64];
  const uint4 *c4 = reinterpret_cast<const uint4 *>(count + (size_t)blockIdx.x * kScanTile);
  const uint4 u0 = c4[2 * threadIdx.x], u1 = c4[2 * threadIdx.x + 1];
  uint32_t s = u0.x + u0.y + u0.z + u0.w + u1.x + u1.y + u1.z + u1.w;
  s = block_reduce_sum(s, lds);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = s;
}

__global__ __launch_bounds__(kScanThreads) void k_bin_scan(uint32_t *count, const uint32_t *tile_sum,
                                                           uint32_t *start, uint32_t ntiles) {
  __shared__ uint32_t lds[kScanThreads / 64];
  __shared__ uint32_t wsum[kScanThreads / 64];
  uint32_t part = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kScanThreads) part += tile_sum[j];
  const uint32_t offset = block_reduce_sum(part, lds);

  uint4 *c4 = reinterpret_cast<uint4 *>(count + (size_t)blockIdx.x * kScanTile);
  const uint4 u0 = c4[2 * threadIdx.x], u1 = c4[2 * threadIdx.x + 1];
  const uint32_t v[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
  uint32_t ex[8];
  uint32_t tot = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    ex[k] = tot;
    tot += v[k];
  }
  // inclusive wave scan of thread totals
  uint32_t inc = tot;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += t;
  }
  const uint32_t w = threadIdx.x >> 6;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t wbase = 0;
  for (uint32_t k = 0; k < w; ++k) wbase += wsum[k];
  const uint32_t texcl = offset + wbase + inc - tot;
  uint4 *s4 = reinterpret_cast<uint4 *>(start + (size_t)blockIdx.x * kScanTile);
  s4[2 * threadIdx.x] = make_uint4(texcl + ex[0], texcl + ex[1], texcl + ex[2], texcl + ex[3]);
  s4[2 * threadIdx.x + 1] = make_uint4(texcl + ex[4], texcl + ex[5], texcl + ex[6], texcl + ex[7]);
  c4[2 * threadIdx.x] = make_uint4(0, 0, 0, 0);
  c4[2 * threadIdx.x + 1] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == ntiles - 1 && threadIdx.x == kScanThreads - 1) start[(size_t)ntiles * kScanTile] = texcl + tot;
}

template <typename REC>
__global__ __launch_bounds__(256) void k_bin_scatter(const double *rx, const double *ry, const double *rz, size_t n,
                                                     GridParams g, const uint32_t *slot, const uint32_t *start,
                                                     REC *rec) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr bool f32 = sizeof(REC) == sizeof(RecordF32);
  double x = rx[i], y = ry[i], z = rz[i];
  binning_point<f32>(x, y, z);
  const CellInfo c = cell_of(x, y, z, g);
  const uint32_t p = start[c.bucket] + slot[i];
  REC r;
  r.x = x;
  r.y = y;
  r.z = z;
  r.idx = (uint32_t)i;
  if constexpr (!f32) r.pad = 0;
  rec[p] = r;
}

struct PairScanArgs {
  PairSink sink;
  GridParams g;
  const void *rec;
  const uint32_t *start;
};

template <int FORM>
__device__ __forceinline__ bool prefilter(const PairScanArgs &a, const void *recs, uint32_t q, double px, double py,
                                          double pz, float fx, float fy, float fz, uint32_t &oj) {
  if constexpr (FORM == kFormFp64) {
    const RecordF64 r = static_cast<const RecordF64 *>(recs)[q];
    const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    oj = r.idx;
    return !(dot3(dx, dy, dz, dx, dy, dz) > a.sink.fp.cutoff2);
  } else {
    const RecordF32 r = static_cast<const RecordF32 *>(recs)[q];
    const float dx = fx - r.x, dy = fy - r.y, dz = fz - r.z;
    oj = r.idx;
    return !(dx * dx + dy * dy + dz * dz > a.sink.fp.cutoff2_f32);
  }
}

template <int FORM, bool DIGEST>
__global__ __launch_bounds__(256) void k_pair_scan(PairScanArgs a) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  ScanAcc acc;
  if (p < a.sink.o.n) {
    double px = 0, py = 0, pz = 0;
    float fx = 0, fy = 0, fz = 0;
    uint32_t oi;
    if constexpr (FORM == kFormFp64) {
      const RecordF64 r = static_cast<const RecordF64 *>(a.rec)[p];
      px = r.x;
      py = r.y;
      pz = r.z;
      oi = r.idx;
    } else {
      const RecordF32 r = static_cast<const RecordF32 *>(a.rec)[p];
      fx = r.x;
      fy = r.y;
      fz = r.z;
      px = fx;
      py = fy;
      pz = fz;
      oi = r.idx;
    }
    const uint32_t base = cell_of(px, py, pz, a.g).base;
    const uint32_t mx = (1u << a.g.bx) - 1u, my = (1u << a.g.by) - 1u, mz = (1u << a.g.bz) - 1u;
    const uint32_t sx = a.g.by + a.g.bz, sy = a.g.bz;
    const uint32_t bxs = base >> sx, bys = (base >> sy) & my, bzs = base & mz;
    const uint32_t bz1 = (bzs + 1u) & mz;
#pragma unroll 1
    for (int c = 0; c < 4; ++c) {
      const uint32_t fxb = (bxs + (uint32_t)(c >> 1)) & mx;
      const uint32_t fyb = (bys + (uint32_t)(c & 1)) & my;
      const uint32_t row = (fxb << sx) | (fyb << sy);
      // z-run 1: buckets bzs..bzs+1 when contiguous, else bzs then 0
      uint32_t lo = a.start[row + bzs];
      uint32_t hi = a.start[row + bzs + (bz1 ? 2u : 1u)];
#pragma unroll 1
      for (int run = 0; run < 2; ++run) {
        uint32_t q = lo > p + 1 ? lo : p + 1;
#pragma unroll 1
        for (; q < hi; ++q) {
          uint32_t oj;
          if (prefilter<FORM>(a, a.rec, q, px, py, pz, fx, fy, fz, oj)) handle_candidate<DIGEST>(a.sink, oi, oj, acc);
        }
        if (bz1 != 0u) break;
        lo = a.start[row];
        hi = a.start[row + 1];
      }
    }
  }
  flush_counters<DIGEST>(a.sink, acc);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_bin_count(int form, const DeviceObjects &o, const GridParams &g, uint32_t *count, uint32_t *slot,
                            hipStream_t s) {
  if (o.n == 0) return hipSuccess;
  const dim3 grid((unsigned)((o.n + 255) / 256));
  if (form == kFormFp64)
    hipLaunchKernelGGL(k_bin_count<false>, grid, dim3(256), 0, s, o.rx, o.ry, o.rz, o.n, g, count, slot);
  else
    hipLaunchKernelGGL(k_bin_count<true>, grid, dim3(256), 0, s, o.rx, o.ry, o.rz, o.n, g, count, slot);
  return hipGetLastError();
}

hipError_t launch_bin_reduce(const GridParams &g, const uint32_t *count, uint32_t *tile_sum, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_reduce, dim3(g.nbuckets / kScanTile), dim3(kScanThreads), 0, s, count, tile_sum);
  return hipGetLastError();
}

hipError_t launch_bin_scan(const GridParams &g, uint32_t *count, const uint32_t *tile_sum, uint32_t *start,
                           hipStream_t s) {
  const uint32_t ntiles = g.nbuckets / kScanTile;
  hipLaunchKernelGGL(k_bin_scan, dim3(ntiles), dim3(kScanThreads), 0, s, count, tile_sum, start, ntiles);
  return hipGetLastError();
}

hipError_t launch_bin_scatter(int form, const DeviceObjects &o, const GridParams &g, const uint32_t *slot,
                              const uint32_t *start, void *rec, hipStream_t s) {
  if (o.n == 0) return hipSuccess;
  const dim3 grid((unsigned)((o.n + 255) / 256));
  if (form == kFormFp64)
    hipLaunchKernelGGL(k_bin_scatter<RecordF64>, grid, dim3(256), 0, s, o.rx, o.ry, o.rz, o.n, g, slot, start,
                       static_cast<RecordF64 *>(rec));
  else
    hipLaunchKernelGGL(k_bin_scatter<RecordF32>, grid, dim3(256), 0, s, o.rx, o.ry, o.rz, o.n, g, slot, start,
                       static_cast<RecordF32 *>(rec));
  return hipGetLastError();
}

hipError_t launch_pair_scan(int form, bool digest, const DeviceObjects &o, const GridParams &g,
                            const FunctorParams &fp, const void *rec, const uint32_t *start,
                            LoggedPair *log, uint64_t log_cap, unsigned long long *log_count,
                            ScanCounters *counters, uint64_t iteration, hipStream_t s) {
  if (o.n == 0) return hipSuccess;
  PairScanArgs a;
  a.sink = PairSink{o, fp, log, log_cap, log_count, counters, iteration, nullptr};
  a.g = g;
  a.rec = rec;
  a.start = start;
  const dim3 grid((unsigned)((o.n + 255) / 256));
  const bool f64 = (form == kFormFp64);
  if (f64 && !digest) hipLaunchKernelGGL((k_pair_scan<kFormFp64, false>), grid, dim3(256), 0, s, a);
  if (f64 && digest) hipLaunchKernelGGL((k_pair_scan<kFormFp64, true>), grid, dim3(256), 0, s, a);
  if (!f64 && !digest) hipLaunchKernelGGL((k_pair_scan<kFormSplit, false>), grid, dim3(256), 0, s, a);
  if (!f64 && digest) hipLaunchKernelGGL((k_pair_scan<kFormSplit, true>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace ladds

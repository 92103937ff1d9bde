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

namespace ladds {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 2048 buckets per tile

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

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

// ---------------------------------------------------------------------------
// exact FP64 predicate: CollisionFunctor::AoSFunctor + distanceByLinearInterpolation
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return exact_dot3(a0, a1, a2, b0, b1, b2);
}

// Returns 0 (no report), 1 (collision) or 2 (evaded); i < j are object indices, evaluated
// as AoSFunctor(i, j).
__device__ __forceinline__ int collision_functor(const DeviceObjects &o, uint32_t i, uint32_t j,
                                              const FunctorParams &fp, PairRecord &out) {
#pragma clang fp contract(off)
  const int ai = o.activity[i], aj = o.activity[j];
  if (ai < 0 || aj < 0) return 0;  // deleted (dummy) objects, CollisionFunctor.cpp:50-52
  const double Ri = o.radius_m[i], Rj = o.radius_m[j];
  const bool evaded = (ai != 0 && aj != 0) || (ai != 0 && Rj >= fp.min_det_radius) ||
                      (aj != 0 && Ri >= fp.min_det_radius);
  const uint64_t pi = o.parent[i], pj = o.parent[j];
  if (pi != ~0ull && pi == pj) return 0;
  const double rix = o.rx[i], riy = o.ry[i], riz = o.rz[i];
  const double rjx = o.rx[j], rjy = o.ry[j], rjz = o.rz[j];
  const double dx = rix - rjx, dy = riy - rjy, dz = riz - rjz;
  const double d0 = dot3(dx, dy, dz, dx, dy, dz);
  if (d0 > fp.cutoff2) return 0;
  const double vix = o.vx[i], viy = o.vy[i], viz = o.vz[i];
  const double vjx = o.vx[j], vjy = o.vy[j], vjz = o.vz[j];
  const double dt = fp.dt;
  // DistanceApproximation.cpp:25-55
  const double oix = rix - vix * dt, oiy = riy - viy * dt, oiz = riz - viz * dt;
  const double ojx = rjx - vjx * dt, ojy = rjy - vjy * dt, ojz = rjz - vjz * dt;
  const double vi_ri = dot3(vix, viy, viz, oix, oiy, oiz);
  const double vi_rj = dot3(vix, viy, viz, ojx, ojy, ojz);
  const double vj_ri = dot3(vjx, vjy, vjz, oix, oiy, oiz);
  const double vj_rj = dot3(vjx, vjy, vjz, ojx, ojy, ojz);
  const double nominator = vi_rj + vj_ri - vi_ri - vj_rj;
  const double two_vi_vj = 2.0 * dot3(vix, viy, viz, vjx, vjy, vjz);
  const double vi_square = dot3(vix, viy, viz, vix, viy, viz);
  const double vj_square = dot3(vjx, vjy, vjz, vjx, vjy, vjz);
  const double denominator = vi_square + vj_square - two_vi_vj;
  double t = nominator / denominator;
  t = (t < 0.0) ? 0.0 : ((dt < t) ? dt : t);  // std::clamp: NaN passes through
  const double p2x = ojx + vjx * t, p2y = ojy + vjy * t, p2z = ojz + vjz * t;
  const double drx = (oix + vix * t) - p2x;
  const double dry = (oiy + viy * t) - p2y;
  const double drz = (oiz + viz * t) - p2z;
  const double dl = dot3(drx, dry, drz, drx, dry, drz);
  if (evaded) {
    if (dl > fp.evasion2) return 0;
  } else if (fp.threshold_mode == 1) {
    if (dl > fp.threshold2) return 0;
  } else {
    const double scaled = (Ri + Rj) * fp.cdf_km;
    if (dl > scaled * scaled) return 0;
  }
  out.mid[0] = p2x + drx * 0.5;
  out.mid[1] = p2y + dry * 0.5;
  out.mid[2] = p2z + drz * 0.5;
  out.d2 = dl;
  const uint64_t idi = o.id[i], idj = o.id[j];
  if (idi < idj) {
    out.id_a = idi;
    out.id_b = idj;
    out.idx_a = i;
    out.idx_b = j;
  } else {
    out.id_a = idj;
    out.id_b = idi;
    out.idx_a = j;
    out.idx_b = i;
  }
  return evaded ? 2 : 1;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct PairScanArgs {
  DeviceObjects o;
  GridParams g;
  FunctorParams fp;
  const void *rec;
  const uint32_t *start;
  LoggedPair *log;
  uint64_t log_cap;
  unsigned long long *log_count;
  ScanCounters *counters;
  uint64_t iteration;
};

struct ScanAcc {
  unsigned long long cand = 0, hit = 0, ev = 0, d0 = 0, d1 = 0, e0 = 0, e1 = 0;
};

template <bool DIGEST>
__device__ __forceinline__ void handle_candidate(const PairScanArgs &a, uint32_t oi, uint32_t oj, ScanAcc &acc) {
  const uint32_t i = oi < oj ? oi : oj;
  const uint32_t j = oi < oj ? oj : oi;
  PairRecord r;
  const int kind = collision_functor(a.o, i, j, a.fp, r);
  acc.cand += 1;
  if constexpr (DIGEST) {
    if (kind != 0) {
      const uint64_t h = mix64(r.id_a ^ mix64(r.id_b));
      const uint64_t h2 = mix64(h ^ 0xD1B54A32D192ED03ull);
      if (kind == 1) {
        acc.hit += 1;
        acc.d0 += h;
        acc.d1 += h2;
      } else {
        acc.ev += 1;
        acc.e0 += h;
        acc.e1 += h2;
      }
    }
  } else {
    // wavefront compaction: one atomic per wavefront for the lanes holding a result
    const unsigned long long hm = __ballot(kind == 1);
    const unsigned long long em = __ballot(kind == 2);
    const unsigned long long all = hm | em;
    if (all) {
      const int leader = __ffsll((long long)all) - 1;
      unsigned long long base = 0;
      if ((int)lane_id() == leader) {
        base = atomicAdd(a.log_count, (unsigned long long)__popcll(all));
        if (hm) atomicAdd(&a.counters->hits, (unsigned long long)__popcll(hm));
        if (em) atomicAdd(&a.counters->evaded, (unsigned long long)__popcll(em));
      }
      base = __shfl(base, leader, 64);
      if (kind != 0) {
        const unsigned long long below = all & ((1ull << lane_id()) - 1ull);
        const unsigned long long pos = base + (unsigned long long)__popcll(below);
        if (pos < a.log_cap) {
          LoggedPair lp;
          lp.pair = r;
          lp.iteration = a.iteration;
          lp.evaded = (kind == 2) ? 1u : 0u;
          lp.reserved = 0;
          a.log[pos] = lp;
        }
      }
    }
  }
}

template <int FORM>
__device__ __forceinline__ bool prefilter(const PairScanArgs &a, const void *recs, uint32_t q, double px, double py,
                                          double pz, float fx, float fy, float fz, uint32_t &oj) {
  if constexpr (FORM == kFormFp64) {
    const RecordF64 r = static_cast<const RecordF64 *>(recs)[q];
    const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    oj = r.idx;
    return !(dot3(dx, dy, dz, dx, dy, dz) > a.fp.cutoff2);
  } else {
    const RecordF32 r = static_cast<const RecordF32 *>(recs)[q];
    const float dx = fx - r.x, dy = fy - r.y, dz = fz - r.z;
    oj = r.idx;
    return !(dx * dx + dy * dy + dz * dz > a.fp.cutoff2_f32);
  }
}

template <int FORM, bool DIGEST>
__global__ __launch_bounds__(256) void k_pair_scan(PairScanArgs a) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  ScanAcc acc;
  if (p < a.o.n) {
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
          if (prefilter<FORM>(a, a.rec, q, px, py, pz, fx, fy, fz, oj)) handle_candidate<DIGEST>(a, oi, oj, acc);
        }
        if (bz1 != 0u) break;
        lo = a.start[row];
        hi = a.start[row + 1];
      }
    }
  }
  // one atomic per wavefront for the counters
  const unsigned long long cand = wave_reduce_sum(acc.cand);
  if constexpr (DIGEST) {
    const unsigned long long hit = wave_reduce_sum(acc.hit), ev = wave_reduce_sum(acc.ev);
    const unsigned long long d0 = wave_reduce_sum(acc.d0), d1 = wave_reduce_sum(acc.d1);
    const unsigned long long e0 = wave_reduce_sum(acc.e0), e1 = wave_reduce_sum(acc.e1);
    if (lane_id() == 0) {
      if (hit) {
        atomicAdd(&a.counters->hits, hit);
        atomicAdd(&a.counters->digest[0], d0);
        atomicAdd(&a.counters->digest[1], d1);
      }
      if (ev) {
        atomicAdd(&a.counters->evaded, ev);
        atomicAdd(&a.counters->edigest[0], e0);
        atomicAdd(&a.counters->edigest[1], e1);
      }
    }
  }
  if (lane_id() == 0 && cand) atomicAdd(&a.counters->candidates, cand);
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
  a.o = o;
  a.g = g;
  a.fp = fp;
  a.rec = rec;
  a.start = start;
  a.log = log;
  a.log_cap = log_cap;
  a.log_count = log_count;
  a.counters = counters;
  a.iteration = iteration;
  const dim3 grid((unsigned)((o.n + 255) / 256));
  const bool f64 = (form == kFormFp64);
  if (f64 && !digest) hipLaunchKernelGGL((k_pair_scan<kFormFp64, false>), grid, dim3(256), 0, s, a);
  if (f64 && digest) hipLaunchKernelGGL((k_pair_scan<kFormFp64, true>), grid, dim3(256), 0, s, a);
  if (!f64 && !digest) hipLaunchKernelGGL((k_pair_scan<kFormSplit, false>), grid, dim3(256), 0, s, a);
  if (!f64 && digest) hipLaunchKernelGGL((k_pair_scan<kFormSplit, true>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace ladds

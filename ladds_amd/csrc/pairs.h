// pairs.h -- the exact FP64 conjunction predicate and the hit sink shared by the
// pair-scan kernels.
//
// collision_functor restates CollisionFunctor::AoSFunctor (src/ladds/CollisionFunctor.cpp:40-131)
// with utils::distanceByLinearInterpolation (src/ladds/utils/DistanceApproximation.cpp:13-58)
// operation for operation, without FMA contraction, so its results are bit-identical to
// the reference's (x86-64 baseline flags) for the same inputs.
#pragma once

#include <hip/hip_runtime.h>

#include "binning.h"
#include "engine.h"

namespace ladds {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// exact FP64 predicate: CollisionFunctor::AoSFunctor + distanceByLinearInterpolation
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return exact_dot3(a0, a1, a2, b0, b1, b2);
}

// What the predicate reads of one object: object `i` of the object arrays, or a copy of
// its fields stored beside the sorted records of a dense plane pair (FatRecord).
struct SoAObject {
  const DeviceObjects &o;
  uint32_t i;
  __device__ __forceinline__ int activity() const { return o.activity[i]; }
  __device__ __forceinline__ double radius() const { return o.radius_m[i]; }
  __device__ __forceinline__ uint64_t parent() const { return o.parent[i]; }
  __device__ __forceinline__ uint64_t id() const { return o.id[i]; }
  __device__ __forceinline__ double rx() const { return o.rx[i]; }
  __device__ __forceinline__ double ry() const { return o.ry[i]; }
  __device__ __forceinline__ double rz() const { return o.rz[i]; }
  __device__ __forceinline__ double vx() const { return o.vx[i]; }
  __device__ __forceinline__ double vy() const { return o.vy[i]; }
  __device__ __forceinline__ double vz() const { return o.vz[i]; }
};

struct alignas(16) FatRecord {
  double r[3], v[3];
  double radius_m;
  uint64_t parent, id;
  int32_t activity;
  uint32_t pad;
};

__device__ __forceinline__ FatRecord fat_record_of(const DeviceObjects &o, uint32_t i) {
  FatRecord f;
  f.r[0] = o.rx[i];
  f.r[1] = o.ry[i];
  f.r[2] = o.rz[i];
  f.v[0] = o.vx[i];
  f.v[1] = o.vy[i];
  f.v[2] = o.vz[i];
  f.radius_m = o.radius_m[i];
  f.parent = o.parent[i];
  f.id = o.id[i];
  f.activity = o.activity[i];
  f.pad = 0;
  return f;
}

struct FatObject {
  const FatRecord &f;
  __device__ __forceinline__ int activity() const { return f.activity; }
  __device__ __forceinline__ double radius() const { return f.radius_m; }
  __device__ __forceinline__ uint64_t parent() const { return f.parent; }
  __device__ __forceinline__ uint64_t id() const { return f.id; }
  __device__ __forceinline__ double rx() const { return f.r[0]; }
  __device__ __forceinline__ double ry() const { return f.r[1]; }
  __device__ __forceinline__ double rz() const { return f.r[2]; }
  __device__ __forceinline__ double vx() const { return f.v[0]; }
  __device__ __forceinline__ double vy() const { return f.v[1]; }
  __device__ __forceinline__ double vz() const { return f.v[2]; }
};

// Returns 0 (no report), 1 (collision) or 2 (evaded); objects `a` (index i) and `b`
// (index j), i < j, evaluated as AoSFunctor(i, j).
template <typename OBJ>
__device__ __forceinline__ int collision_functor(const OBJ &a, const OBJ &b, uint32_t i, uint32_t j,
                                              const FunctorParams &fp, PairRecord &out) {
#pragma clang fp contract(off)
  const int ai = a.activity(), aj = b.activity();
  if (ai < 0 || aj < 0) return 0;  // deleted (dummy) objects, CollisionFunctor.cpp:50-52
  const double Ri = a.radius(), Rj = b.radius();
  const bool evaded = (ai != 0 && aj != 0) || (ai != 0 && Rj >= fp.min_det_radius) ||
                      (aj != 0 && Ri >= fp.min_det_radius);
  const uint64_t pi = a.parent(), pj = b.parent();
  if (pi != ~0ull && pi == pj) return 0;
  const double rix = a.rx(), riy = a.ry(), riz = a.rz();
  const double rjx = b.rx(), rjy = b.ry(), rjz = b.rz();
  const double dx = rix - rjx, dy = riy - rjy, dz = riz - rjz;
  const double d0 = dot3(dx, dy, dz, dx, dy, dz);
  if (d0 > fp.cutoff2) return 0;
  const double vix = a.vx(), viy = a.vy(), viz = a.vz();
  const double vjx = b.vx(), vjy = b.vy(), vjz = b.vz();
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
  const uint64_t idi = a.id(), idj = b.id();
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

// Where hits go: the device pair log (list mode) or the counters (digest mode).
struct PairSink {
  DeviceObjects o;
  FunctorParams fp;
  LoggedPair *log;
  uint64_t log_cap;
  unsigned long long *log_count;
  ScanCounters *counters;
  uint64_t iteration;
  const StepState *step_state;  // non-null (graph replay): iteration = iter_base + step
};

__device__ __forceinline__ unsigned long long sink_iteration(const PairSink &a) {
  return a.step_state ? a.step_state->iter_base + a.step_state->step : a.iteration;
}

struct ScanAcc {
  unsigned long long cand = 0, hit = 0, ev = 0, d0 = 0, d1 = 0, e0 = 0, e1 = 0;
};

// A predicate result to the sink: digest sums, or the log (wavefront compaction).
template <bool DIGEST>
__device__ __forceinline__ void sink_result(const PairSink &a, int kind, const PairRecord &r, ScanAcc &acc,
                                            uint32_t scenario) {
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
        if (hm) {
          atomicAdd(&a.counters->hits, (unsigned long long)__popcll(hm));
          atomicMax(&a.counters->last_hit, sink_iteration(a) + 1ull);
        }
        if (em) atomicAdd(&a.counters->evaded, (unsigned long long)__popcll(em));
      }
      base = __shfl(base, leader, 64);
      if (kind != 0) {
        const unsigned long long below = all & ((1ull << lane_id()) - 1ull);
        const unsigned long long pos = base + (unsigned long long)__popcll(below);
        if (pos < a.log_cap) {
          LoggedPair lp;
          lp.pair = r;
          lp.iteration = sink_iteration(a);
          lp.evaded = (kind == 2) ? 1u : 0u;
          lp.scenario = scenario;
          a.log[pos] = lp;
        }
      }
    }
  }
}

// The predicate of objects oi and oj read from the object arrays.
template <bool DIGEST>
__device__ __forceinline__ void handle_candidate(const PairSink &a, uint32_t oi, uint32_t oj, ScanAcc &acc,
                                                 uint32_t scenario = 0u) {
  const uint32_t i = oi < oj ? oi : oj;
  const uint32_t j = oi < oj ? oj : oi;
  PairRecord r;
  const int kind = collision_functor(SoAObject{a.o, i}, SoAObject{a.o, j}, i, j, a.fp, r);
  sink_result<DIGEST>(a, kind, r, acc, scenario);
}

// The same predicate with both objects' fields read from their FatRecords fi (object oi)
// and fj (object oj).
template <bool DIGEST>
__device__ __forceinline__ void handle_candidate(const PairSink &a, const FatRecord &fi, const FatRecord &fj,
                                                 uint32_t oi, uint32_t oj, ScanAcc &acc, uint32_t scenario) {
  const bool lo = oi < oj;
  PairRecord r;
  const int kind = collision_functor(FatObject{lo ? fi : fj}, FatObject{lo ? fj : fi}, lo ? oi : oj, lo ? oj : oi,
                                     a.fp, r);
  sink_result<DIGEST>(a, kind, r, acc, scenario);
}

// Flush the per-lane counters: one atomic per wavefront.  All lanes of the wave must call it.
template <bool DIGEST>
__device__ __forceinline__ void flush_counters(const PairSink &a, const ScanAcc &acc) {
  const unsigned long long cand = wave_reduce_sum(acc.cand);
  if constexpr (DIGEST) {
    const unsigned long long hit = wave_reduce_sum(acc.hit), ev = wave_reduce_sum(acc.ev);
    const unsigned long long d0 = wave_reduce_sum(acc.d0), d1 = wave_reduce_sum(acc.d1);
    const unsigned long long e0 = wave_reduce_sum(acc.e0), e1 = wave_reduce_sum(acc.e1);
    if (lane_id() == 0) {
      if (hit) {
        atomicAdd(&a.counters->hits, hit);
        atomicMax(&a.counters->last_hit, sink_iteration(a) + 1ull);
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

}  // namespace ladds

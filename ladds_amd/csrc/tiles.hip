// tiles.hip -- atomic-free binning and the LDS-tiled pair scan.
//
// Replaces AutoPas::iteratePairwise(&CollisionFunctor) over LinkedCells
// (src/ladds/Simulation.cpp:89-92,211-222).  The folded grid of binning.h is viewed
// as 2^bx x-planes of 2^pb buckets (pb = by + bz).  Per detection:
//
//   keys       bucket of every object's binning point (written by the propagation
//              kernel, or by k_bin_keys when no propagation precedes the scan)
//   partition  one workgroup per tile of consecutive objects: LDS histogram of the
//              planes, LDS prefix, then every object's record is written into the
//              tile's own region grouped by plane (contiguous, coalesced writes) and
//              the tile's per-plane offsets are stored.  No global atomics.
//   pairs      one workgroup per plane: gathers the plane's records and those of the
//              next plane from every tile, counting-sorts both by bucket in LDS, and
//              scans the 3x3x3 cells around each record from LDS (cells of side >= R,
//              so a plane is one cell thick and holds half the objects of a 2R plane;
//              buckets two cells tall or wide are scanned two at a time, on the side of
//              the half of the bucket the record lies in).
//              Pairs inside the plane are taken from the lower sorted position, pairs
//              with the next plane from the plane below, so every pair within the
//              search radius is seen exactly once.  Candidates go through the exact
//              FP64 predicate (pairs.h); hits are compacted per wavefront.
//   A plane whose records exceed the LDS capacity is sorted into a global scratch
//   arena instead (same code, global storage).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "binning.h"
#include "launch.h"
#include "pairs.h"

namespace ladds {

namespace {

constexpr int kPartThreads = 1024;
// Pair-kernel workgroups of 256, 512 or 1024 threads (TileParams::pair_threads).  The LDS
// footprint of a plane pair fixes how many workgroups share a CU; larger workgroups give
// the latency-bound scan more waves when only one fits (the FP64 records are twice the size).
constexpr int kMaxPairThreads = 1024;
constexpr int kPairRegCapacity = 2048;  // records kept in registers between the two sort passes (at most)
constexpr uint32_t kChunkRecords = 512;  // own records per chunk of a dense plane

// In-place exclusive scan of a[0..n) in LDS by the whole block; returns the total.
// `wsum` holds blockDim/64 + 1 entries.
__device__ uint32_t block_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *wsum) {
  const uint32_t nt = blockDim.x, tid = threadIdx.x;
  const uint32_t per = (n + nt - 1) / nt;
  const uint32_t lo = min(n, tid * per), hi = min(n, lo + per);
  uint32_t tot = 0;
  for (uint32_t k = lo; k < hi; ++k) tot += a[k];
  uint32_t inc = tot;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += t;
  }
  const uint32_t w = tid >> 6, nw = nt >> 6;
  __syncthreads();  // a[] reads above are complete before anyone rewrites
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < nw; ++k) {
      const uint32_t v = wsum[k];
      wsum[k] = run;
      run += v;
    }
    wsum[nw] = run;
  }
  __syncthreads();
  uint32_t run = wsum[w] + inc - tot;
  for (uint32_t k = lo; k < hi; ++k) {
    const uint32_t v = a[k];
    a[k] = run;
    run += v;
  }
  const uint32_t total = wsum[nw];
  __syncthreads();
  return total;
}

template <typename REC>
__device__ __forceinline__ void binning_point_of(const REC &r, double &x, double &y, double &z) {
  x = (double)r.x;
  y = (double)r.y;
  z = (double)r.z;
}

}  // namespace

// ---------------------------------------------------------------------------
// keys and partition
// ---------------------------------------------------------------------------
// LDS words of the partition kernel before its staging area (32-byte aligned).
__host__ __device__ constexpr uint32_t partition_header_words(uint32_t nplanes) {
  return (nplanes + kPartThreads / 64 + 1 + 7u) & ~7u;
}

template <bool F32>
__global__ __launch_bounds__(256) void k_bin_keys(const double *rx, const double *ry, const double *rz, size_t n,
                                                  GridParams g, uint32_t *keys) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = rx[i], y = ry[i], z = rz[i];
  binning_point<F32>(x, y, z);
  keys[i] = cell_of(x, y, z, g).bucket;
}

// Records are placed in LDS first and the tile's region is then written with contiguous
// stores (scattered 16/32-byte stores into the region cost ~3x the bytes in partial-line
// write-backs).  A tile whose records exceed the LDS staging capacity (`stage_cap`
// records) is staged in several passes over consecutive ranges of its region.
template <typename REC, int kPartItems>
__global__ __launch_bounds__(kPartThreads) void k_tile_partition(const double *rx, const double *ry,
                                                                  const double *rz, size_t n, const uint32_t *keys,
                                                                  TileParams tp, REC *items, uint32_t *tps,
                                                                  uint32_t *arena_top, uint32_t stage_cap) {
  extern __shared__ uint32_t lds[];
  uint32_t *cnt = lds;                // nplanes
  uint32_t *wsum = lds + tp.nplanes;  // kPartThreads / 64 + 1
  REC *staged = reinterpret_cast<REC *>(lds + partition_header_words(tp.nplanes));
  const uint32_t P = tp.nplanes, pb = tp.plane_bits;
  // objects [base, end) of this tile
  size_t base, end;
  if (tp.tile_begin) {
    base = tp.tile_begin[blockIdx.x];
    end = tp.tile_begin[blockIdx.x + 1];
  } else {
    base = (size_t)blockIdx.x * tp.tile_size;
    end = min(n, base + tp.tile_size);
  }
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) cnt[p] = 0;
  // the pair scan's counters: arena fill and the persistent kernel's queue positions
  if (blockIdx.x == 0 && threadIdx.x < 1 + kXcdCount) arena_top[threadIdx.x * kQueueStride] = 0;
  __syncthreads();
  // every thread owns kPartItems objects of the tile; their loads are issued together
  uint32_t key[kPartItems];
#pragma unroll
  for (int k = 0; k < kPartItems; ++k) {
    const size_t i = base + threadIdx.x + (size_t)k * kPartThreads;
    key[k] = (i < end) ? keys[i] : ~0u;
  }
#pragma unroll
  for (int k = 0; k < kPartItems; ++k)
    if (key[k] != ~0u) atomicAdd(&cnt[key[k] >> pb], 1u);
  __syncthreads();
  const uint32_t total = block_exclusive_scan(cnt, P, wsum);
  uint32_t *row = tps + (size_t)blockIdx.x * (P + 1);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) row[p] = cnt[p];
  if (threadIdx.x == 0) row[P] = total;
  __syncthreads();
  constexpr bool f32 = sizeof(REC) == sizeof(RecordF32);
  double px[kPartItems], py[kPartItems], pz[kPartItems];
#pragma unroll
  for (int k = 0; k < kPartItems; ++k) {
    const size_t i = base + threadIdx.x + (size_t)k * kPartThreads;
    if (key[k] != ~0u) {
      px[k] = rx[i];
      py[k] = ry[i];
      pz[k] = rz[i];
    }
  }
  uint32_t pos[kPartItems];
#pragma unroll
  for (int k = 0; k < kPartItems; ++k) pos[k] = key[k] != ~0u ? atomicAdd(&cnt[key[k] >> pb], 1u) : ~0u;
#pragma unroll 1
  for (uint32_t lo = 0; lo < total; lo += stage_cap) {
    const uint32_t hi = min(total, lo + stage_cap);
#pragma unroll
    for (int k = 0; k < kPartItems; ++k) {
      if (pos[k] < lo || pos[k] >= hi) continue;
      double x = px[k], y = py[k], z = pz[k];
      binning_point<f32>(x, y, z);
      REC r;
      r.x = x;
      r.y = y;
      r.z = z;
      r.idx = (uint32_t)(base + threadIdx.x + (size_t)k * kPartThreads);
      if constexpr (!f32) r.pad = 0;
      staged[pos[k] - lo] = r;
    }
    __syncthreads();
    for (uint32_t j = lo + threadIdx.x; j < hi; j += kPartThreads) items[base + j] = staged[j - lo];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pair scan over plane tiles
// ---------------------------------------------------------------------------
struct PairTileArgs {
  PairSink sink;
  GridParams g;
  TileParams tp;
  const void *items;
  const uint32_t *tps;
  void *scratch;
  uint32_t *arena_top;
};

template <int FORM>
using RecOf = typename std::conditional<FORM == kFormFp64, RecordF64, RecordF32>::type;

template <int FORM>
__device__ __forceinline__ bool within(const RecOf<FORM> &a, const RecOf<FORM> &b, const FunctorParams &fp) {
  if constexpr (FORM == kFormFp64) {
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return !(exact_dot3(dx, dy, dz, dx, dy, dz) > fp.cutoff2);
  } else {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return !(dx * dx + dy * dy + dz * dz > fp.cutoff2_f32);
  }
}

// Cells have side >= the search radius R, so partners within R lie in the 3x3x3 cells
// around an object.  Within its own plane a pair is taken from the lower sorted
// position; pairs with the next plane are all taken here (the next plane's own scan
// never looks back), pairs with the previous plane by the previous plane's workgroup.
// Per object: six rows (three y-rows of its own plane, three of the next), or, with rows
// two cells tall (tp.half_rows), four (two of each plane, on the side of the half of its
// row the object lies in: fewer rows of the same buckets, so fewer folded aliases to
// reject); in each row three buckets, or two with buckets two cells wide along z
// (tp.half_buckets), the ones on the side of the half of its bucket.  In each row
// the buckets fz-1, fz, fz+1 (mod Kz) are one contiguous run of the sorted array plus,
// where the fold wraps (fz = 0 or Kz-1), a one-bucket run at the other end of the row.
// Both runs are read unconditionally (the second one empty when there is no wrap), so
// the lanes of a wavefront take the same path.  cur[b] is the end of bucket b in the
// sorted array and cur[-1] the start of bucket 0 (0 for the own plane, n_own for the
// next), so bucket b spans [cur[b-1], cur[b]).  The rows are
// walked by one loop so that the candidate path (the exact predicate and the wavefront
// compaction) exists once in the code.
// `fat` (or null): the predicate records at the sorted positions, for plane pairs sorted in
// the global arena.
template <int FORM, bool DIGEST, bool CHUNKED>
__device__ __forceinline__ void scan_plane(const PairTileArgs &a, const RecOf<FORM> *sorted,
                                           const uint32_t *cur_own, const uint32_t *cur_halo, uint32_t n_own,
                                           uint32_t threads, uint32_t scen, uint32_t chunk, uint32_t nchunk,
                                           const FatRecord *fat, ScanAcc &acc) {
  const GridParams &g = a.g;
  const uint32_t my = (1u << g.by) - 1u, mz = (1u << g.bz) - 1u;
  const uint32_t sy = g.bz;
#pragma unroll 1
  for (uint32_t s = threadIdx.x; s < n_own; s += threads) {
    const RecOf<FORM> me = sorted[s];
    // chunks split the plane's objects by object index: the placement order differs
    // between the chunks' own sorts, the index does not
    if (CHUNKED && me.idx % nchunk != chunk) continue;
    double x, y, z;
    binning_point_of(me, x, y, z);
    uint32_t ly, lz;  // 1: the record lies in the lower half of its row (bucket)
    const uint32_t lb = yz_bucket(y, z, g, ly, lz);
    const uint32_t fy = (lb >> sy) & my, fz = lb & mz;
    // buckets [zlo, zhi] of each row: fz-1, fz, fz+1 (one cell wide) or, two cells wide,
    // fz-1, fz (lower half) or fz, fz+1 (upper half); where that crosses the fold, the
    // bucket at the other end of the row (zw) is scanned as a second run
    uint32_t zlo, zhi;
    bool wraps;
    if (a.tp.half_buckets == 0u) {
      zlo = fz ? fz - 1u : 0u;
      zhi = fz < mz ? fz + 1u : mz;
      wraps = (fz == 0u) || (fz == mz);
    } else {
      const bool wl = fz == 0u && lz != 0u, wh = fz == mz && lz == 0u;
      zlo = wl ? 0u : fz - lz;
      zhi = wh ? mz : fz + 1u - lz;
      wraps = wl || wh;
    }
    const uint32_t zw = fz ? 0u : mz;  // the wrapped bucket when `wraps`
    const bool half = a.tp.half_rows != 0u;
#pragma unroll 1
    for (uint32_t kind = 0; kind < (half ? 4u : 6u); ++kind) {
      const bool next_plane = (kind & 1u) != 0u;
      // Own-plane pairs: with one workgroup per plane each pair is taken from its lower
      // sorted position.  The in-LDS placement order within a bucket depends on atomic
      // timing, so chunks of one plane (each sorting the plane itself) cannot share that
      // rule: they take a pair from its lower object index instead.
      const bool by_index = CHUNKED && !next_plane;
      // the row scanned, as an offset d (mod Ky) from the record's own row
      uint32_t d;
      if (!half) {
        // rows one cell tall: rows y-1, y, y+1 of both planes; own row y-1 sorts entirely
        // below this record unless the fold wraps it (fy = 0)
        if (!CHUNKED && kind == 0u && fy != 0u) continue;
        d = (kind >> 1) - 1u;
      } else if (next_plane || CHUNKED) {
        // rows two cells tall: partners within R (<= half a row) lie in rows y-1, y for a
        // record in the lower half of its row, y, y+1 for one in the upper half
        d = (kind >> 1) - ly;
      } else {
        // own plane, lower sorted position: row y (kind 2) and the row on the record's
        // side (kind 0), y+1 for the upper half, y-1 for the lower half, which sorts
        // entirely below this record unless the fold wraps it (fy = 0)
        if (kind == 0u && ly != 0u && fy != 0u) continue;
        d = kind == 2u ? 0u : (ly ? ~0u : 1u);
      }
      const uint32_t *cur = next_plane ? cur_halo : cur_own;
      const uint32_t row = ((fy + d) & my) << sy;
      const uint32_t b_lo = row | zlo, b_w = row | zw;
      const uint32_t lo0 = cur[(int32_t)b_lo - 1], hi0 = cur[row | zhi];
      const uint32_t w0 = cur[(int32_t)b_w - 1], w1 = cur[b_w];
      const uint32_t lo1 = wraps ? w0 : 0u, hi1 = wraps ? w1 : 0u;
      const uint32_t from = (next_plane || by_index) ? 0u : s + 1u;
      uint32_t q = max(lo0, from), end = hi0;
#pragma unroll 1
      for (int part = 0; part < 2; ++part) {
        // two candidates per iteration: their LDS reads are issued together
#pragma unroll 1
        for (; q < end; q += 2u) {
          const bool two = q + 1u < end;
          const RecOf<FORM> o0 = sorted[q];
          const RecOf<FORM> o1 = sorted[two ? q + 1u : q];
          bool c0 = within<FORM>(me, o0, a.sink.fp);
          bool c1 = two && within<FORM>(me, o1, a.sink.fp);
          if (CHUNKED && by_index) {
            c0 = c0 && o0.idx > me.idx;
            c1 = c1 && o1.idx > me.idx;
          }
          // candidates are rare (~1 in 10^4 tests at 1M objects): the exact predicate's
          // path is laid out as the cold one
          if (__builtin_expect(c0 || c1, 0)) {
#pragma unroll 1
            for (int u = 0; u < 2; ++u) {
              if (!(u ? c1 : c0)) continue;
              const uint32_t oj = u ? o1.idx : o0.idx;
              if (fat)
                handle_candidate<DIGEST>(a.sink, fat[s], fat[q + (uint32_t)u], me.idx, oj, acc, scen);
              else
                handle_candidate<DIGEST>(a.sink, me.idx, oj, acc, scen);
            }
          }
        }
        q = max(lo1, from);
        end = hi1;
      }
    }
  }
}

// One work item of the pair scan: plane pair `wid` = (scenario, plane, chunk), sorted in
// LDS up to tp.cap_items records, beyond that in the global scratch arena.
// FAT: plane pairs sorted in the arena carry the predicate's fields beside their records
// (instantiated only for launches whose shape left plane pairs beyond the LDS, so that the
// common all-in-LDS case keeps the smaller register footprint).
template <int FORM, bool DIGEST, int THREADS, bool FAT>
__device__ __forceinline__ void pair_plane(const PairTileArgs &a, uint32_t *lds, uint32_t wid, ScanAcc &acc) {
  constexpr int kPairRegItems = kPairRegCapacity / THREADS < 8 ? kPairRegCapacity / THREADS : 8;
  using REC = RecOf<FORM>;
  const TileParams &tp = a.tp;
  const uint32_t P = tp.nplanes, L = 1u << tp.plane_bits;
  // a scenario's tiles are [t0, t0 + T)
  const uint32_t chunk = wid % tp.nchunk, pidx = wid / tp.nchunk;
  const uint32_t scen = pidx / P;
  const uint32_t plane = pidx - scen * P, next = (plane + 1u) & (P - 1u);  // P >= 4
  const uint32_t t0 = tp.scen_tile ? tp.scen_tile[scen] : 0u;
  const uint32_t T = tp.scen_tile ? tp.scen_tile[scen + 1] - t0 : tp.ntiles;
  // LDS layout.  Segments: the plane's records in each of the T tiles, then the next
  // plane's records in each tile; seg[] is their prefix (2T + 1 entries), so record j of
  // the plane pair (own records first) lies in segment t with seg[t] <= j < seg[t+1].
  uint32_t *seg = lds;
  uint32_t *src = seg + (2u * T + 1u);     // 2T source offsets
  // Bucket cursors of both planes in one array [own L][1][next L] (cur_own[-1] = 0): one
  // prefix over it gives positions in a single sorted array, own records first, and leaves
  // cur_next[-1] = n_own, the start of the next plane's bucket 0.
  uint32_t *cur_own = src + 2u * T + 1u;
  uint32_t *cur_next = cur_own + L + 1u;
  uint32_t *wsum = cur_next + L;           // THREADS / 64 + 1
  uint32_t *misc = wsum + (kMaxPairThreads / 64 + 1);  // [0] scratch offset
  REC *lds_items = reinterpret_cast<REC *>(lds + tp.lds_header_words);

  const REC *items = static_cast<const REC *>(a.items);
  for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
    const uint32_t g = t0 + t;
    const uint32_t *row = a.tps + (size_t)g * (P + 1);
    const uint32_t o0 = row[plane], o1 = row[plane + 1];
    const uint32_t n0 = row[next], n1 = row[next + 1];
    const uint32_t tile_base = tp.tile_begin ? tp.tile_begin[g] : g * tp.tile_size;
    seg[t] = o1 - o0;
    seg[T + t] = n1 - n0;
    src[t] = tile_base + o0;
    src[T + t] = tile_base + n0;
  }
  for (uint32_t b = threadIdx.x; b < 2u * L + 1u; b += blockDim.x) cur_own[b] = 0;
  if (threadIdx.x == 0) cur_own[-1] = 0;
  __syncthreads();
  const uint32_t total = block_exclusive_scan(seg, 2u * T, wsum);
  const uint32_t n_own = seg[T];
  if (threadIdx.x == 0) seg[2u * T] = total;
  // chunks this plane uses: ~kChunkRecords own records each, at most tp.nchunk
  const uint32_t cp = min(tp.nchunk, max(1u, (n_own + kChunkRecords - 1u) / kChunkRecords));
  if (n_own != 0u && chunk < cp) {
    const bool in_lds = total <= tp.cap_items;
    if (!in_lds && threadIdx.x == 0) {
      const uint32_t off = atomicAdd(a.arena_top, total);
      misc[0] = off;
      misc[1] = (off <= tp.arena_cap && total <= tp.arena_cap - off) ? 1u : 0u;
      if (!misc[1]) atomicAdd(&a.sink.counters->overflow, 1ull);
    }
    __syncthreads();
    if (!in_lds && !misc[1]) return;  // no room: the plane is reported, never written out of bounds
    REC *sorted = in_lds ? lds_items : static_cast<REC *>(a.scratch) + misc[0];
    // dense plane pairs sorted in the arena: the predicate's fields beside each record
    FatRecord *fat = (FAT && !in_lds)
                         ? reinterpret_cast<FatRecord *>(static_cast<char *>(a.scratch) +
                                                         (size_t)tp.arena_cap * kArenaSlotBytes) + misc[0]
                         : nullptr;
    // own records at [0, n_own), the next plane's at [n_own, total)
    // Each thread reads a block of consecutive records: one binary search for the
    // block's first record, then a forward walk over the segment boundaries.
    auto find = [&](uint32_t j) -> uint32_t {
      uint32_t lo = 0, hi = 2u * T;
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid] <= j)
          lo = mid;
        else
          hi = mid;
      }
      return lo;
    };
    auto local_bucket = [&](const REC &r) -> uint32_t {
      double x, y, z;
      binning_point_of(r, x, y, z);
      return yz_bucket(y, z, a.g);
    };
    if (total <= (uint32_t)(kPairRegItems * THREADS)) {
      // records stay in registers between the histogram and the placement pass
      REC reg[kPairRegItems];
      uint32_t lb[kPairRegItems];
      const uint32_t ipt = (total + THREADS - 1u) / THREADS;  // records per thread
      const uint32_t jb = threadIdx.x * ipt;
      const uint32_t je = min(total, jb + ipt);
      uint32_t t = jb < total ? find(jb) : 0u;
#pragma unroll
      for (int k = 0; k < kPairRegItems; ++k) {
        const uint32_t j = jb + k;
        if (j < je) {
          while (seg[t + 1u] <= j) ++t;
          reg[k] = items[src[t] + (j - seg[t])];
        }
      }
#pragma unroll
      for (int k = 0; k < kPairRegItems; ++k) {
        const uint32_t j = jb + k;
        if (j < je) {
          lb[k] = local_bucket(reg[k]);
          atomicAdd(&(j < n_own ? cur_own : cur_next)[lb[k]], 1u);
        }
      }
      __syncthreads();
      block_exclusive_scan(cur_own, 2u * L + 1u, wsum);
#pragma unroll
      for (int k = 0; k < kPairRegItems; ++k) {
        const uint32_t j = jb + k;
        if (j < je) sorted[atomicAdd(&(j < n_own ? cur_own : cur_next)[lb[k]], 1u)] = reg[k];
      }
      __syncthreads();
    } else {
      // pass 1: bucket histogram; pass 2: placement (records read twice)
      constexpr uint32_t kBatch = 4;
#pragma unroll 1
      for (int pass = 0; pass < 2; ++pass) {
#pragma unroll 1
        for (uint32_t jb = threadIdx.x * kBatch; jb < total; jb += kBatch * THREADS) {
          REC r[kBatch];
          uint32_t t = find(jb);
#pragma unroll
          for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t j = jb + u;
            if (j < total) {
              while (seg[t + 1u] <= j) ++t;
              r[u] = items[src[t] + (j - seg[t])];
            }
          }
#pragma unroll
          for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t j = jb + u;
            if (j >= total) continue;
            uint32_t *cur = j < n_own ? cur_own : cur_next;
            const uint32_t b = local_bucket(r[u]);
            if (pass == 0)
              atomicAdd(&cur[b], 1u);
            else
              sorted[atomicAdd(&cur[b], 1u)] = r[u];
          }
        }
        __syncthreads();
        if (pass == 0) block_exclusive_scan(cur_own, 2u * L + 1u, wsum);
      }
    }
    // cur[b] now holds the end of bucket b.  Arena plane pairs: the predicate's fields of
    // every sorted record, written in sorted order (the records are visible to the whole
    // workgroup after the barrier above)
    if (fat) {
      for (uint32_t p = threadIdx.x; p < total; p += THREADS) fat[p] = fat_record_of(a.sink.o, sorted[p].idx);
      __syncthreads();
    }
    if (cp > 1u) {
      if (in_lds)
        scan_plane<FORM, DIGEST, true>(a, lds_items, cur_own, cur_next, n_own, THREADS, scen, chunk, cp, nullptr,
                                       acc);
      else
        scan_plane<FORM, DIGEST, true>(a, sorted, cur_own, cur_next, n_own, THREADS, scen, chunk, cp, fat, acc);
    } else {
      if (in_lds)
        scan_plane<FORM, DIGEST, false>(a, lds_items, cur_own, cur_next, n_own, THREADS, scen, 0u, 1u, nullptr,
                                        acc);
      else
        scan_plane<FORM, DIGEST, false>(a, sorted, cur_own, cur_next, n_own, THREADS, scen, 0u, 1u, fat, acc);
    }
  }
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD b mod 8), each
// XCD with its own L2.  Plane pair w reads plane w+1 again as its halo, so runs of
// kXcdRun consecutive plane pairs are given to one XCD (the runs still alternate over the
// XCDs, which keeps their loads balanced), and most halo reads find the records the
// neighbouring plane pair loaded in that L2.  Applies when G is a multiple of 8 kXcdRun.
constexpr uint32_t kXcds = kXcdCount;
constexpr uint32_t kXcdRun = 8;
__device__ __forceinline__ uint32_t xcd_local_item(uint32_t b, uint32_t g) {
  if (g % (kXcds * kXcdRun)) return b;
  const uint32_t xcd = b % kXcds, slot = b / kXcds;
  return (slot / kXcdRun) * (kXcds * kXcdRun) + xcd * kXcdRun + slot % kXcdRun;
}

// Occupancy target of the pair kernel: 5 waves per SIMD lets the register allocator keep
// the split form at 77 VGPRs (list mode), 6 waves per SIMD, instead of 103 (4 waves) with
// no spills; the FP64 form spills under it below 512 threads and keeps the default there.
template <int FORM, int THREADS>
constexpr int kPairWavesPerEu = (FORM == kFormSplit || THREADS >= 512) ? 5 : 1;

static_assert(sizeof(FatRecord) == kFatRecordBytes, "predicate record size");
static_assert(sizeof(RecordF64) <= kArenaSlotBytes && sizeof(RecordF32) <= kArenaSlotBytes, "arena slot");

// PERSIST false: one workgroup per plane pair (grid nplanes * nscen * nchunk).
// PERSIST true (tp.persistent_groups workgroups, as many as the CUs hold at once): each
// workgroup loops over plane pairs.  Workgroup b (on XCD b mod 8, as workgroups are dealt to
// the XCDs in turn) takes the next run r of its XCD's queue (a.arena_top[kQueueStride *
// (1 + xcd)], cleared by the partition; one atomic per run), which is run r of the runs of 8
// consecutive items that XCD owns in the one-workgroup-per-item mapping (xcd_local_item), and
// scans its items, until the run starts past the last item.  Every wavefront of the
// workgroup reads the same run, so all leave the loop together.  A separate instantiation:
// the loop around the scan costs registers (96 VGPRs against 77) that the one-item kernel
// keeps free.
template <int FORM, bool DIGEST, int THREADS, bool FAT, bool PERSIST>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(kPairWavesPerEu<FORM, THREADS>)))
void k_pair_tiles(PairTileArgs a) {
  extern __shared__ uint32_t lds[];
  ScanAcc acc;
  const uint32_t nchunk = a.tp.nchunk;
  if constexpr (!PERSIST) {
    uint32_t item = xcd_local_item(blockIdx.x, gridDim.x);
    if (a.tp.plane_order) item = a.tp.plane_order[item / nchunk] * nchunk + item % nchunk;
    pair_plane<FORM, DIGEST, THREADS, FAT>(a, lds, item, acc);
  } else {
    // the LDS header's last word, which pair_plane leaves alone (a static __shared__ word
    // would push the kernel past the dynamic LDS limit it is configured with)
    uint32_t &next_pos = lds[a.tp.lds_header_words - 1];
    const uint32_t total = a.tp.nplanes * a.tp.nscen * nchunk;
    const uint32_t xcd = blockIdx.x % kXcds;
    uint32_t *queue = a.arena_top + kQueueStride * (1 + xcd);
    for (;;) {
      if (threadIdx.x == 0) next_pos = atomicAdd(queue, 1u);
      __syncthreads();  // also: the previous item's LDS is no longer read
      const uint32_t run = next_pos;
      __syncthreads();  // every thread has read next_pos before thread 0 writes the next one
      const uint32_t first = (run * kXcds + xcd) * kXcdRun;
      if (first >= total) break;
      for (uint32_t i = 0; i < kXcdRun && first + i < total; ++i) {
        uint32_t item = first + i;
        if (a.tp.plane_order) item = a.tp.plane_order[item / nchunk] * nchunk + item % nchunk;
        pair_plane<FORM, DIGEST, THREADS, FAT>(a, lds, item, acc);
        __syncthreads();  // the next item's set-up rewrites the LDS this one read
      }
    }
  }
  flush_counters<DIGEST>(a.sink, acc);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_bin_keys(int form, const DeviceObjects &o, const GridParams &g, uint32_t *keys, hipStream_t s) {
  if (o.n == 0) return hipSuccess;
  const dim3 grid((unsigned)((o.n + 255) / 256));
  if (form == kFormFp64)
    hipLaunchKernelGGL(k_bin_keys<false>, grid, dim3(256), 0, s, o.rx, o.ry, o.rz, o.n, g, keys);
  else
    hipLaunchKernelGGL(k_bin_keys<true>, grid, dim3(256), 0, s, o.rx, o.ry, o.rz, o.n, g, keys);
  return hipGetLastError();
}

template <typename REC, int ITEMS>
void launch_partition_items(const DeviceObjects &o, const uint32_t *keys, const TileParams &tp, REC *out,
                            uint32_t *tps, uint32_t *arena_top, hipStream_t s, size_t stage_bytes) {
  const dim3 grid(tp.ntiles), block(kPartThreads);
  const size_t head = sizeof(uint32_t) * partition_header_words(tp.nplanes);
  // staging capacity: the whole tile when it fits the LDS (and the caller's window), else the
  // largest multiple of 1024 records that does (two or more passes).  With at most 2^14
  // planes the header leaves room for 3072 FP64 records, so a pass holds >= 1024 records.
  size_t cap = (size_t)ITEMS * kPartThreads;
  if (stage_bytes && sizeof(REC) * cap > stage_bytes)
    cap = std::max<size_t>(1024, stage_bytes / sizeof(REC) / 1024 * 1024);
  if (head + sizeof(REC) * cap > kPairTileLdsBytes) cap = (kPairTileLdsBytes - head) / sizeof(REC) / 1024 * 1024;
  hipLaunchKernelGGL((k_tile_partition<REC, ITEMS>), grid, block, head + sizeof(REC) * cap, s, o.rx, o.ry, o.rz, o.n,
                     keys, tp, out, tps, arena_top, (uint32_t)cap);
}

template <typename REC>
hipError_t launch_tile_partition_as(const DeviceObjects &o, const uint32_t *keys, const TileParams &tp, void *items,
                                    uint32_t *tps, uint32_t *arena_top, hipStream_t s, size_t stage_bytes) {
  REC *out = static_cast<REC *>(items);
  switch (tp.tile_size) {
    case 4096:
      launch_partition_items<REC, 4>(o, keys, tp, out, tps, arena_top, s, stage_bytes);
      break;
    case 8192:
      launch_partition_items<REC, 8>(o, keys, tp, out, tps, arena_top, s, stage_bytes);
      break;
    case 16384:
      launch_partition_items<REC, 16>(o, keys, tp, out, tps, arena_top, s, stage_bytes);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_tile_partition(int form, const DeviceObjects &o, const uint32_t *keys, const TileParams &tp,
                                 void *items, uint32_t *tps, uint32_t *arena_top, hipStream_t s, size_t stage_bytes) {
  if (o.n == 0) return hipSuccess;
  if (form == kFormFp64)
    return launch_tile_partition_as<RecordF64>(o, keys, tp, items, tps, arena_top, s, stage_bytes);
  return launch_tile_partition_as<RecordF32>(o, keys, tp, items, tps, arena_top, s, stage_bytes);
}

// The four (form, output) instantiations of one workgroup size, arena layout and mode.
template <int THREADS, bool FAT, bool PERSIST>
hipError_t set_pair_lds_limit_of(int max_lds) {
  const void *fs[] = {reinterpret_cast<const void *>(&k_pair_tiles<kFormFp64, false, THREADS, FAT, PERSIST>),
                      reinterpret_cast<const void *>(&k_pair_tiles<kFormFp64, true, THREADS, FAT, PERSIST>),
                      reinterpret_cast<const void *>(&k_pair_tiles<kFormSplit, false, THREADS, FAT, PERSIST>),
                      reinterpret_cast<const void *>(&k_pair_tiles<kFormSplit, true, THREADS, FAT, PERSIST>)};
  for (const void *f : fs) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <int THREADS>
hipError_t set_pair_lds_limit(int max_lds) {
  hipError_t e = set_pair_lds_limit_of<THREADS, false, false>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit_of<THREADS, true, false>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit_of<THREADS, false, true>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit_of<THREADS, true, true>(max_lds);
  return e;
}

hipError_t configure_tile_kernels() {
  const int max_lds = 160 * 1024;
  hipError_t e = set_pair_lds_limit<128>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit<256>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit<384>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit<512>(max_lds);
  if (e == hipSuccess) e = set_pair_lds_limit<1024>(max_lds);
  auto set = [&](const void *f) {
    if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
  };
  set(reinterpret_cast<const void *>(&k_tile_partition<RecordF64, 4>));
  set(reinterpret_cast<const void *>(&k_tile_partition<RecordF64, 8>));
  set(reinterpret_cast<const void *>(&k_tile_partition<RecordF64, 16>));
  set(reinterpret_cast<const void *>(&k_tile_partition<RecordF32, 4>));
  set(reinterpret_cast<const void *>(&k_tile_partition<RecordF32, 8>));
  set(reinterpret_cast<const void *>(&k_tile_partition<RecordF32, 16>));
  return e;
}

template <int THREADS, bool PERSIST>
int pair_groups_per_cu_as(int form) {
  int n = 0;
  const void *f = form == kFormFp64
                      ? reinterpret_cast<const void *>(&k_pair_tiles<kFormFp64, false, THREADS, false, PERSIST>)
                      : reinterpret_cast<const void *>(&k_pair_tiles<kFormSplit, false, THREADS, false, PERSIST>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, THREADS, 0) != hipSuccess) return 0;
  return n;
}

template <bool PERSIST>
int pair_groups_per_cu_mode(int form, uint32_t threads) {
  switch (threads) {
    case 128:
      return pair_groups_per_cu_as<128, PERSIST>(form);
    case 256:
      return pair_groups_per_cu_as<256, PERSIST>(form);
    case 384:
      return pair_groups_per_cu_as<384, PERSIST>(form);
    case 512:
      return pair_groups_per_cu_as<512, PERSIST>(form);
    case 1024:
      return pair_groups_per_cu_as<1024, PERSIST>(form);
    default:
      return 0;
  }
}

int pair_tiles_groups_per_cu(int form, uint32_t threads, bool persistent) {
  return persistent ? pair_groups_per_cu_mode<true>(form, threads) : pair_groups_per_cu_mode<false>(form, threads);
}

uint32_t pair_tiles_header_words(const TileParams &tp) {
  const uint32_t L = 1u << tp.plane_bits;
  const uint32_t T = tp.nscen > 1 ? tp.max_scen_tiles : tp.ntiles;
  // segment tables, bucket cursors, wavefront sums, 4 scratch words, and a last word that
  // only the persistent kernel uses (its queue position)
  uint32_t w = 4 * T + 2 + 2 * L + 2 + kMaxPairThreads / 64 + 1 + 4 + 1;
  return (w + 7u) & ~7u;  // 32-byte alignment of the record array
}

template <int THREADS, bool FAT, bool PERSIST>
void launch_pair_variant(bool f64, bool digest, dim3 grid, dim3 block, size_t shm, hipStream_t s,
                         const PairTileArgs &a) {
  if (f64 && !digest) hipLaunchKernelGGL((k_pair_tiles<kFormFp64, false, THREADS, FAT, PERSIST>), grid, block, shm, s, a);
  if (f64 && digest) hipLaunchKernelGGL((k_pair_tiles<kFormFp64, true, THREADS, FAT, PERSIST>), grid, block, shm, s, a);
  if (!f64 && !digest) hipLaunchKernelGGL((k_pair_tiles<kFormSplit, false, THREADS, FAT, PERSIST>), grid, block, shm, s, a);
  if (!f64 && digest) hipLaunchKernelGGL((k_pair_tiles<kFormSplit, true, THREADS, FAT, PERSIST>), grid, block, shm, s, a);
}

template <int THREADS>
void launch_pair_tiles_as(int form, bool digest, const PairTileArgs &a, hipStream_t s) {
  const TileParams &tp = a.tp;
  const size_t rec = (form == kFormFp64) ? sizeof(RecordF64) : sizeof(RecordF32);
  const size_t shm = sizeof(uint32_t) * tp.lds_header_words + rec * tp.cap_items;
  const dim3 grid(tp.persistent_groups ? tp.persistent_groups : tp.nplanes * tp.nscen * tp.nchunk), block(THREADS);
  const bool f64 = (form == kFormFp64), persist = tp.persistent_groups != 0u;
  if (tp.arena_fat && persist) launch_pair_variant<THREADS, true, true>(f64, digest, grid, block, shm, s, a);
  if (tp.arena_fat && !persist) launch_pair_variant<THREADS, true, false>(f64, digest, grid, block, shm, s, a);
  if (!tp.arena_fat && persist) launch_pair_variant<THREADS, false, true>(f64, digest, grid, block, shm, s, a);
  if (!tp.arena_fat && !persist) launch_pair_variant<THREADS, false, false>(f64, digest, grid, block, shm, s, a);
}

hipError_t launch_pair_tiles(int form, bool digest, const DeviceObjects &o, const GridParams &g,
                             const TileParams &tp, const FunctorParams &fp, const void *items, const uint32_t *tps,
                             void *scratch, uint32_t *arena_top, LoggedPair *log, uint64_t log_cap,
                             unsigned long long *log_count, ScanCounters *counters, uint64_t iteration,
                             const StepState *step_state, hipStream_t s) {
  if (o.n == 0) return hipSuccess;
  PairTileArgs a;
  a.sink = PairSink{o, fp, log, log_cap, log_count, counters, iteration, step_state};
  a.g = g;
  a.tp = tp;
  a.items = items;
  a.tps = tps;
  a.scratch = scratch;
  a.arena_top = arena_top;
  switch (tp.pair_threads) {
    case 128:
      launch_pair_tiles_as<128>(form, digest, a, s);
      break;
    case 256:
      launch_pair_tiles_as<256>(form, digest, a, s);
      break;
    case 512:
      launch_pair_tiles_as<512>(form, digest, a, s);
      break;
    case 384:
      launch_pair_tiles_as<384>(form, digest, a, s);
      break;
    case 1024:
      launch_pair_tiles_as<1024>(form, digest, a, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ladds

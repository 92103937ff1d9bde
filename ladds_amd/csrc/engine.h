// engine.h -- internal types shared by the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ladds {

// Precision forms of the propagation (values of ladds_config.precision).
enum : int { kFormFp64 = 0, kFormSplit = 1, kFormSplitAll32 = 2 };

// Uniform environment of one Yoshida sub-step (identical for every object):
// Earth rotation angle and Sun/Moon positions, precomputed on the host in FP64.
// The third-body terms use Battin's form a = -GM/|s|^3 (1+q)^-3/2 (r + f(q) s)
// (Battin 1987, eq. 8.61) which has no cancellation, so FP32 evaluation keeps
// the relative accuracy of the tidal term itself.
template <typename T>
struct SubstepEnv {
  T ct, st;               // cos/sin of the Earth rotation angle
  T sun[3];               // Sun position [km]
  T sun_inv_s2;           // 1 / |s|^2
  T sun_k;                // GM_sun / |s|^3
  T moon[3];              // Moon position [km]
  T moon_inv_s2;          // 1 / |m|^2
  T moon_k;               // GM_moon / |m|^3
  T sun_acc[3];           // GM_sun s / |s|^3   (indirect term of the direct third-body form)
  T moon_acc[3];          // GM_moon m / |m|^3
  T pad[2];
};
static_assert(sizeof(SubstepEnv<double>) == 20 * 8, "env layout");
static_assert(sizeof(SubstepEnv<float>) == 20 * 4, "env layout");

// Folded spatial grid: cell (ix, iy, iz) -> bucket ((ix & mx) * Ky + (iy & my)) * Kz + (iz & mz).
// Neighbouring cells stay adjacent in bucket order (z fastest), so buckets hold
// spatially coherent runs; cells that alias into one bucket are at least
// K * cell_size apart and are rejected by the distance test.
struct GridParams {
  double inv_cell;       // 1 / cell size in y and z [1/km]
  double inv_cell_x;     // 1 / cell size in x (the plane width) [1/km], >= cell size
  double inv_cell_y;     // plane tiles: 1 / bucket row height in y [1/km], the cell size or twice it
  double inv_cell_z;     // plane tiles: 1 / bucket width in z [1/km], the cell size or twice it
  uint32_t bx, by, bz;   // log2 of Kx, Ky, Kz
  uint32_t nbuckets;     // Kx * Ky * Kz
};

// Plane-tile decomposition of the folded grid (tiles.hip): planes are the 2^bx
// x-slices of 2^plane_bits buckets; objects are partitioned in tiles of tile_size.
struct TileParams {
  uint32_t tile_size;        // objects per partition tile
  uint32_t ntiles;
  uint32_t plane_bits;       // by + bz
  uint32_t nplanes;          // 2^bx
  uint32_t cap_items;        // records of one plane pair that fit the pair kernel's LDS
  uint32_t lds_header_words; // LDS words before the record array in the pair kernel
  uint32_t pair_threads;     // pair-kernel workgroup size: 128, 256, 512 or 1024
  // Batched ensembles: independent scenarios, each with its own tiles and planes; objects
  // of different scenarios never pair.  nscen == 1: one population, tiles of tile_size.
  uint32_t nscen;
  uint32_t max_scen_tiles;     // most tiles of one scenario (sizes the pair kernel's LDS)
  const uint32_t *tile_begin;  // nscen > 1: first object of every tile (ntiles + 1 entries)
  const uint32_t *scen_tile;   // nscen > 1: first tile of every scenario (nscen + 1 entries)
  // Dense planes: a plane of n_own objects is scanned by min(nchunk, ceil(n_own / 512))
  // workgroups, chunk c taking the objects with index = c (mod chunks); every chunk sorts
  // the plane itself (grid: nchunk workgroups per plane, the unused ones exit).
  uint32_t nchunk;
  uint32_t arena_cap;          // records of the global scratch arena
  // 1: the arena holds arena_cap slots of kArenaSlotBytes (sorted records) followed by
  // arena_cap predicate records (kFatRecordBytes each) at the same positions, and the pair
  // kernel's arena planes read the predicate's fields from them (set when plane pairs
  // exceeded the LDS capacity at set-up: dense, clustered populations)
  uint32_t arena_fat;
  // The plane pairs (scenario * nplanes + plane) in dispatch order (nscen * nplanes
  // entries), or null for index order.
  const uint32_t *plane_order;
  // 1: bucket rows are two cells tall (GridParams::inv_cell_y), and an object scans two rows
  // of the next plane, chosen by the half of its row it lies in; 0: rows one cell tall,
  // three rows of the next plane.
  uint32_t half_rows;
  // 1: buckets are two cells wide along z (GridParams::inv_cell_z), and an object scans two
  // buckets of a row, chosen by the half of its bucket it lies in; 0: three buckets.
  uint32_t half_buckets;
  // > 0: the pair kernel runs this many workgroups (as many as the CUs hold at once), each
  // taking plane pairs from its XCD's queue until none is left; 0: one workgroup per plane
  // pair.
  uint32_t persistent_groups;
};

// Words of the pair scan's device counters: [0] the scratch arena's fill and, one 128-byte
// line each, [kQueueStride * (1 + x)] the next run of XCD x's queue (persistent pair
// kernel).  The partition kernel clears them.
constexpr uint32_t kXcdCount = 8;
constexpr uint32_t kQueueStride = 32;
constexpr uint32_t kScanCounterWords = kQueueStride * (1 + kXcdCount);

// Sorted record of the plain FP64 form and of the split form: the position the
// object is binned by, and its index in the object arrays.
struct alignas(32) RecordF64 {
  double x, y, z;
  uint32_t idx;
  uint32_t pad;
};
struct alignas(16) RecordF32 {
  float x, y, z;
  uint32_t idx;
};

// Device-side step position for graph replay: a captured step reads which environment
// and iteration it is at from here, and its last kernel advances `step`.
struct StepState {
  uint32_t step;      // index of the step in the staged environment segment
  uint32_t pad;
  uint64_t iter_base; // iteration label = iter_base + step (modulo 2^64)
};

// Device view of the object arrays.
struct DeviceObjects {
  size_t n;
  double *rx, *ry, *rz, *vx, *vy, *vz;
  double *radius_m, *aom, *bc_inv, *cr;
  double *srp64, *drag64;   // kSrpCoef * cr * aom, kDragCoef * bc_inv
  float *srp32, *drag32;
  int32_t *activity;
  uint64_t *id, *parent;
};

struct PairRecord {  // same layout as ladds_pair
  uint64_t id_a, id_b;
  uint32_t idx_a, idx_b;
  double d2;
  double mid[3];
};

struct PairRecord32 {  // same layout as ladds_pair_record
  uint32_t id_a, id_b;
  float d2;
};

struct LoggedPair {  // same layout as ladds_logged_pair
  PairRecord pair;
  uint64_t iteration;
  uint32_t evaded;
  uint32_t scenario;
};

// Counters written by the pair scan and the propagation (device memory).
struct ScanCounters {
  unsigned long long hits, evaded, candidates;
  unsigned long long overflow;  // planes skipped because the scratch arena was exhausted
  unsigned long long digest[2], edigest[2];
  unsigned long long burnups;   // objects the propagation flagged as burnt up
  unsigned long long last_hit;  // 1 + the latest iteration that had a collision (0: none)
};

// Spawn protection (Simulation.cpp:287-293, 399-402, 438, 470-473): the simulation loop
// counts iterations since the last one whose detection found collisions; when the count
// is 100 at the start of an iteration, every parent id is reset, so fragments of one
// breakup can conjunct again.  Device copy of the count, used while objects with a parent
// id exist.
struct SpawnState {
  unsigned long long quiet;    // the count at the start of the first iteration not yet accounted
  unsigned long long seen;     // ScanCounters::hits when the last detection was accounted
  unsigned long long cleared;  // parent ids were reset
  unsigned long long pad;
};

struct FunctorParams {
  double cutoff2;         // cutoff^2 [km^2]
  double cdf_km;          // collisionDistanceFactor / 1000
  double evasion2;        // evasionTrackingCutoff^2 [km^2]
  double min_det_radius;  // [m]
  double threshold2;      // absolute mode threshold^2
  double dt;              // interpolation window [s]
  int threshold_mode;
  float cutoff2_f32;      // inflated FP32 prefilter radius^2 (split form)
};

}  // namespace ladds

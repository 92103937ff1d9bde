// launch.h -- host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace ladds {

hipError_t upload_atmosphere();

hipError_t launch_propagate(int form, const DeviceObjects &o, const SubstepEnv<double> *env64,
                            const SubstepEnv<float> *env32, int nsteps, unsigned mask, const double cdt[4],
                            const double ddt[3], const GridParams *grid, uint32_t *keys, uint32_t *count,
                            uint32_t *slot, double burnup_r2, const StepState *step_state, hipStream_t stream,
                            const DeviceObjects *out = nullptr, uint32_t step_offset = 0,
                            unsigned long long *burnups = nullptr);
// Graph replay bookkeeping: set the step position, advance it by one step.
hipError_t launch_set_step(StepState *st, uint32_t step, uint64_t iter_base, hipStream_t stream);
hipError_t launch_advance_step(StepState *st, hipStream_t stream);
// Copies `words` 8-byte words from device memory to mapped host memory (one workgroup).
hipError_t launch_publish(const unsigned long long *src, unsigned long long *dst, int words, hipStream_t stream);

hipError_t launch_bin_count(int form, const DeviceObjects &o, const GridParams &g, uint32_t *count, uint32_t *slot,
                            hipStream_t s);
hipError_t launch_bin_reduce(const GridParams &g, const uint32_t *count, uint32_t *tile_sum, hipStream_t s);
hipError_t launch_bin_scan(const GridParams &g, uint32_t *count, const uint32_t *tile_sum, uint32_t *start,
                           hipStream_t s);
hipError_t launch_bin_scatter(int form, const DeviceObjects &o, const GridParams &g, const uint32_t *slot,
                              const uint32_t *start, void *rec, hipStream_t s);
hipError_t launch_pair_scan(int form, bool digest, const DeviceObjects &o, const GridParams &g,
                            const FunctorParams &fp, const void *rec, const uint32_t *start,
                            LoggedPair *log, uint64_t log_cap, unsigned long long *log_count,
                            ScanCounters *counters, uint64_t iteration, hipStream_t s);

// Plane-tile pipeline (tiles.hip)
hipError_t launch_bin_keys(int form, const DeviceObjects &o, const GridParams &g, uint32_t *keys, hipStream_t s);
// stage_bytes: most LDS bytes of records a partition workgroup stages per pass (0: the whole
// tile when it fits); a smaller window lets the partition share CUs with other kernels.
hipError_t launch_tile_partition(int form, const DeviceObjects &o, const uint32_t *keys, const TileParams &tp,
                                 void *items, uint32_t *tps, uint32_t *arena_top, hipStream_t s,
                                 size_t stage_bytes = 0);
hipError_t launch_pair_tiles(int form, bool digest, const DeviceObjects &o, const GridParams &g,
                             const TileParams &tp, const FunctorParams &fp, const void *items, const uint32_t *tps,
                             void *scratch, uint32_t *arena_top, LoggedPair *log, uint64_t log_cap,
                             unsigned long long *log_count, ScanCounters *counters, uint64_t iteration,
                             const StepState *step_state, hipStream_t s);
uint32_t pair_tiles_header_words(const TileParams &tp);
// Pair-kernel workgroups of `threads` threads a CU holds by its registers alone (list
// mode; the persistent instantiation when `persistent`; 0 when the query fails).
int pair_tiles_groups_per_cu(int form, uint32_t threads, bool persistent = false);
// Workgroup sizes of the pair kernel.
inline bool valid_pair_threads(int t) { return t == 128 || t == 256 || t == 384 || t == 512 || t == 1024; }
hipError_t configure_tile_kernels();
constexpr uint32_t kPairTileLdsBytes = 160 * 1024;
// LDS reserved per pair-kernel workgroup beyond its request (allocation granularity).
constexpr uint32_t kLdsMargin = 1024;
constexpr uint32_t kPartitionTile = 8192;
// Global scratch arena of the pair scan: per record slot a sorted record (32 bytes
// reserved, the FP64 record's size) and the fields the exact predicate reads (80 bytes).
constexpr size_t kArenaSlotBytes = 32;
constexpr size_t kFatRecordBytes = 80;

// Search radius of the pair scan: the cutoff in the FP64 form; in the split form the
// FP32 prefilter radius, inflated so that FP32 rounding of positions (< 2^-24 |x| per
// axis, < 0.014 km for |x| < 1e5 km) and of the FP32 distance cannot reject a pair
// the exact FP64 test keeps.
inline double search_radius(int form, double cutoff_km) {
  return form == kFormFp64 ? cutoff_km : cutoff_km * (1.0 + 1e-5) + 0.1;
}

// Burn-up compaction: alive flags are activity >= 0.
hipError_t launch_count_alive(const DeviceObjects &o, uint32_t *tile_sum, uint32_t ntiles, hipStream_t s);
hipError_t launch_compact(const DeviceObjects &src, const DeviceObjects &dst, const uint32_t *tile_sum,
                          uint32_t ntiles, uint32_t *n_alive, hipStream_t s);

constexpr uint32_t kCompactTile = 2048;

// Spawn protection.  launch_clear_parents with st == nullptr resets every parent id; with a
// SpawnState it does so only when the quiet-iteration count reaches 100 within the next
// `span` iteration starts.  launch_spawn_post then advances the count over those
// iterations (resetting it when `detect` and the scan found collisions).
hipError_t launch_clear_parents(SpawnState *st, uint64_t span, uint64_t *parent, size_t n, hipStream_t s);
hipError_t launch_spawn_post(SpawnState *st, uint64_t span, const ScanCounters *counters, bool detect,
                             hipStream_t s);

// Output records (export.hip): float32 AoS state + uint32 ids; one iteration's pairs as
// {uint32, uint32, float32}.  counts[2] must be zero before launch_log_records.
// State of the objects at the given indices, AoS {rx, ry, rz, vx, vy, vz} per entry.
hipError_t launch_gather_state(const DeviceObjects &o, const uint32_t *idx, size_t count, double *out, hipStream_t s);
hipError_t launch_particle_records(const DeviceObjects &o, size_t n, float *pos, float *vel, uint32_t *ids,
                                   hipStream_t s);
hipError_t launch_log_records(const LoggedPair *log, const unsigned long long *log_count, uint64_t log_cap,
                              uint64_t iteration, PairRecord32 *hits, uint64_t hcap, PairRecord32 *ev, uint64_t ecap,
                              unsigned long long *counts, hipStream_t s);

}  // namespace ladds

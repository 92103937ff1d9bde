// launch.h -- host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace ladds {

hipError_t upload_atmosphere();

hipError_t launch_propagate(int form, const DeviceObjects &o, const SubstepEnv<double> *env64,
                            const SubstepEnv<float> *env32, int nsteps, unsigned mask, const double cdt[4],
                            const double ddt[3], const GridParams *grid, uint32_t *count, uint32_t *slot,
                            double burnup_r2, hipStream_t stream);

hipError_t launch_bin_count(const DeviceObjects &o, const GridParams &g, uint32_t *count, uint32_t *slot,
                            hipStream_t s);
hipError_t launch_bin_reduce(const GridParams &g, const uint32_t *count, uint32_t *tile_sum, hipStream_t s);
hipError_t launch_bin_scan(const GridParams &g, uint32_t *count, const uint32_t *tile_sum, uint32_t *start,
                           hipStream_t s);
hipError_t launch_bin_scatter(int form, const DeviceObjects &o, const GridParams &g, const uint32_t *slot,
                              const uint32_t *start, void *rec, uint32_t *sidx, hipStream_t s);
hipError_t launch_pair_scan(int form, bool digest, const DeviceObjects &o, const GridParams &g,
                            const FunctorParams &fp, const void *rec, const uint32_t *sidx, const uint32_t *start,
                            LoggedPair *log, uint64_t log_cap, unsigned long long *log_count,
                            ScanCounters *counters, uint64_t iteration, hipStream_t s);

// Burn-up compaction: alive flags are activity >= 0.
hipError_t launch_count_alive(const DeviceObjects &o, uint32_t *tile_sum, uint32_t ntiles, hipStream_t s);
hipError_t launch_compact(const DeviceObjects &src, const DeviceObjects &dst, const uint32_t *tile_sum,
                          uint32_t ntiles, uint32_t *n_alive, hipStream_t s);

constexpr uint32_t kCompactTile = 2048;

}  // namespace ladds

// noalloc.cpp -- the conjunction digest on caller-owned device arrays, without allocation.
//
// ladds_conjunction_digest_noalloc is ladds_conjunction_digest (the count and 128-bit
// digest of Simulation::collisionDetection's pair sets, Simulation.cpp:211-222) for callers
// that cannot allocate device memory: code being captured into a HIP graph, serving loops,
// harnesses that forbid hipMalloc.  The objects stay in the caller's device memory; the
// scan's scratch (cell keys, plane-sorted records, tile offsets, the arena for dense plane
// pairs, counters) is the file-scope device array `scratch`, of a fixed size that holds
// LADDS_NOALLOC_MAX_OBJECTS objects.  The grid shape is fixed from the object count alone (no host
// pass over the positions): 1024 x-planes, buckets per plane from the count, records of a
// plane pair beyond the LDS capacity sorted in the arena.  Every call on one GPU shares that
// GPU's array: calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/ladds_amd.h"
#include "engine.h"
#include "launch.h"

namespace ladds {

namespace {

constexpr size_t kMaxObjects = LADDS_NOALLOC_MAX_OBJECTS;
constexpr uint32_t kPlaneBitsX = 10;  // 1024 x-planes
constexpr uint32_t kTile = 4096;      // partition tile
constexpr uint32_t kMaxTiles = (uint32_t)((kMaxObjects + kTile - 1) / kTile);
constexpr uint32_t kThreads = 256;    // pair-kernel workgroup

// Layout of the scratch array (each part 256-byte aligned).
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kKeysOff = 0;
constexpr size_t kRecOff = kKeysOff + align256(sizeof(uint32_t) * kMaxObjects);
constexpr size_t kTpsOff = kRecOff + align256(sizeof(RecordF64) * kMaxObjects);
constexpr size_t kArenaOff = kTpsOff + align256(sizeof(uint32_t) * (size_t)kMaxTiles * ((1u << kPlaneBitsX) + 1));
constexpr size_t kArenaRecords = 2 * kMaxObjects;  // every record at most twice (own plane and halo)
constexpr size_t kTopOff = kArenaOff + align256(sizeof(RecordF64) * kArenaRecords);
constexpr size_t kCountersOff = kTopOff + align256(sizeof(uint32_t) * kScanCounterWords);  // the pair scan's counter words
constexpr size_t kScratchBytes = kCountersOff + align256(sizeof(ScanCounters));

static_assert(sizeof(RecordF64) >= sizeof(RecordF32), "records");

}  // namespace

// The scratch buffer.  Host code never uses this name as an address: hipGetSymbolAddress
// gives each device's copy of it, kept in g_scratch[device].
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

__global__ void k_clear_counters(ScanCounters *c) {
  unsigned long long *w = reinterpret_cast<unsigned long long *>(c);
  for (uint32_t k = threadIdx.x; k < sizeof(ScanCounters) / sizeof(unsigned long long); k += blockDim.x) w[k] = 0ull;
}

constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char *g_scratch[kMaxDevices] = {};
bool g_configured[kMaxDevices] = {};

thread_local char g_error[256] = "";

int fail(int code, const char *msg) {
  std::snprintf(g_error, sizeof(g_error), "%s", msg);
  return code;
}

int hipfail(hipError_t e, const char *what) {
  std::snprintf(g_error, sizeof(g_error), "%s: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? LADDS_E_NOMEM : LADDS_E_HIP;
}

// This device's scratch address, looked up once per device with the device current.
hipError_t scratch_of(int device, unsigned char **out) {
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<unsigned char *>(p);
  }
  if (!g_configured[device]) {
    const hipError_t e = configure_tile_kernels();  // LDS limits of the pair kernel (no allocation)
    if (e != hipSuccess) return e;
    g_configured[device] = true;
  }
  *out = g_scratch[device];
  return hipSuccess;
}

uint32_t ceil_log2(uint64_t v) {
  uint32_t b = 0;
  while ((1ull << b) < v) ++b;
  return b;
}

}  // namespace

}  // namespace ladds

using namespace ladds;

extern "C" int ladds_conjunction_digest_noalloc(const ladds_device_objects *objs, const ladds_config *cfg,
                                                double dt_interp, uint64_t *n, uint64_t digest[2], uint64_t *ne,
                                                uint64_t edigest[2], void *stream) {
  if (!objs || !cfg || !n || !digest || !ne || !edigest) return fail(LADDS_E_INVALID, "null argument");
  *n = *ne = 0;
  digest[0] = digest[1] = edigest[0] = edigest[1] = 0;
  if (objs->n > kMaxObjects)
    return fail(LADDS_E_INVALID, "ladds_conjunction_digest_noalloc: more objects than LADDS_NOALLOC_MAX_OBJECTS");
  if (!(cfg->cutoff_km > 0.0) || cfg->precision < 0 || cfg->precision > 2 || cfg->threshold_mode < 0 ||
      cfg->threshold_mode > 1 || !(dt_interp >= 0.0))
    return fail(LADDS_E_INVALID, "invalid cutoff, precision, threshold mode or dt");
  if (cfg->device < 0 || cfg->device >= kMaxDevices) return fail(LADDS_E_NODEVICE, "device ordinal out of range");
  if (objs->n < 2) return LADDS_OK;
  if (!objs->rx || !objs->ry || !objs->rz || !objs->vx || !objs->vy || !objs->vz || !objs->radius_m ||
      !objs->activity || !objs->id || !objs->parent)
    return fail(LADDS_E_INVALID, "every object array is required (device pointers)");
  hipError_t e = hipSetDevice(cfg->device);
  if (e != hipSuccess) return hipfail(e, "hipSetDevice");
  unsigned char *base = nullptr;
  e = scratch_of(cfg->device, &base);
  if (e != hipSuccess) return hipfail(e, "hipGetSymbolAddress(scratch)");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int form = cfg->precision;
  const size_t N = objs->n;

  DeviceObjects o{};
  o.n = N;
  o.rx = const_cast<double *>(objs->rx);
  o.ry = const_cast<double *>(objs->ry);
  o.rz = const_cast<double *>(objs->rz);
  o.vx = const_cast<double *>(objs->vx);
  o.vy = const_cast<double *>(objs->vy);
  o.vz = const_cast<double *>(objs->vz);
  o.radius_m = const_cast<double *>(objs->radius_m);
  o.activity = const_cast<int32_t *>(objs->activity);
  o.id = const_cast<uint64_t *>(objs->id);
  o.parent = const_cast<uint64_t *>(objs->parent);

  // grid: cells of side >= the search radius, 2^10 x-planes, ~1 object per in-plane bucket
  const double radius = search_radius(form, cfg->cutoff_km);
  GridParams g{};
  g.inv_cell = 1.0 / (radius * (1.0 + 1e-9));
  g.inv_cell_x = g.inv_cell;
  g.inv_cell_y = g.inv_cell;  // buckets of one cell (tp.half_rows = tp.half_buckets = 0)
  g.inv_cell_z = g.inv_cell;
  g.bx = kPlaneBitsX;
  const uint32_t pb = std::min<uint32_t>(12u, std::max<uint32_t>(4u, ceil_log2(N / (1u << kPlaneBitsX) + 1)));
  g.by = pb / 2;
  g.bz = pb - g.by;
  g.nbuckets = 1u << (g.bx + pb);

  TileParams tp{};
  tp.tile_size = kTile;
  tp.ntiles = (uint32_t)((N + kTile - 1) / kTile);
  tp.plane_bits = pb;
  tp.nplanes = 1u << kPlaneBitsX;
  tp.nscen = 1;
  tp.pair_threads = kThreads;
  tp.nchunk = 1;
  tp.lds_header_words = pair_tiles_header_words(tp);
  const size_t rec = form == kFormFp64 ? sizeof(RecordF64) : sizeof(RecordF32);
  // LDS for two workgroups per CU
  const size_t hdr = 4 * (size_t)tp.lds_header_words;
  tp.cap_items = (uint32_t)((kPairTileLdsBytes / 2 - hdr - kLdsMargin) / rec / 64 * 64);
  tp.arena_cap = (uint32_t)(sizeof(RecordF64) * kArenaRecords / rec);

  FunctorParams fp{};
  fp.cutoff2 = cfg->cutoff_km * cfg->cutoff_km;
  fp.cdf_km = cfg->collision_distance_factor / 1000.;
  fp.evasion2 = cfg->evasion_cutoff_km * cfg->evasion_cutoff_km;
  fp.min_det_radius = cfg->min_detection_radius_m;
  fp.threshold2 = cfg->threshold_km * cfg->threshold_km;
  fp.dt = dt_interp;
  fp.threshold_mode = cfg->threshold_mode;
  const double inflated = search_radius(kFormSplit, cfg->cutoff_km);
  fp.cutoff2_f32 = std::nextafter((float)(inflated * inflated), INFINITY);

  uint32_t *keys = reinterpret_cast<uint32_t *>(base + kKeysOff);
  void *records = base + kRecOff;
  uint32_t *tps = reinterpret_cast<uint32_t *>(base + kTpsOff);
  void *arena = base + kArenaOff;
  uint32_t *arena_top = reinterpret_cast<uint32_t *>(base + kTopOff);
  ScanCounters *counters = reinterpret_cast<ScanCounters *>(base + kCountersOff);

  // counters start at zero (a fill kernel: nothing here is a memset node if captured); the
  // partition kernel resets the arena top itself
  hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, counters);
  e = hipGetLastError();
  if (e == hipSuccess) e = launch_bin_keys(form, o, g, keys, s);
  if (e == hipSuccess) e = launch_tile_partition(form, o, keys, tp, records, tps, arena_top, s);
  if (e == hipSuccess)
    e = launch_pair_tiles(form, true, o, g, tp, fp, records, tps, arena, arena_top, nullptr, 0, nullptr, counters, 0,
                          nullptr, s);
  if (e != hipSuccess) return hipfail(e, "conjunction scan launch");
  ScanCounters host{};
  e = hipMemcpyAsync(&host, counters, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hipfail(e, "counter read-back");
  if (host.overflow) return fail(LADDS_E_OVERFLOW, "conjunction scan arena exhausted");
  *n = host.hits;
  *ne = host.evaded;
  digest[0] = host.digest[0];
  digest[1] = host.digest[1];
  edigest[0] = host.edigest[0];
  edigest[1] = host.edigest[1];
  return LADDS_OK;
}

namespace ladds {
// Message of the last failed ladds_conjunction_digest_noalloc call on this thread (read by
// ladds_last_error(NULL)); empty when there is none.
const char *noalloc_error() { return g_error; }
}  // namespace ladds

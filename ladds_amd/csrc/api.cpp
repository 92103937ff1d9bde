// api.cpp -- the C ABI (include/ladds_amd.h): context, memory, and the per-step schedule.
//
// A context replaces the objects Simulation::run builds (src/ladds/Simulation.cpp:522-549):
// the AutoPas container (initAutoPas :59-120), the accumulator + Yoshida integrator
// (initIntegrator :122-148) and the CollisionFunctor built per detection
// (collisionDetection :211-222).  All state lives in HBM; host calls only enqueue
// kernels on the context's stream, except where results are returned.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <array>
#include <unordered_set>
#include <vector>

#include "../../include/ladds_amd.h"
#include "constants.h"
#include "engine.h"
#include "ephemeris.h"
#include "launch.h"

using namespace ladds;

namespace ladds {
const char *noalloc_error();
}

namespace {

struct EventPair {
  hipEvent_t a, b;
};

struct KernelProfile {
  std::vector<EventPair> used;
  uint64_t launches = 0;
  double ms = 0.0;  // resolved time
};

constexpr const char *kKernelNames[] = {"propagate",   "bin_count", "bin_reduce",     "bin_scan",
                                        "bin_scatter", "pair_scan", "bin_keys",       "tile_partition",
                                        "pair_tiles",  "compact"};

struct HostReadback {
  ScanCounters sc;
  unsigned long long log_count;
  SpawnState spawn;
};
// The scan counters and the pair-log count share one device block laid out as the start of
// HostReadback, so one copy reads both back.
struct DeviceReadback {
  ScanCounters sc;
  unsigned long long log_count;
};
static_assert(offsetof(HostReadback, log_count) == offsetof(DeviceReadback, log_count) &&
                  sizeof(DeviceReadback) == offsetof(HostReadback, spawn),
              "one copy reads the counters and the log count");

// Pair-scan pipelines (ladds_config.scan_variant).
enum : int { kVariantAuto = 0, kVariantAtomicSort = 1, kVariantPlaneTiles = 2 };

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count);
}

uint32_t ceil_log2(uint64_t v) {
  uint32_t b = 0;
  while ((1ull << b) < v) ++b;
  return b;
}

// One buffer of the overlapped simulation loop: a step's state (r, v, activity), its cell
// keys, its plane-sorted records and tile offsets and the pair kernel's scratch arena.
struct PipeBuffer {
  double *r[3] = {nullptr, nullptr, nullptr}, *v[3] = {nullptr, nullptr, nullptr};
  int32_t *act = nullptr;
  uint32_t *keys = nullptr;
  size_t cap = 0;  // objects the state, keys and records hold
  void *rec = nullptr;
  uint32_t *tps = nullptr;
  size_t tps_alloc = 0;
  uint32_t *arena_top = nullptr;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
};

}  // namespace

struct ladds_ctx {
  ladds_config cfg{};
  int form = kFormFp64;
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0;
  int cus = 0;  // compute units of the device

  size_t n = 0, cap = 0;
  DeviceObjects d{};
  DeviceObjects d_alt{};  // compaction target

  GridParams grid{};
  TileParams tiles{};
  int variant = kVariantPlaneTiles;
  uint32_t *count = nullptr, *slot = nullptr, *start = nullptr, *tile_sum = nullptr;
  void *rec = nullptr;
  uint32_t *tps = nullptr, *arena_top = nullptr;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  size_t tps_alloc = 0;
  uint32_t *n_alive = nullptr;
  uint32_t grid_buckets_alloc = 0;
  size_t tile_sum_alloc = 0;

  // Environments of one segment of up to kEnvSegment steps, evaluated on the device by
  // k_substep_env in stream order before the steps that read them.
  SubstepEnv<double> *env64 = nullptr;
  SubstepEnv<float> *env32 = nullptr;

  LoggedPair *log = nullptr;
  uint64_t log_cap = 0;
  unsigned long long *log_count = nullptr;
  ScanCounters *counters = nullptr;
  // Pinned host copies of the counters, the log count and the spawn state: one
  // asynchronous read-back and one synchronisation at the end of ladds_simulate.
  HostReadback *h_rb = nullptr;         // pinned, coherent: kernels store into it (d_rb_host)
  HostReadback *d_rb_host = nullptr;     // device address of h_rb
  // Pinned staging for pair-log entries returned by ladds_conjunctions, and for the
  // object states gathered by ladds_gather_state.
  LoggedPair *h_log = nullptr;
  size_t h_log_cap = 0;
  void *h_x = nullptr;
  size_t h_x_bytes = 0;
  // The device pair log may hold entries of ladds_simulate not yet cleared.
  bool log_nonempty = false;

  // Spawn protection (Simulation.cpp:287-293): `quiet` is the count of iterations since the
  // last collision at the start of the next iteration to run.  While objects with a parent
  // id exist (`protected_objects`), ladds_simulate keeps the count on the device
  // (`d_spawn`) and resets the parent ids there; otherwise the count is kept on the host
  // from the latest iteration with collisions.
  // `slot` holds the plane-tile bucket keys of the current state and grid (written by the
  // last ladds_propagate step), so a detection can skip k_bin_keys.
  bool keys_valid = false;

  uint64_t quiet = 0;  // from 0 at ladds_create, as Simulation.h:227; continued across calls
  bool protected_objects = false;
  SpawnState *d_spawn = nullptr;

  // Ensemble scenarios (ladds_set_scenarios): offsets of S scenarios, S + 1 entries; empty
  // for a single population.  Device tile tables of the plane-tile pipeline.
  std::vector<uint64_t> scen_off;
  uint32_t *d_tile_begin = nullptr, *d_scen_tile = nullptr;
  size_t tile_begin_alloc = 0, scen_tile_alloc = 0;
  uint32_t *d_plane_order = nullptr;
  size_t plane_order_alloc = 0;

  // Autotuned plane-tile shape (ladds_autotune); zero / negative values keep the defaults.
  double tune_cell_km = 0.0;
  int tune_pb = -1;
  int tune_threads = 0;
  int tune_tile = 0;
  int tune_chunks = 0;
  double tune_fx = 0.0;  // plane width / cell size
  int tune_wgs = 0;      // pair-kernel workgroups per CU its LDS is sized for (0: automatic)
  int tune_order = 0;    // pair-kernel dispatch: 0 plane order, 1 heaviest runs of planes first
  int tune_ens_plane = 0;  // ensembles: objects per x-plane the planes are sized for (0: 64)
  int tune_rows = 0;       // bucket row height along y in cells: 2 or 1 (0: 1)
  int tune_bcells = 0;     // bucket width along z in cells: 2 or 1 (0: 1)
  int tune_persistent = 0; // 1: pair kernel as persistent workgroups with per-XCD queues
  int tune_pair_streams = 0; // pair scans of the overlapped loop on 1 or 2 streams (0: by form)

  StepState *d_step = nullptr;  // graph replay position (ladds_config.step_graph)

  // Overlapped simulation.  The eager loop cycles through three buffers (four when the pair
  // scans alternate over `stream2` and `stream4`): step k+2's propagation (`stream`) runs
  // beside step k+1's partition (`stream3`) and step k's pair scan.  The step graphs (ladds_config.step_graph) alternate buffers 0 and 1,
  // the partition on `stream`.  Buffer 0 is the context's own arrays (`d`, `slot`, `rec`,
  // `tps`, `scratch`, `arena_top`); pipe[1] .. pipe[3] hold the others.
  hipStream_t stream2 = nullptr, stream3 = nullptr;
  hipStream_t stream4 = nullptr;   // second pair-scan stream (pair_streams 2)
  hipEvent_t ev_join4 = nullptr;   // end of ladds_simulate: join of the two pair-scan streams
  hipEvent_t ev_prop[4] = {}, ev_part[4] = {}, ev_pair[4] = {};
  hipEvent_t ev_join = nullptr;  // end of ladds_simulate: stream waits for stream2
  PipeBuffer pipe[4];

  // Staging for the output records (ladds_particle_records, ladds_log_records).
  void *xbuf = nullptr;
  size_t xbuf_bytes = 0;
  unsigned long long *xcount = nullptr;

  uint64_t step_count = 0;  // steps since t0
  // Environment segment staged by ladds_propagate: steps [env_first, env_first + env_len)
  // of the clock (env_t0, env_dt) are in env64/env32 from index 0 (reused by later calls of
  // ladds_propagate and ladds_simulate while they cover the steps; see env_window).
  double env_t0 = 0.0, env_dt = 0.0;
  uint64_t env_first = 0, env_len = 0;
  // ladds_propagate returned with its steps still running on `stream`; entry points that
  // touch device memory from the host first wait for them (settle)
  bool pending = false;
  double c[4], dcoef[3];

  bool profiling = false;
  std::map<std::string, KernelProfile> prof;
  std::vector<EventPair> event_pool;

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
  int hipfail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return (e == hipErrorOutOfMemory) ? LADDS_E_NOMEM : LADDS_E_HIP;
  }
  double time() const { return cfg.t0 + (double)step_count * cfg.dt; }

  EventPair get_events() {
    if (!event_pool.empty()) {
      EventPair p = event_pool.back();
      event_pool.pop_back();
      return p;
    }
    EventPair p{};
    hipEventCreate(&p.a);
    hipEventCreate(&p.b);
    return p;
  }
};

#define HIPCHK(ctx, call)                                 \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return (ctx)->hipfail(e_, #call); \
  } while (0)

namespace {

// Launch wrapper: records events around the launch when profiling is on.
template <typename F>
hipError_t timed_on(ladds_ctx *ctx, const char *name, hipStream_t stream, F &&launch) {
  if (!ctx->profiling) return launch();
  EventPair ev = ctx->get_events();
  (void)hipEventRecord(ev.a, stream);
  const hipError_t e = launch();
  (void)hipEventRecord(ev.b, stream);
  KernelProfile &kp = ctx->prof[name];
  kp.used.push_back(ev);
  kp.launches += 1;
  return e;
}

template <typename F>
hipError_t timed(ladds_ctx *ctx, const char *name, F &&launch) {
  return timed_on(ctx, name, ctx->stream, launch);
}

void resolve_profiles(ladds_ctx *ctx) {
  hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) hipStreamSynchronize(ctx->stream2);
  if (ctx->stream3) hipStreamSynchronize(ctx->stream3);
  for (auto &kv : ctx->prof) {
    for (EventPair &ev : kv.second.used) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess) kv.second.ms += ms;
      ctx->event_pool.push_back(ev);
    }
    kv.second.used.clear();
  }
}

void free_objects(DeviceObjects &o) {
  hipFree(o.rx);
  hipFree(o.ry);
  hipFree(o.rz);
  hipFree(o.vx);
  hipFree(o.vy);
  hipFree(o.vz);
  hipFree(o.radius_m);
  hipFree(o.aom);
  hipFree(o.bc_inv);
  hipFree(o.cr);
  hipFree(o.srp64);
  hipFree(o.drag64);
  hipFree(o.srp32);
  hipFree(o.drag32);
  hipFree(o.activity);
  hipFree(o.id);
  hipFree(o.parent);
  o = DeviceObjects{};
}

hipError_t alloc_objects(DeviceObjects &o, size_t cap) {
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  chk(dalloc(&o.rx, cap));
  chk(dalloc(&o.ry, cap));
  chk(dalloc(&o.rz, cap));
  chk(dalloc(&o.vx, cap));
  chk(dalloc(&o.vy, cap));
  chk(dalloc(&o.vz, cap));
  chk(dalloc(&o.radius_m, cap));
  chk(dalloc(&o.aom, cap));
  chk(dalloc(&o.bc_inv, cap));
  chk(dalloc(&o.cr, cap));
  chk(dalloc(&o.srp64, cap));
  chk(dalloc(&o.drag64, cap));
  chk(dalloc(&o.srp32, cap));
  chk(dalloc(&o.drag32, cap));
  chk(dalloc(&o.activity, cap));
  chk(dalloc(&o.id, cap));
  chk(dalloc(&o.parent, cap));
  return e;
}

// Grid geometry.  Atomic counting sort: cells of side >= 2R, 2x2x2 search blocks, about
// one object per bucket.  Plane tiles: cells of side >= R, 3x3x3 neighbourhoods; enough
// x-planes for the x-extent of the population (no aliasing of x-slabs), about one object
// per bucket within a plane, and an LDS capacity sized from the largest plane pair.
int setup_grid(ladds_ctx *ctx, size_t n, const double *rx) {
  ctx->keys_valid = false;
  const bool tiles = ctx->variant == kVariantPlaneTiles;
  const double radius = search_radius(ctx->form, ctx->cfg.cutoff_km);
  const double min_cell = (tiles ? 1.0 : 2.0) * radius * (1.0 + 1e-9);
  double cell = ctx->cfg.cell_size_km > 0.0 ? ctx->cfg.cell_size_km : min_cell;
  if (tiles && ctx->tune_cell_km > 0.0) cell = ctx->tune_cell_km;
  if (!(cell >= min_cell))
    return ctx->fail(LADDS_E_INVALID, tiles ? "cell_size_km must be >= the search radius (cutoff, or the split "
                                              "prefilter radius)"
                                            : "cell_size_km must be >= 2 * the search radius");
  GridParams g{};
  g.inv_cell = 1.0 / cell;
  // plane width: cells may be longer in x than in y and z (fewer, fuller planes)
  const double cell_x = (tiles && ctx->tune_fx > 1.0) ? cell * ctx->tune_fx : cell;
  g.inv_cell_x = 1.0 / cell_x;
  // buckets two cells tall (wide): an object then scans two rows of the next plane (two
  // buckets of a row), not three
  const bool half_rows = tiles && ctx->tune_rows == 2;
  const bool half_buckets = tiles && ctx->tune_bcells == 2;
  g.inv_cell_y = half_rows ? 0.5 * g.inv_cell : g.inv_cell;
  g.inv_cell_z = half_buckets ? 0.5 * g.inv_cell : g.inv_cell;
  uint32_t bits;
  std::vector<uint32_t> plane_count;
  if (tiles) {
    double xmax = 1.0e4;
    if (rx && n) {
      xmax = 0.0;
      for (size_t i = 0; i < n; ++i) xmax = std::max(xmax, std::fabs(rx[i]));
    }
    const double nx = std::min(1.0e6, 2.0 * xmax / cell_x + 3.0);
    g.bx = std::min<uint32_t>(14u, std::max<uint32_t>(2u, ceil_log2((uint64_t)nx)));
    // a scenario of an ensemble is a population of its own: its planes hold ~64 objects (or
    // the tuner's choice), folding distant x-slabs into one plane so that the per-plane
    // set-up of the pair kernel is spread over enough objects
    const size_t nscen = ctx->scen_off.size() > 2 ? ctx->scen_off.size() - 1 : 1;
    const size_t per_pop = std::max<size_t>(n / nscen, 1);
    const size_t ens_plane = ctx->tune_ens_plane > 0 ? (size_t)ctx->tune_ens_plane : 64;
    if (nscen > 1) g.bx = std::min<uint32_t>(g.bx, std::max<uint32_t>(2u, ceil_log2(per_pop / ens_plane + 1)));
    const uint64_t used = std::min<uint64_t>((uint64_t)nx, 1ull << g.bx);
    uint32_t pb = std::min<uint32_t>(12u, std::max<uint32_t>(4u, ceil_log2(per_pop / used + 1)));
    if (ctx->tune_pb >= 4 && ctx->tune_pb <= 12) pb = (uint32_t)ctx->tune_pb;
    // rows twice as tall as the buckets are wide fold y with one bit fewer, so that both
    // folds span similar lengths
    g.by = (half_rows && !half_buckets) ? (pb - 1) / 2 : pb / 2;
    g.bz = pb - g.by;
    bits = g.bx + pb;
    if (rx && n) {
      // per scenario plane populations (one population: a single scenario)
      const uint32_t P = 1u << g.bx, mx = P - 1u;
      plane_count.assign((size_t)P * nscen, 0u);
      for (size_t sc = 0; sc < nscen; ++sc) {
        const size_t lo = nscen > 1 ? ctx->scen_off[sc] : 0, hi = nscen > 1 ? ctx->scen_off[sc + 1] : n;
        for (size_t i = lo; i < hi; ++i) {
          double u = rx[i] * g.inv_cell_x;
          u = std::fmin(std::fmax(u, -1.0e9), 1.0e9);
          plane_count[sc * P + ((uint32_t)(int32_t)std::floor(u) & mx)] += 1;
        }
      }
    }
  } else {
    bits = std::min<uint32_t>(28u, std::max<uint32_t>(12u, ceil_log2(std::max<size_t>(n, 1))));
    g.bx = (bits + 2) / 3;
    g.by = (bits + 1) / 3;
    g.bz = bits / 3;
  }
  g.nbuckets = 1u << bits;
  ctx->grid = g;
  if (!tiles && g.nbuckets > ctx->grid_buckets_alloc) {
    hipFree(ctx->count);
    hipFree(ctx->start);
    HIPCHK(ctx, dalloc(&ctx->count, g.nbuckets));
    HIPCHK(ctx, dalloc(&ctx->start, (size_t)g.nbuckets + 1));
    HIPCHK(ctx, hipMemset(ctx->count, 0, sizeof(uint32_t) * g.nbuckets));
    ctx->grid_buckets_alloc = g.nbuckets;
  }
  // burn-up compaction reuses tile_sum: one entry per 2048 objects
  const size_t ts_need = std::max<size_t>((size_t)g.nbuckets / 2048, n / 2048) + 1;
  if (ts_need > ctx->tile_sum_alloc) {
    hipFree(ctx->tile_sum);
    HIPCHK(ctx, dalloc(&ctx->tile_sum, ts_need));
    ctx->tile_sum_alloc = ts_need;
  }
  TileParams tp{};
  tp.tile_size = (ctx->tune_tile == 4096 || ctx->tune_tile == 16384) ? (uint32_t)ctx->tune_tile : kPartitionTile;
  tp.ntiles = (uint32_t)std::max<size_t>(1, (n + tp.tile_size - 1) / tp.tile_size);
  tp.nscen = 1;
  if (tiles && ctx->scen_off.size() > 2) {
    // tiles never straddle scenarios: scenario s owns tiles [scen_tile[s], scen_tile[s+1])
    const size_t S = ctx->scen_off.size() - 1;
    std::vector<uint32_t> tb, st(S + 1);
    uint32_t most = 1;
    for (size_t sc = 0; sc < S; ++sc) {
      st[sc] = (uint32_t)tb.size();
      for (uint64_t o = ctx->scen_off[sc]; o < ctx->scen_off[sc + 1]; o += tp.tile_size) tb.push_back((uint32_t)o);
      most = std::max<uint32_t>(most, (uint32_t)tb.size() - st[sc]);
    }
    st[S] = (uint32_t)tb.size();
    tb.push_back((uint32_t)n);
    if (tb.size() > ctx->tile_begin_alloc) {
      hipFree(ctx->d_tile_begin);
      HIPCHK(ctx, dalloc(&ctx->d_tile_begin, tb.size()));
      ctx->tile_begin_alloc = tb.size();
    }
    if (st.size() > ctx->scen_tile_alloc) {
      hipFree(ctx->d_scen_tile);
      HIPCHK(ctx, dalloc(&ctx->d_scen_tile, st.size()));
      ctx->scen_tile_alloc = st.size();
    }
    HIPCHK(ctx, hipMemcpy(ctx->d_tile_begin, tb.data(), sizeof(uint32_t) * tb.size(), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_scen_tile, st.data(), sizeof(uint32_t) * st.size(), hipMemcpyHostToDevice));
    tp.ntiles = (uint32_t)std::max<size_t>(1, tb.size() - 1);
    tp.nscen = (uint32_t)S;
    tp.max_scen_tiles = most;
    tp.tile_begin = ctx->d_tile_begin;
    tp.scen_tile = ctx->d_scen_tile;
  }
  tp.plane_bits = g.by + g.bz;
  tp.nplanes = 1u << g.bx;
  tp.half_rows = half_rows ? 1u : 0u;
  tp.half_buckets = half_buckets ? 1u : 0u;
  if (tiles && !plane_count.empty() && plane_count.size() == (size_t)tp.nplanes * tp.nscen &&
      ctx->tune_order == 1) {
    // Dispatch order of the pair kernel's plane pairs (scenario x plane): workgroups start
    // in index order and each XCD takes runs of 8 consecutive plane pairs, so runs of 8
    // planes of one scenario are put in descending order of their work (longest first: the
    // last workgroups on a CU are the short ones).  Work of a plane pair ~ 6 x own records
    // (scan) + next records (gather).  Measured at 1M objects: -6% for FP64 records with
    // 512 threads, +7% for split records with 256 threads, so the tuner decides.
    const uint32_t P = tp.nplanes, run = 8;
    std::vector<std::pair<uint64_t, uint32_t>> w;
    for (uint32_t sc = 0; sc < tp.nscen; ++sc) {
      const uint32_t *pc = plane_count.data() + (size_t)sc * P;
      for (uint32_t r0 = 0; r0 < P; r0 += run) {
        uint64_t s = 0;
        for (uint32_t p = r0; p < r0 + run && p < P; ++p) s += 6ull * pc[p] + pc[(p + 1) % P];
        w.emplace_back(s, sc * P + r0);
      }
    }
    std::stable_sort(w.begin(), w.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    std::vector<uint32_t> order;
    order.reserve((size_t)P * tp.nscen);
    for (const auto &e : w)
      for (uint32_t k = 0; k < run && (e.second % P) + k < P; ++k) order.push_back(e.second + k);
    if (order.size() > ctx->plane_order_alloc) {
      hipFree(ctx->d_plane_order);
      ctx->d_plane_order = nullptr;
      ctx->plane_order_alloc = 0;
      HIPCHK(ctx, dalloc(&ctx->d_plane_order, order.size()));
      ctx->plane_order_alloc = order.size();
    }
    HIPCHK(ctx, hipMemcpy(ctx->d_plane_order, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice));
    tp.plane_order = ctx->d_plane_order;
  }
  // ladds_config.block_size overrides; by default FP64 records (32 B) leave room for one
  // workgroup per CU, which then needs 16 waves, while the split form fits several.
  const int bs = ctx->cfg.block_size;
  tp.pair_threads = valid_pair_threads(bs)
                        ? (uint32_t)bs
                        : (tp.nscen > 1 ? 128u : (ctx->form == kFormFp64 ? 1024u : 256u));
  if (valid_pair_threads(ctx->tune_threads)) tp.pair_threads = (uint32_t)ctx->tune_threads;
  tp.lds_header_words = pair_tiles_header_words(tp);  // depends on the workgroup size
  const size_t rec = (ctx->form == kFormFp64) ? sizeof(RecordF64) : sizeof(RecordF32);
  size_t pair_max = 0;
  if (!plane_count.empty()) {
    const size_t P = tp.nplanes;
    for (size_t sc = 0; sc < plane_count.size() / P; ++sc)
      for (size_t p = 0; p < P; ++p)
        pair_max = std::max<size_t>(pair_max, plane_count[sc * P + p] + plane_count[sc * P + (p + 1) % P]);
  }
  // dense planes (clusters): several workgroups per plane
  size_t own_max = 0;
  for (uint32_t c : plane_count) own_max = std::max<size_t>(own_max, c);
  tp.nchunk = (uint32_t)std::min<size_t>(16, std::max<size_t>(1, own_max / (4 * (size_t)tp.pair_threads)));
  if (ctx->cfg.scan_chunks >= 1 && ctx->cfg.scan_chunks <= 32) tp.nchunk = (uint32_t)ctx->cfg.scan_chunks;
  if (ctx->tune_chunks >= 1 && ctx->tune_chunks <= 32) tp.nchunk = (uint32_t)ctx->tune_chunks;
  // LDS of a plane pair: room for the fullest plane pair now, then grown to the whole LDS
  // share of as many workgroups per CU as that size allows (slack for the planes' drift at
  // no cost in occupancy, which bounds the latency-bound scan).  Plane pairs beyond the
  // capacity are sorted in the global scratch arena.
  {
    const size_t hdr = 4 * (size_t)tp.lds_header_words;
    // the header (segment tables of every tile, bucket cursors) grows with the tile count:
    // it must leave room for a workgroup's worth of records
    if (tiles && hdr + 256 * rec + kLdsMargin > kPairTileLdsBytes)
      return ctx->fail(LADDS_E_INVALID, "too many partition tiles or buckets per plane for the pair kernel's LDS: "
                                        "use fewer objects per context or a larger partition tile");
    const size_t fit = (kPairTileLdsBytes - hdr) / rec;
    const size_t need = std::min(fit, std::max<size_t>(1024, (pair_max + 63) / 64 * 64));
    // workgroups per CU the kernel's registers allow (16 waves per CU when unknown)
    const int by_regs = pair_tiles_groups_per_cu(ctx->form, tp.pair_threads);
    const size_t wave_cap = by_regs > 0 ? (size_t)by_regs
                                        : std::max<size_t>(1, 16 / std::max<uint32_t>(1, tp.pair_threads / 64));
    size_t wgs = std::min(wave_cap, kPairTileLdsBytes / (hdr + need * rec + kLdsMargin));
    if (ctx->tune_wgs >= 1 && ctx->tune_wgs <= 8) wgs = (size_t)ctx->tune_wgs;
    wgs = std::max<size_t>(1, wgs);
    // persistent pair kernel: as many workgroups as the CUs hold at once (its own
    // instantiation's registers and the LDS share; a multiple of the 8 XCDs, each with its
    // share of the queues)
    tp.persistent_groups = 0u;
    if (tiles && ctx->tune_persistent == 1 && ctx->cus > 0) {
      const int p_regs = pair_tiles_groups_per_cu(ctx->form, tp.pair_threads, true);
      const size_t per_cu = p_regs > 0 ? std::min<size_t>(wgs, (size_t)p_regs) : wgs;
      tp.persistent_groups = (uint32_t)(((size_t)ctx->cus * per_cu + 7) / 8 * 8);
    }
    const size_t share = kPairTileLdsBytes / wgs;
    const size_t cap = share > hdr + kLdsMargin ? (share - hdr - kLdsMargin) / rec / 64 * 64 : 0;
    tp.cap_items = (uint32_t)std::min(fit, std::max<size_t>(cap, 256));
  }
  // scratch arena for planes beyond the LDS capacity: every chunk of a plane sorts it.  When
  // plane pairs exceed the LDS capacity now (dense, clustered populations), each slot also
  // holds the predicate's fields of its object: their many candidates then read two
  // contiguous records instead of gathering from the object arrays (2.3x per step on
  // 8M clustered objects); sparse populations keep the leaner kernel.
  tp.arena_fat = pair_max > tp.cap_items ? 1u : 0u;
  const size_t slot_bytes = kArenaSlotBytes + (tp.arena_fat ? kFatRecordBytes : 0);
  const size_t arena_need = std::max<size_t>(2 * n, 1024) * tp.nchunk;
  const size_t arena_bytes = arena_need * slot_bytes;
  if (tiles && arena_bytes > ctx->scratch_bytes) {
    hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    HIPCHK(ctx, hipMalloc(&ctx->scratch, arena_bytes));
    ctx->scratch_bytes = arena_bytes;
  }
  tp.arena_cap = (uint32_t)std::min<size_t>(ctx->scratch_bytes / slot_bytes, 0xFFFFFFFFu);
  ctx->tiles = tp;
  const size_t tps_need = (size_t)tp.ntiles * (tp.nplanes + 1);
  if (tps_need > ctx->tps_alloc) {
    hipFree(ctx->tps);
    HIPCHK(ctx, dalloc(&ctx->tps, tps_need));
    ctx->tps_alloc = tps_need;
  }
  return LADDS_OK;
}

FunctorParams functor_params(const ladds_ctx *ctx, double dt_interp) {
  const ladds_config &c = ctx->cfg;
  FunctorParams fp{};
  fp.cutoff2 = c.cutoff_km * c.cutoff_km;
  fp.cdf_km = c.collision_distance_factor / 1000.;
  fp.evasion2 = c.evasion_cutoff_km * c.evasion_cutoff_km;
  fp.min_det_radius = c.min_detection_radius_m;
  fp.threshold2 = c.threshold_km * c.threshold_km;
  fp.dt = dt_interp;
  fp.threshold_mode = c.threshold_mode;
  const double inflated = search_radius(kFormSplit, c.cutoff_km);
  fp.cutoff2_f32 = std::nextafter((float)(inflated * inflated), INFINITY);
  return fp;
}

constexpr int kEnvSegment = 8192;  // steps per environment segment
constexpr int kChunk = 1024;       // steps per propagation launch of ladds_propagate

int alloc_env(ladds_ctx *ctx) {
  if (ctx->env64) return LADDS_OK;
  const size_t need = 3 * (size_t)kEnvSegment;
  HIPCHK(ctx, dalloc(&ctx->env64, need));
  HIPCHK(ctx, dalloc(&ctx->env32, need));
  return LADDS_OK;
}

// Environments of steps [first, first + count) of the segment starting at the context's
// current step, evaluated on the context's stream (ordered before the propagation that
// reads them and after the one that read the previous segment's).
int stage_env(ladds_ctx *ctx, int first, int count, hipStream_t stream = nullptr) {
  const double tau[3] = {ctx->c[0], ctx->c[0] + ctx->c[1], ctx->c[0] + ctx->c[1] + ctx->c[2]};
  HIPCHK(ctx, launch_substep_env(ctx->cfg.t0, ctx->cfg.dt, ctx->step_count, tau, first, count, ctx->env64,
                                 ctx->env32, stream ? stream : ctx->stream));
  return LADDS_OK;
}

// Offset in env64/env32 of the context's current step.  The segment staged last serves while
// it covers the next `count` steps of the same clock (environments depend only on t0, dt and
// the step index); otherwise a segment of kEnvSegment steps is staged from the current step,
// on `stream` (default: the context's), after the propagations that read the previous one.
int env_window(ladds_ctx *ctx, int count, size_t *off, hipStream_t stream = nullptr) {
  const bool cached = ctx->env_len && ctx->env_t0 == ctx->cfg.t0 && ctx->env_dt == ctx->cfg.dt &&
                      ctx->step_count >= ctx->env_first &&
                      ctx->step_count + (uint64_t)count <= ctx->env_first + ctx->env_len;
  if (!cached) {
    ctx->env_len = 0;
    const int rc = stage_env(ctx, 0, kEnvSegment, stream);
    if (rc != LADDS_OK) return rc;
    ctx->env_t0 = ctx->cfg.t0;
    ctx->env_dt = ctx->cfg.dt;
    ctx->env_first = ctx->step_count;
    ctx->env_len = (uint64_t)kEnvSegment;
  }
  *off = (size_t)(ctx->step_count - ctx->env_first);
  return LADDS_OK;
}

// Waits for steps ladds_propagate left running (see ladds_ctx::pending).
int settle(ladds_ctx *ctx) {
  if (!ctx->pending) return LADDS_OK;
  ctx->pending = false;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return LADDS_OK;
}
#define LADDS_SETTLE(ctx)                  \
  do {                                     \
    const int settle_rc = settle(ctx);     \
    if (settle_rc != LADDS_OK) return settle_rc; \
  } while (0)

int ensure_log(ladds_ctx *ctx, uint64_t cap) {
  if (cap <= ctx->log_cap) return LADDS_OK;
  LoggedPair *nl = nullptr;
  HIPCHK(ctx, dalloc(&nl, cap));
  if (ctx->log) {
    HIPCHK(ctx, hipMemcpyAsync(nl, ctx->log, sizeof(LoggedPair) * ctx->log_cap, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(ctx->log);
  }
  ctx->log = nl;
  ctx->log_cap = cap;
  return LADDS_OK;
}

int launch_propagate_steps(ladds_ctx *ctx, size_t env_offset, int nsteps, bool bin, double burnup_r2,
                           const StepState *step_state = nullptr) {
  double cdt[4], ddt[3];
  for (int k = 0; k < 4; ++k) cdt[k] = ctx->c[k] * ctx->cfg.dt;
  for (int k = 0; k < 3; ++k) ddt[k] = ctx->dcoef[k] * ctx->cfg.dt;
  const hipError_t e = timed(ctx, "propagate", [&] {
    return launch_propagate(ctx->form, ctx->d, ctx->env64 + 3 * env_offset, ctx->env32 + 3 * env_offset, nsteps,
                            ctx->cfg.components, cdt, ddt, &ctx->grid,
                            (bin && ctx->variant == kVariantPlaneTiles) ? ctx->slot : nullptr,
                            (bin && ctx->variant == kVariantAtomicSort) ? ctx->count : nullptr, ctx->slot, burnup_r2,
                            step_state, ctx->stream, nullptr, 0, &ctx->counters->burnups);
  });
  if (e != hipSuccess) return ctx->hipfail(e, "propagate launch");
  return LADDS_OK;
}

// Spawn-protection bookkeeping around the pair scan of a detection that closes `span`
// iterations (0: none, no objects with a parent id).
hipError_t spawn_pre(ladds_ctx *ctx, uint64_t span, hipStream_t stream) {
  return span ? launch_clear_parents(ctx->d_spawn, span, ctx->d.parent, ctx->n, stream) : hipSuccess;
}
hipError_t spawn_post(ladds_ctx *ctx, uint64_t span, bool detect, hipStream_t stream) {
  return span ? launch_spawn_post(ctx->d_spawn, span, ctx->counters, detect, stream) : hipSuccess;
}

// Binning (after the propagation's binning pass when `binned`) + pair scan on the current state.
int launch_detection(ladds_ctx *ctx, bool binned, double dt_interp, bool digest, uint64_t iteration,
                     const StepState *step_state = nullptr, uint64_t spawn_span = 0) {
  const FunctorParams fp = functor_params(ctx, dt_interp);
  hipError_t e = spawn_pre(ctx, spawn_span, ctx->stream);
  if (ctx->variant == kVariantPlaneTiles) {
    if (!binned)
      e = timed(ctx, "bin_keys", [&] { return launch_bin_keys(ctx->form, ctx->d, ctx->grid, ctx->slot, ctx->stream); });
    if (e == hipSuccess)
      e = timed(ctx, "tile_partition", [&] {
        return launch_tile_partition(ctx->form, ctx->d, ctx->slot, ctx->tiles, ctx->rec, ctx->tps, ctx->arena_top,
                                     ctx->stream);
      });
    if (e == hipSuccess)
      e = timed(ctx, "pair_tiles", [&] {
        return launch_pair_tiles(ctx->form, digest, ctx->d, ctx->grid, ctx->tiles, fp, ctx->rec, ctx->tps,
                                 ctx->scratch, ctx->arena_top, ctx->log, ctx->log_cap, ctx->log_count, ctx->counters,
                                 iteration, step_state, ctx->stream);
      });
    if (e == hipSuccess) e = spawn_post(ctx, spawn_span, true, ctx->stream);
    if (e != hipSuccess) return ctx->hipfail(e, "detection launch");
    return LADDS_OK;
  }
  if (!binned)
    e = timed(ctx, "bin_count",
              [&] { return launch_bin_count(ctx->form, ctx->d, ctx->grid, ctx->count, ctx->slot, ctx->stream); });
  if (e == hipSuccess)
    e = timed(ctx, "bin_reduce", [&] { return launch_bin_reduce(ctx->grid, ctx->count, ctx->tile_sum, ctx->stream); });
  if (e == hipSuccess)
    e = timed(ctx, "bin_scan",
              [&] { return launch_bin_scan(ctx->grid, ctx->count, ctx->tile_sum, ctx->start, ctx->stream); });
  if (e == hipSuccess)
    e = timed(ctx, "bin_scatter", [&] {
      return launch_bin_scatter(ctx->form, ctx->d, ctx->grid, ctx->slot, ctx->start, ctx->rec, ctx->stream);
    });
  if (e == hipSuccess)
    e = timed(ctx, "pair_scan", [&] {
      return launch_pair_scan(ctx->form, digest, ctx->d, ctx->grid, fp, ctx->rec, ctx->start, ctx->log,
                              ctx->log_cap, ctx->log_count, ctx->counters, iteration, ctx->stream);
    });
  if (e == hipSuccess) e = spawn_post(ctx, spawn_span, true, ctx->stream);
  if (e != hipSuccess) return ctx->hipfail(e, "detection launch");
  return LADDS_OK;
}

// Pipeline buffers 1 .. nbuf-1 of the overlapped loop, sized like buffer 0.
// Stream `to` follows what is queued on `from` so far (event `ev`, recorded there).
int follow(ladds_ctx *ctx, hipEvent_t ev, hipStream_t from, hipStream_t to) {
  HIPCHK(ctx, hipEventRecord(ev, from));
  HIPCHK(ctx, hipStreamWaitEvent(to, ev, 0));
  return LADDS_OK;
}

int ensure_overlap_buffers(ladds_ctx *ctx, int nbuf) {
  for (int b = 1; b < nbuf; ++b) {
    PipeBuffer &pb = ctx->pipe[b];
    if (pb.cap < ctx->cap) {
      for (int k = 0; k < 3; ++k) {
        hipFree(pb.r[k]);
        hipFree(pb.v[k]);
        pb.r[k] = pb.v[k] = nullptr;
      }
      hipFree(pb.act);
      hipFree(pb.keys);
      hipFree(pb.rec);
      pb.act = nullptr;
      pb.keys = nullptr;
      pb.rec = nullptr;
      pb.cap = 0;
      for (int k = 0; k < 3; ++k) {
        HIPCHK(ctx, dalloc(&pb.r[k], ctx->cap));
        HIPCHK(ctx, dalloc(&pb.v[k], ctx->cap));
      }
      HIPCHK(ctx, dalloc(&pb.act, ctx->cap));
      HIPCHK(ctx, dalloc(&pb.keys, ctx->cap));
      HIPCHK(ctx, dalloc(reinterpret_cast<RecordF64 **>(&pb.rec), ctx->cap));
      pb.cap = ctx->cap;
    }
    if (!pb.arena_top) HIPCHK(ctx, dalloc(&pb.arena_top, kScanCounterWords));
    if (pb.tps_alloc < ctx->tps_alloc) {
      hipFree(pb.tps);
      pb.tps = nullptr;
      pb.tps_alloc = 0;
      HIPCHK(ctx, dalloc(&pb.tps, ctx->tps_alloc));
      pb.tps_alloc = ctx->tps_alloc;
    }
    if (pb.scratch_bytes < ctx->scratch_bytes) {
      hipFree(pb.scratch);
      pb.scratch = nullptr;
      pb.scratch_bytes = 0;
      HIPCHK(ctx, hipMalloc(&pb.scratch, ctx->scratch_bytes));
      pb.scratch_bytes = ctx->scratch_bytes;
    }
  }
  return LADDS_OK;
}

// View of state buffer b (0: the context's own arrays).
DeviceObjects state_view(const ladds_ctx *ctx, int b) {
  DeviceObjects v = ctx->d;
  if (b) {
    const PipeBuffer &pb = ctx->pipe[b];
    v.rx = pb.r[0];
    v.ry = pb.r[1];
    v.rz = pb.r[2];
    v.vx = pb.v[0];
    v.vy = pb.v[1];
    v.vz = pb.v[2];
    v.activity = pb.act;
  }
  return v;
}

uint32_t *keys_of(ladds_ctx *ctx, int b) { return b ? ctx->pipe[b].keys : ctx->slot; }
void *rec_of(ladds_ctx *ctx, int b) { return b ? ctx->pipe[b].rec : ctx->rec; }
uint32_t *tps_of(ladds_ctx *ctx, int b) { return b ? ctx->pipe[b].tps : ctx->tps; }
void *scratch_of(ladds_ctx *ctx, int b) { return b ? ctx->pipe[b].scratch : ctx->scratch; }
uint32_t *arena_top_of(ladds_ctx *ctx, int b) { return b ? ctx->pipe[b].arena_top : ctx->arena_top; }

// Propagate from state buffer `cur` into `nxt` (on `stream`), writing the cell keys of the
// result when `keys`.  With `step_state` the environments start at its step + step_offset.
int launch_prop_into(ladds_ctx *ctx, int cur, int nxt, size_t env_offset, int nsteps, bool keys, double burnup_r2,
                     const StepState *step_state, uint32_t step_offset, hipStream_t stream = nullptr) {
  const DeviceObjects in = state_view(ctx, cur), out = state_view(ctx, nxt);
  double cdt[4], ddt[3];
  for (int k = 0; k < 4; ++k) cdt[k] = ctx->c[k] * ctx->cfg.dt;
  for (int k = 0; k < 3; ++k) ddt[k] = ctx->dcoef[k] * ctx->cfg.dt;
  if (!stream) stream = ctx->stream;
  const hipError_t e = timed_on(ctx, "propagate", stream, [&] {
    return launch_propagate(ctx->form, in, ctx->env64 + 3 * env_offset, ctx->env32 + 3 * env_offset, nsteps,
                            ctx->cfg.components, cdt, ddt, &ctx->grid, keys ? keys_of(ctx, nxt) : nullptr, nullptr,
                            ctx->slot, burnup_r2, step_state, stream, &out, step_offset,
                            &ctx->counters->burnups);
  });
  if (e != hipSuccess) return ctx->hipfail(e, "propagation launch");
  return LADDS_OK;
}

// Partition state buffer b into its records (on `stream`), staging at most stage_bytes of
// records in LDS per pass (0: the whole tile when it fits).
int launch_part_of(ladds_ctx *ctx, int b, hipStream_t stream, size_t stage_bytes = 0) {
  const hipError_t e = timed_on(ctx, "tile_partition", stream, [&] {
    return launch_tile_partition(ctx->form, state_view(ctx, b), keys_of(ctx, b), ctx->tiles, rec_of(ctx, b),
                                 tps_of(ctx, b), arena_top_of(ctx, b), stream, stage_bytes);
  });
  if (e != hipSuccess) return ctx->hipfail(e, "partition launch");
  return LADDS_OK;
}

// Propagate from state buffer `cur` into 1 - cur and, when `partition`, partition the result
// (both on `stream`).  With `step_state` the environments start at its step + step_offset.
int launch_prop_part(ladds_ctx *ctx, int cur, size_t env_offset, int nsteps, bool partition, double burnup_r2,
                     const StepState *step_state, uint32_t step_offset) {
  const int nxt = 1 - cur;
  int rc = launch_prop_into(ctx, cur, nxt, env_offset, nsteps, partition, burnup_r2, step_state, step_offset);
  if (rc == LADDS_OK && partition) rc = launch_part_of(ctx, nxt, ctx->stream);
  return rc;
}

// Pair scan of state buffer b (records partitioned into buffer b) on `stream`, with the
// spawn-protection bookkeeping of `spawn_span` iterations around it.
int launch_pairs_of(ladds_ctx *ctx, int b, double dt_interp, bool digest, uint64_t iteration,
                    const StepState *step_state, hipStream_t stream, uint64_t spawn_span = 0) {
  const FunctorParams fp = functor_params(ctx, dt_interp);
  hipError_t e = spawn_pre(ctx, spawn_span, stream);
  if (e == hipSuccess)
    e = timed_on(ctx, "pair_tiles", stream, [&] {
      return launch_pair_tiles(ctx->form, digest, state_view(ctx, b), ctx->grid, ctx->tiles, fp, rec_of(ctx, b),
                               tps_of(ctx, b), scratch_of(ctx, b), arena_top_of(ctx, b), ctx->log, ctx->log_cap,
                               ctx->log_count, ctx->counters, iteration, step_state, stream);
    });
  if (e == hipSuccess) e = spawn_post(ctx, spawn_span, true, stream);
  if (e != hipSuccess) return ctx->hipfail(e, "pair scan launch");
  return LADDS_OK;
}

// LDS window of the overlapped loop's partitions.  A workgroup staging a
// whole 8,192-record tile takes 136 KB of LDS and waits for a CU the pair kernel has left;
// staging part of it per pass it runs beside pair workgroups.  Measured at 1M objects with
// one pair-scan stream and three buffers: 1.41e10 -> 1.47e10 object-steps/s with 32 KB (16,
// 48 and 64 KB: 1.45-1.46e10); with two pair-scan streams and four buffers, 64 KB (two
// passes): 70.9 -> 69.7 us per step, 20-step calls 80.4 -> 78.2 (16 KB: slower).  FP64
// records: with one pair-scan stream the passes only cost (1.02e10 -> 0.98e10 with 32 or
// 64 KB against the whole tile), with two streams and four buffers 64 KB gains (93.7 -> 92.3
// us per step, 20-step calls 103.9 -> 102.0).
constexpr size_t kPipelineStageBytes = 64 * 1024;

// One iteration of the eager overlapped loop: propagate buffer `cur` into nxt = cur + 1
// (mod nbuf) on `stream` and, when `detect`, partition nxt on `stream3` once the propagation
// is done; its pair scan (issue_scan) is queued by the next iteration.  Buffer nxt was last
// read by the partition and the pair scan of the step nbuf iterations back
// (`pending_next`): the propagation waits for that pair scan, which itself followed its
// partition.  So step k+2's propagation, step k+1's partition and step k's pair scan can
// all run at once, and the partition (bound by HBM traffic) no longer holds up the next
// propagation.
//
// `fill` (the call's first iteration, when it detects): nothing runs yet, so the
// propagation, partition and pair scan all go on `stream2` in order, with no hop between
// streams on the way to the first pair scan; the next propagation, on `stream`, waits for
// this one only.
int overlapped_iteration(ladds_ctx *ctx, int nbuf, int cur, bool pending_next, size_t env_offset, int nsteps,
                         bool detect, double burnup_r2, double dt_interp, bool digest, uint64_t iteration,
                         uint64_t spawn_span, bool fill) {
  const int nxt = (cur + 1) % nbuf;
  const size_t stage = kPipelineStageBytes;
  if (fill && detect) {
    int rc = launch_prop_into(ctx, cur, nxt, env_offset, nsteps, true, burnup_r2, nullptr, 0, ctx->stream2);
    if (rc == LADDS_OK) rc = follow(ctx, ctx->ev_prop[nxt], ctx->stream2, ctx->stream);
    if (rc == LADDS_OK) rc = launch_part_of(ctx, nxt, ctx->stream2, stage);
    if (rc == LADDS_OK) rc = launch_pairs_of(ctx, nxt, dt_interp, digest, iteration, nullptr, ctx->stream2, spawn_span);
    if (rc == LADDS_OK) HIPCHK(ctx, hipEventRecord(ctx->ev_pair[nxt], ctx->stream2));
    return rc;
  }
  // A hipStreamWaitEvent on a pending event costs 10-23 us of host time (2-3 us for a
  // launch), enough with three a step to make the host nearly as slow as the GPU: the wait
  // for the pair scan that last read buffer nxt is left out once that scan is done.
  if (pending_next) {
    const hipError_t q = hipEventQuery(ctx->ev_pair[nxt]);  // (a pending event's answer is no error)
    if (q == hipErrorNotReady)
      HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_pair[nxt], 0));
    else
      HIPCHK(ctx, q);
  }
  int rc = launch_prop_into(ctx, cur, nxt, env_offset, nsteps, detect, burnup_r2, nullptr, 0);
  if (rc != LADDS_OK || !detect) return rc;
  rc = follow(ctx, ctx->ev_prop[nxt], ctx->stream, ctx->stream3);
  if (rc == LADDS_OK) rc = launch_part_of(ctx, nxt, ctx->stream3, stage);
  if (rc == LADDS_OK) HIPCHK(ctx, hipEventRecord(ctx->ev_part[nxt], ctx->stream3));
  return rc;
}

// The pair scan of a step whose partition overlapped_iteration queued, issued one
// iteration later.
struct PendingScan {
  bool valid = false;
  int b = 0;
  double dt_interp = 0.0;
  uint64_t iteration = 0, span = 0;
};

// Queue the pair scan of buffer d.b on `ps` once its partition has finished, with no
// wait packet.  A wait packet in front of the pair kernel (hipStreamWaitEvent on the
// partition) cost ~12 us between consecutive pair scans at 1M objects even though the
// partition had finished ~50 us before: the host waits instead (the partition of step k+1
// ends while the pair scan of step k still runs, so the next pair scan is queued before
// the pair stream runs dry).
int issue_scan(ladds_ctx *ctx, const PendingScan &d, bool digest, hipStream_t ps) {
  for (;;) {
    const hipError_t q = hipEventQuery(ctx->ev_part[d.b]);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) return ctx->hipfail(q, "partition event");
  }
  int rc = launch_pairs_of(ctx, d.b, d.dt_interp, digest, d.iteration, nullptr, ps, d.span);
  if (rc == LADDS_OK) HIPCHK(ctx, hipEventRecord(ctx->ev_pair[d.b], ps));
  return rc;
}

// Pair-scan streams of the overlapped loop.  On one stream each pair scan starts only after
// the previous one has ended, and ~12 us pass between them (its last workgroups drain, the
// next launch ramps up).  Alternating two streams lets a pair scan start while the previous
// one drains; the scans share only atomic counters and the pair log.  Measured at 1M
// objects: split 72.5 -> 69.0 us per step; FP64 101 -> 104 us (two 80 KB FP64 pair
// workgroups fill a CU's LDS, so the overlapping scans only contend).  Spawn protection
// accounts the scans in order: one stream then.
int pair_streams(const ladds_ctx *ctx) {
  if (ctx->tune_pair_streams) return ctx->tune_pair_streams;
  return ctx->form == kFormFp64 ? 1 : 2;
}

// Software-pipelined step graph for state buffer b: the pair scan of the state in b (on a
// forked branch) runs beside the propagation of b into 1 - b and its partition (one step
// ahead, environment step + 1); then the device step advances.  Two graphs, b = 0 and 1,
// alternate.  Nothing is memset inside: the partition kernel resets its scratch arena.
int capture_step_graph(ladds_ctx *ctx, int b, double burnup_r2, bool digest, hipGraphExec_t *exec) {
  *exec = nullptr;
  hipGraph_t graph = nullptr;
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = LADDS_OK;
  hipError_t e = hipEventRecord(ctx->ev_part[0], ctx->stream);  // fork
  if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream2, ctx->ev_part[0], 0);
  if (e == hipSuccess)
    rc = launch_pairs_of(ctx, b, ctx->cfg.dt, digest, 0, ctx->d_step, ctx->stream2, ctx->protected_objects ? 1 : 0);
  if (e == hipSuccess && rc == LADDS_OK) rc = launch_prop_part(ctx, b, 0, 1, true, burnup_r2, ctx->d_step, 1u);
  if (e == hipSuccess && rc == LADDS_OK) e = hipEventRecord(ctx->ev_pair[0], ctx->stream2);  // join
  if (e == hipSuccess && rc == LADDS_OK) e = hipStreamWaitEvent(ctx->stream, ctx->ev_pair[0], 0);
  if (e == hipSuccess && rc == LADDS_OK) e = launch_advance_step(ctx->d_step, ctx->stream);
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  if (rc == LADDS_OK && e == hipSuccess) e = ec;
  if (rc == LADDS_OK && e == hipSuccess) e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (rc != LADDS_OK) return rc;
  if (e != hipSuccess) return ctx->hipfail(e, "step graph capture");
  return LADDS_OK;
}

int compact_if_needed(ladds_ctx *ctx) {
  if (ctx->n == 0) return LADDS_OK;
  const uint32_t ntiles = (uint32_t)((ctx->n + kCompactTile - 1) / kCompactTile);
  if (!ctx->d_alt.rx) HIPCHK(ctx, alloc_objects(ctx->d_alt, ctx->cap));
  DeviceObjects dst = ctx->d_alt;
  dst.n = ctx->n;
  hipError_t e = launch_count_alive(ctx->d, ctx->tile_sum, ntiles, ctx->stream);
  if (e == hipSuccess)
    e = timed(ctx, "compact",
              [&] { return launch_compact(ctx->d, dst, ctx->tile_sum, ntiles, ctx->n_alive, ctx->stream); });
  if (e != hipSuccess) return ctx->hipfail(e, "compaction launch");
  uint32_t alive = 0;
  HIPCHK(ctx, hipMemcpyAsync(&alive, ctx->n_alive, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (alive != ctx->n) {
    std::vector<int32_t> act;
    if (ctx->scen_off.size() > 2) {  // the flags of the pre-compaction arrays give the new offsets
      act.resize(ctx->n);
      HIPCHK(ctx, hipMemcpy(act.data(), ctx->d.activity, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost));
    }
    std::swap(ctx->d, ctx->d_alt);
    ctx->keys_valid = false;
    ctx->d.n = alive;
    ctx->d_alt.n = 0;
    ctx->n = alive;
    if (!act.empty()) {
      std::vector<uint64_t> off(ctx->scen_off.size(), 0);
      for (size_t sc = 0; sc + 1 < ctx->scen_off.size(); ++sc) {
        uint64_t kept = 0;
        for (uint64_t i = ctx->scen_off[sc]; i < ctx->scen_off[sc + 1]; ++i) kept += act[i] >= 0 ? 1 : 0;
        off[sc + 1] = off[sc] + kept;
      }
      ctx->scen_off = off;
      std::vector<double> rx(alive);
      if (alive) HIPCHK(ctx, hipMemcpy(rx.data(), ctx->d.rx, sizeof(double) * alive, hipMemcpyDeviceToHost));
      const int rc = setup_grid(ctx, alive, rx.data());
      if (rc != LADDS_OK) return rc;
    }
  }
  return LADDS_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *ladds_version(void) { return "ladds_amd 0.1 (gfx950)"; }
int ladds_abi_version(void) { return LADDS_ABI_VERSION; }

int ladds_device_count(int *count) {
  if (!count) return LADDS_E_INVALID;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return LADDS_OK;
}

void ladds_config_default(ladds_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->dt = 10.0;                       // cfg/default_cfg.yaml:12
  cfg->t0 = 0.0;                        // Simulation.cpp:143
  cfg->components = LADDS_KEP;
  cfg->precision = LADDS_FP64;
  cfg->cutoff_km = 80.0;                // cfg/default_cfg.yaml:59
  cfg->collision_distance_factor = 1.0; // Simulation.cpp:233
  cfg->min_detection_radius_m = 0.05;   // Simulation.cpp:245
  cfg->evasion_cutoff_km = 0.1;         // Simulation.cpp:246
  cfg->threshold_mode = LADDS_THRESHOLD_RADIUS_SCALED;
  cfg->threshold_km = 1.0;
  cfg->cell_size_km = 0.0;
  cfg->scan_variant = 0;
  cfg->block_size = 0;
  cfg->device = 0;
  cfg->pair_output = LADDS_PAIRS_LIST;
  cfg->log_capacity = 1u << 20;
  cfg->step_graph = 0;
  cfg->scan_chunks = 0;
  cfg->tune_simulate = 0;
}

int ladds_create(const ladds_config *cfg, ladds_ctx **out) {
  if (!cfg || !out) return LADDS_E_INVALID;
  *out = nullptr;
  if (!(cfg->dt > 0.0) || !(cfg->cutoff_km > 0.0) || cfg->precision < 0 || cfg->precision > 2 ||
      cfg->scan_variant < 0 || cfg->scan_variant > 2 ||
      cfg->threshold_mode < 0 || cfg->threshold_mode > 1 || cfg->pair_output < 0 || cfg->pair_output > 1)
    return LADDS_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) return LADDS_E_NODEVICE;
  ladds_ctx *ctx = new ladds_ctx();
  ctx->cfg = *cfg;
  ctx->form = cfg->precision;
  ctx->device = cfg->device;
  ctx->variant = (cfg->scan_variant == kVariantAtomicSort) ? kVariantAtomicSort : kVariantPlaneTiles;
  yoshida_coefficients(ctx->c, ctx->dcoef);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream4, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join4, hipEventDisableTiming);
  for (int b = 0; b < 4 && e == hipSuccess; ++b) {
    e = hipEventCreateWithFlags(&ctx->ev_prop[b], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_part[b], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_pair[b], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void **>(&ctx->h_rb), sizeof(HostReadback),
                      hipHostMallocMapped | hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void **>(&ctx->d_rb_host), ctx->h_rb, 0);
  if (e == hipSuccess) e = upload_atmosphere();
  if (e == hipSuccess) e = configure_tile_kernels();
  if (e == hipSuccess) e = dalloc(&ctx->arena_top, kScanCounterWords);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  if (e == hipSuccess) {
    DeviceReadback *rb = nullptr;
    e = dalloc(&rb, 1);
    if (e == hipSuccess) {
      ctx->counters = &rb->sc;  // the block's start: freed through this pointer
      ctx->log_count = &rb->log_count;
      e = hipMemset(rb, 0, sizeof(DeviceReadback));
    }
  }
  if (e == hipSuccess) e = dalloc(&ctx->n_alive, 1);
  if (e != hipSuccess) {
    ladds_destroy(ctx);
    return (e == hipErrorOutOfMemory) ? LADDS_E_NOMEM : LADDS_E_HIP;
  }
  const int rc = ensure_log(ctx, std::max<uint64_t>(cfg->log_capacity, 1024));
  if (rc != LADDS_OK) {
    ladds_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return LADDS_OK;
}

int ladds_destroy(ladds_ctx *ctx) {
  if (!ctx) return LADDS_E_INVALID;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) hipStreamSynchronize(ctx->stream2);
  if (ctx->stream3) hipStreamSynchronize(ctx->stream3);
  if (ctx->stream4) hipStreamSynchronize(ctx->stream4);
  for (PipeBuffer &pb : ctx->pipe) {
    for (int k = 0; k < 3; ++k) {
      hipFree(pb.r[k]);
      hipFree(pb.v[k]);
    }
    hipFree(pb.act);
    hipFree(pb.keys);
    hipFree(pb.rec);
    hipFree(pb.tps);
    hipFree(pb.arena_top);
    hipFree(pb.scratch);
  }
  for (int b = 0; b < 4; ++b) {
    if (ctx->ev_prop[b]) hipEventDestroy(ctx->ev_prop[b]);
    if (ctx->ev_part[b]) hipEventDestroy(ctx->ev_part[b]);
    if (ctx->ev_pair[b]) hipEventDestroy(ctx->ev_pair[b]);
  }
  if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
  if (ctx->ev_join4) hipEventDestroy(ctx->ev_join4);
  if (ctx->h_rb) hipHostFree(ctx->h_rb);
  if (ctx->h_log) hipHostFree(ctx->h_log);
  if (ctx->h_x) hipHostFree(ctx->h_x);
  hipFree(ctx->d_spawn);
  if (ctx->stream2) hipStreamDestroy(ctx->stream2);
  if (ctx->stream3) hipStreamDestroy(ctx->stream3);
  if (ctx->stream4) hipStreamDestroy(ctx->stream4);
  free_objects(ctx->d);
  free_objects(ctx->d_alt);
  hipFree(ctx->count);
  hipFree(ctx->slot);
  hipFree(ctx->start);
  hipFree(ctx->tile_sum);
  hipFree(ctx->rec);
  hipFree(ctx->n_alive);
  hipFree(ctx->tps);
  hipFree(ctx->scratch);
  hipFree(ctx->arena_top);
  hipFree(ctx->env64);
  hipFree(ctx->env32);
  hipFree(ctx->log);
  hipFree(ctx->counters);  // with the log count (one DeviceReadback block)
  hipFree(ctx->xbuf);
  hipFree(ctx->xcount);
  for (auto &kv : ctx->prof)
    for (EventPair &ev : kv.second.used) ctx->event_pool.push_back(ev);
  for (EventPair &ev : ctx->event_pool) {
    hipEventDestroy(ev.a);
    hipEventDestroy(ev.b);
  }
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  hipFree(ctx->d_tile_begin);
  hipFree(ctx->d_scen_tile);
  hipFree(ctx->d_plane_order);
  hipFree(ctx->d_step);
  delete ctx;
  return LADDS_OK;
}

const char *ladds_last_error(const ladds_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  const char *e = noalloc_error();  // the context-free entry point's last error on this thread
  return e[0] ? e : "null context";
}

int ladds_configure(ladds_ctx *ctx, const ladds_config *cfg) {
  if (!ctx || !cfg) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (!(cfg->dt > 0.0) || cfg->threshold_mode < 0 || cfg->threshold_mode > 1 || cfg->pair_output < 0 ||
      cfg->pair_output > 1)
    return ctx->fail(LADDS_E_INVALID, "invalid dt, threshold_mode or pair_output");
  const ladds_config &old = ctx->cfg;
  const bool fixed_changed = cfg->precision != old.precision || cfg->cutoff_km != old.cutoff_km ||
                             cfg->cell_size_km != old.cell_size_km || cfg->scan_variant != old.scan_variant ||
                             cfg->device != old.device;
  if (fixed_changed) {
    if (ctx->n > 0)
      return ctx->fail(LADDS_E_STATE, "precision, cutoff, cell size, scan variant and device are fixed once objects are set");
    if (!(cfg->cutoff_km > 0.0) || cfg->precision < 0 || cfg->precision > 2 || cfg->scan_variant < 0 ||
        cfg->scan_variant > 2 || cfg->device != old.device)
      return ctx->fail(LADDS_E_INVALID, "invalid precision, cutoff or scan variant (device is fixed)");
    ctx->form = cfg->precision;
    ctx->variant = (cfg->scan_variant == kVariantAtomicSort) ? kVariantAtomicSort : kVariantPlaneTiles;
  }
  // A new t0 restarts the clock there; a new dt alone keeps the elapsed time (the clock
  // continues from the current time with the new step).
  const double t_now = ctx->time();
  const bool new_t0 = cfg->t0 != old.t0, new_dt = cfg->dt != old.dt;
  const uint64_t log_cap = cfg->log_capacity;
  ctx->cfg = *cfg;
  if (new_t0) {
    ctx->step_count = 0;
  } else if (new_dt) {
    ctx->cfg.t0 = t_now;
    ctx->step_count = 0;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return ensure_log(ctx, std::max<uint64_t>(log_cap, 1024));
}

}  // extern "C"

namespace {

bool has_parents(const ladds_objects *o) {
  if (!o->parent) return false;
  for (size_t i = 0; i < o->n; ++i)
    if (o->parent[i] != UINT64_MAX) return true;
  return false;
}

int check_objects(ladds_ctx *ctx, const ladds_objects *o) {
  if (o->n > 0 && (!o->rx || !o->ry || !o->rz || !o->vx || !o->vy || !o->vz || !o->radius_m))
    return ctx->fail(LADDS_E_INVALID, "positions, velocities and radius_m are required");
  if (o->activity)
    for (size_t i = 0; i < o->n; ++i)
      if (o->activity[i] < 0 || o->activity[i] > 2) return ctx->fail(LADDS_E_INVALID, "activity must be 0, 1 or 2");
  return LADDS_OK;
}

// Host preparation (derived coefficients, defaults) and upload of objects into
// [offset, offset + o->n) of the device arrays.
int upload_objects(ladds_ctx *ctx, const ladds_objects *o, size_t offset) {
  const size_t n = o->n;
  if (n == 0) return LADDS_OK;
  std::vector<double> radius(o->radius_m, o->radius_m + n), aom(n, 0.0), bc(n, 0.0), cr(n, 1.0), srp64(n),
      drag64(n);
  std::vector<float> srp32(n), drag32(n);
  std::vector<int32_t> act(n, LADDS_PASSIVE);
  std::vector<uint64_t> id(n), parent(n, UINT64_MAX);
  for (size_t i = 0; i < n; ++i) {
    if (o->aom) aom[i] = o->aom[i];
    if (o->bc_inv) bc[i] = o->bc_inv[i];
    if (o->cr) cr[i] = o->cr[i];
    if (o->activity) act[i] = o->activity[i];
    id[i] = o->id ? o->id[i] : (uint64_t)(offset + i);
    if (o->parent) parent[i] = o->parent[i];
    srp64[i] = kSrpCoef * cr[i] * aom[i];
    drag64[i] = kDragCoef * bc[i];
    srp32[i] = (float)srp64[i];
    drag32[i] = (float)drag64[i];
  }
  auto up = [&](auto *dst, const void *src, size_t elem) {
    return hipMemcpyAsync(dst + offset, src, elem * n, hipMemcpyHostToDevice, ctx->stream);
  };
  HIPCHK(ctx, up(ctx->d.rx, o->rx, 8));
  HIPCHK(ctx, up(ctx->d.ry, o->ry, 8));
  HIPCHK(ctx, up(ctx->d.rz, o->rz, 8));
  HIPCHK(ctx, up(ctx->d.vx, o->vx, 8));
  HIPCHK(ctx, up(ctx->d.vy, o->vy, 8));
  HIPCHK(ctx, up(ctx->d.vz, o->vz, 8));
  HIPCHK(ctx, up(ctx->d.radius_m, radius.data(), 8));
  HIPCHK(ctx, up(ctx->d.aom, aom.data(), 8));
  HIPCHK(ctx, up(ctx->d.bc_inv, bc.data(), 8));
  HIPCHK(ctx, up(ctx->d.cr, cr.data(), 8));
  HIPCHK(ctx, up(ctx->d.srp64, srp64.data(), 8));
  HIPCHK(ctx, up(ctx->d.drag64, drag64.data(), 8));
  HIPCHK(ctx, up(ctx->d.srp32, srp32.data(), 4));
  HIPCHK(ctx, up(ctx->d.drag32, drag32.data(), 4));
  HIPCHK(ctx, up(ctx->d.activity, act.data(), 4));
  HIPCHK(ctx, up(ctx->d.id, id.data(), 8));
  HIPCHK(ctx, up(ctx->d.parent, parent.data(), 8));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return LADDS_OK;
}

// Object capacity of at least `need`; with `keep` the first ctx->n objects are preserved.
int reserve_objects(ladds_ctx *ctx, size_t need, bool keep) {
  if (need <= ctx->cap) return LADDS_OK;
  const size_t cap = keep ? std::max(need, 2 * ctx->cap) : need;
  DeviceObjects fresh{};
  const hipError_t ae = alloc_objects(fresh, cap);
  if (ae != hipSuccess) {
    free_objects(fresh);
    return ctx->hipfail(ae, "alloc_objects");
  }
  if (keep && ctx->n) {
    const size_t n = ctx->n;
    const DeviceObjects &o = ctx->d;
    hipError_t e = hipSuccess;
    auto cp = [&](void *dst, const void *src, size_t elem) {
      if (e == hipSuccess) e = hipMemcpyAsync(dst, src, elem * n, hipMemcpyDeviceToDevice, ctx->stream);
    };
    cp(fresh.rx, o.rx, 8);
    cp(fresh.ry, o.ry, 8);
    cp(fresh.rz, o.rz, 8);
    cp(fresh.vx, o.vx, 8);
    cp(fresh.vy, o.vy, 8);
    cp(fresh.vz, o.vz, 8);
    cp(fresh.radius_m, o.radius_m, 8);
    cp(fresh.aom, o.aom, 8);
    cp(fresh.bc_inv, o.bc_inv, 8);
    cp(fresh.cr, o.cr, 8);
    cp(fresh.srp64, o.srp64, 8);
    cp(fresh.drag64, o.drag64, 8);
    cp(fresh.srp32, o.srp32, 4);
    cp(fresh.drag32, o.drag32, 4);
    cp(fresh.activity, o.activity, 4);
    cp(fresh.id, o.id, 8);
    cp(fresh.parent, o.parent, 8);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      free_objects(fresh);
      return ctx->hipfail(e, "object copy");
    }
  }
  fresh.n = ctx->d.n;
  free_objects(ctx->d);
  free_objects(ctx->d_alt);
  ctx->d = fresh;
  hipFree(ctx->slot);
  hipFree(ctx->rec);
  hipFree(ctx->scratch);
  ctx->slot = nullptr;
  ctx->rec = nullptr;
  ctx->scratch = nullptr;  // the scan's scratch arena is sized by setup_grid
  ctx->scratch_bytes = 0;
  HIPCHK(ctx, dalloc(&ctx->slot, cap));
  HIPCHK(ctx, dalloc(reinterpret_cast<RecordF64 **>(&ctx->rec), cap));
  ctx->cap = cap;
  return LADDS_OK;
}

}  // namespace

extern "C" {

int ladds_set_objects(ladds_ctx *ctx, const ladds_objects *o) {
  if (!ctx || !o) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (o->n >= (size_t)UINT32_MAX) return ctx->fail(LADDS_E_INVALID, "at most 2^32-2 objects per context");
  int rc = check_objects(ctx, o);
  if (rc != LADDS_OK) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n = o->n;
  rc = reserve_objects(ctx, n, false);
  if (rc != LADDS_OK) return rc;
  ctx->scen_off.clear();  // a new population is a single scenario until ladds_set_scenarios
  rc = setup_grid(ctx, n, o->rx);
  if (rc != LADDS_OK) return rc;
  ctx->n = n;
  ctx->d.n = n;
  ctx->protected_objects = has_parents(o);
  return upload_objects(ctx, o, 0);
}

int ladds_add_objects(ladds_ctx *ctx, const ladds_objects *o) {
  if (!ctx || !o) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (ctx->scen_off.size() > 2) return ctx->fail(LADDS_E_STATE, "objects cannot be added to an ensemble");
  if (ctx->n + o->n >= (size_t)UINT32_MAX) return ctx->fail(LADDS_E_INVALID, "at most 2^32-2 objects per context");
  int rc = check_objects(ctx, o);
  if (rc != LADDS_OK || o->n == 0) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n0 = ctx->n, n = n0 + o->n;
  rc = reserve_objects(ctx, n, true);
  if (rc != LADDS_OK) return rc;
  rc = upload_objects(ctx, o, n0);
  if (rc != LADDS_OK) return rc;
  ctx->n = n;
  ctx->d.n = n;
  ctx->protected_objects = ctx->protected_objects || has_parents(o);
  std::vector<double> rx(n);
  HIPCHK(ctx, hipMemcpy(rx.data(), ctx->d.rx, sizeof(double) * n, hipMemcpyDeviceToHost));
  return setup_grid(ctx, n, rx.data());
}

int ladds_set_scenarios(ladds_ctx *ctx, size_t nscenarios, const uint64_t *offsets) {
  if (!ctx || nscenarios == 0 || !offsets) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (ctx->n == 0) return ctx->fail(LADDS_E_STATE, "set objects before scenarios");
  if (ctx->variant != kVariantPlaneTiles) return ctx->fail(LADDS_E_STATE, "scenarios need the plane-tile scan");
  if (offsets[0] != 0 || offsets[nscenarios] != ctx->n)
    return ctx->fail(LADDS_E_INVALID, "scenario offsets must start at 0 and end at the object count");
  for (size_t k = 0; k < nscenarios; ++k)
    if (offsets[k + 1] < offsets[k]) return ctx->fail(LADDS_E_INVALID, "scenario offsets must not decrease");
  if (nscenarios > (1u << 20)) return ctx->fail(LADDS_E_INVALID, "too many scenarios");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (nscenarios == 1)
    ctx->scen_off.clear();
  else
    ctx->scen_off.assign(offsets, offsets + nscenarios + 1);
  std::vector<double> rx(ctx->n);
  HIPCHK(ctx, hipMemcpy(rx.data(), ctx->d.rx, sizeof(double) * ctx->n, hipMemcpyDeviceToHost));
  return setup_grid(ctx, ctx->n, rx.data());
}

size_t ladds_num_scenarios(const ladds_ctx *ctx) {
  if (!ctx) return 0;
  return ctx->scen_off.size() > 2 ? ctx->scen_off.size() - 1 : 1;
}

int ladds_get_scenarios(const ladds_ctx *ctx, uint64_t *offsets) {
  if (!ctx || !offsets) return LADDS_E_INVALID;
  if (ctx->scen_off.size() > 2) {
    std::copy(ctx->scen_off.begin(), ctx->scen_off.end(), offsets);
  } else {
    offsets[0] = 0;
    offsets[1] = ctx->n;
  }
  return LADDS_OK;
}

size_t ladds_num_objects(const ladds_ctx *ctx) { return ctx ? ctx->n : 0; }

int ladds_get_state(ladds_ctx *ctx, double *rx, double *ry, double *rz, double *vx, double *vy, double *vz) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t b = 8 * ctx->n;
  double *dst[6] = {rx, ry, rz, vx, vy, vz};
  double *src[6] = {ctx->d.rx, ctx->d.ry, ctx->d.rz, ctx->d.vx, ctx->d.vy, ctx->d.vz};
  for (int k = 0; k < 6; ++k)
    if (dst[k] && b) HIPCHK(ctx, hipMemcpyAsync(dst[k], src[k], b, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return LADDS_OK;
}

int ladds_get_ids(ladds_ctx *ctx, uint64_t *id) {
  if (!ctx || !id) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (ctx->n) HIPCHK(ctx, hipMemcpyAsync(id, ctx->d.id, 8 * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return LADDS_OK;
}

double ladds_get_time(const ladds_ctx *ctx) { return ctx ? ctx->time() : 0.0; }

int ladds_set_time(ladds_ctx *ctx, double t) {
  if (!ctx) return LADDS_E_INVALID;
  ctx->cfg.t0 = t;
  ctx->step_count = 0;
  return LADDS_OK;
}

int ladds_propagate(ladds_ctx *ctx, int nsteps) {
  if (!ctx || nsteps < 0) return LADDS_E_INVALID;
  if (nsteps == 0 || ctx->n == 0) {
    ctx->step_count += (uint64_t)nsteps;
    return LADDS_OK;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  {
    const int rc = alloc_env(ctx);
    if (rc != LADDS_OK) return rc;
  }
  // the last launch also writes the bucket keys of the final state (plane-tile scan), so a
  // detection that follows (iteratePairwise after integrate) skips its binning pass
  const bool tiles = ctx->variant == kVariantPlaneTiles;
  for (int seg0 = 0; seg0 < nsteps; seg0 += kEnvSegment) {
    const int seg = std::min(kEnvSegment, nsteps - seg0);
    // The environment of a whole segment is staged once and serves the calls that follow
    // (integrate(false) once per iteration) until the clock leaves it.
    size_t off = 0;
    int rc = env_window(ctx, seg, &off);
    for (int s = 0; s < seg && rc == LADDS_OK; s += kChunk) {
      const int k = std::min(kChunk, seg - s);
      rc = launch_propagate_steps(ctx, off + (size_t)s, k, tiles && seg0 + s + k == nsteps, 0.0);
    }
    if (rc != LADDS_OK) return rc;
    ctx->step_count += (uint64_t)seg;
  }
  ctx->keys_valid = tiles;
  // no wait here: a detection that follows is ordered on the same stream, and calls that
  // touch device memory from the host settle first
  ctx->pending = true;
  return LADDS_OK;
}

static int run_detection_sync(ladds_ctx *ctx, double dt_interp, bool digest, ScanCounters *out_counters,
                              uint64_t *entries) {
  HIPCHK(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(ScanCounters), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->log_count, 0, sizeof(unsigned long long), ctx->stream));
  int rc = launch_detection(ctx, ctx->keys_valid, dt_interp, digest, 0);
  if (rc != LADDS_OK) return rc;
  unsigned long long cnt = 0;
  HIPCHK(ctx, hipMemcpyAsync(out_counters, ctx->counters, sizeof(ScanCounters), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&cnt, ctx->log_count, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->pending = false;
  *entries = cnt;
  if (out_counters->overflow) return ctx->fail(LADDS_E_OVERFLOW, "conjunction scan scratch arena exhausted");
  return LADDS_OK;
}

// Device time of one detection (keys, partition, pair scan in digest mode) on the
// current state: the minimum over `reps` runs after one warm-up run.
// Smallest relative gain for which the autotuner moves to a candidate shape.
constexpr double kTuneMargin = 0.01;

// Device time of one detection (partition + pair scan, digest output) with the current
// shape: the best of `reps` runs after a warm-up.
static int time_detection(ladds_ctx *ctx, double dt_interp, int reps, double *us) {
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  double best = 1e30;
  int rc = LADDS_OK;
  for (int r = 0; r <= reps && rc == LADDS_OK; ++r) {
    (void)hipEventRecord(e0, ctx->stream);
    HIPCHK(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(ScanCounters), ctx->stream));
    rc = launch_detection(ctx, false, dt_interp, true, 0);
    (void)hipEventRecord(e1, ctx->stream);
    if (hipEventSynchronize(e1) != hipSuccess) rc = ctx->fail(LADDS_E_HIP, "detection timing failed");
    float ms = 0.f;
    if (rc == LADDS_OK && r > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) best = std::min(best, 1000.0 * ms);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *us = best;
  return rc;
}

// Steps of one timed ladds_simulate burst (ladds_config.tune_simulate): long enough that the
// call's fixed cost and the timing noise stay below the tuner's 1% margin.
constexpr int kTuneSteps = 48;

// What a ladds_simulate burst changes, kept aside by the tuner (ladds_config.tune_simulate):
// the state, activity and parent ids (spawn protection may reset them) of the n objects, and
// the host-side step count and spawn-protection count.
struct SimSnapshot {
  double *r[3] = {nullptr, nullptr, nullptr}, *v[3] = {nullptr, nullptr, nullptr};
  int32_t *act = nullptr;
  uint64_t *parent = nullptr;
  uint64_t step_count = 0, quiet = 0;
  bool protected_objects = false;
  void release() {
    for (int k = 0; k < 3; ++k) {
      hipFree(r[k]);
      hipFree(v[k]);
    }
    hipFree(act);
    hipFree(parent);
  }
};

int snapshot_copy(ladds_ctx *ctx, SimSnapshot &s, bool save) {
  const size_t n = ctx->n;
  double *dr[3] = {ctx->d.rx, ctx->d.ry, ctx->d.rz}, *dv[3] = {ctx->d.vx, ctx->d.vy, ctx->d.vz};
  auto cp = [&](void *live, void *kept, size_t bytes) {
    return hipMemcpyAsync(save ? kept : live, save ? live : kept, bytes, hipMemcpyDeviceToDevice, ctx->stream);
  };
  for (int k = 0; k < 3; ++k) {
    HIPCHK(ctx, cp(dr[k], s.r[k], sizeof(double) * n));
    HIPCHK(ctx, cp(dv[k], s.v[k], sizeof(double) * n));
  }
  HIPCHK(ctx, cp(ctx->d.activity, s.act, sizeof(int32_t) * n));
  HIPCHK(ctx, cp(ctx->d.parent, s.parent, sizeof(uint64_t) * n));
  if (save) {
    s.step_count = ctx->step_count;
    s.quiet = ctx->quiet;
    s.protected_objects = ctx->protected_objects;
  } else {
    ctx->step_count = s.step_count;
    ctx->quiet = s.quiet;
    ctx->protected_objects = s.protected_objects;
    ctx->keys_valid = false;
  }
  return LADDS_OK;
}

// Device time per step of kTuneSteps ladds_simulate steps (detection every step, no
// burn-up; events on the context's stream around the call, which returns synchronised) with
// the current shape: the best of `reps` bursts after a warm-up, each from the saved state,
// which is restored at the end with the time, the spawn-protection count and an empty pair
// log.
int time_simulate(ladds_ctx *ctx, int reps, double *us) {
  SimSnapshot snap;
  const size_t n = ctx->n;
  int rc = LADDS_OK;
  hipError_t e = hipSuccess;
  for (int k = 0; k < 3 && e == hipSuccess; ++k) {
    e = dalloc(&snap.r[k], n);
    if (e == hipSuccess) e = dalloc(&snap.v[k], n);
  }
  if (e == hipSuccess) e = dalloc(&snap.act, n);
  if (e == hipSuccess) e = dalloc(&snap.parent, n);
  if (e != hipSuccess) {
    snap.release();
    return ctx->hipfail(e, "tuner snapshot allocation");
  }
  rc = snapshot_copy(ctx, snap, true);
  const bool saved = rc == LADDS_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == LADDS_OK && ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess))
    rc = ctx->hipfail(e, "tuner events");
  double best = 1e30;
  for (int r = 0; r <= reps && rc == LADDS_OK; ++r) {
    (void)hipEventRecord(e0, ctx->stream);
    rc = ladds_simulate(ctx, 0, kTuneSteps, 1, 0.0, nullptr);
    (void)hipEventRecord(e1, ctx->stream);
    float ms = 0.f;
    if (rc == LADDS_OK && (e = hipEventSynchronize(e1)) == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (rc == LADDS_OK && e != hipSuccess) rc = ctx->hipfail(e, "tuner timing");
    if (rc == LADDS_OK) rc = snapshot_copy(ctx, snap, false);
    if (rc == LADDS_OK) {
      e = hipMemsetAsync(ctx->log_count, 0, sizeof(unsigned long long), ctx->stream);
      if (e != hipSuccess) rc = ctx->hipfail(e, "tuner log reset");
      ctx->log_nonempty = false;
    }
    if (rc == LADDS_OK && r > 0) best = std::min(best, 1000.0 * ms);
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (saved && rc != LADDS_OK) {
    // a failed burst: the context still continues from the saved state with an empty log
    // (the first error is the one reported)
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipStreamSynchronize(ctx->stream3);
    (void)snapshot_copy(ctx, snap, false);
    (void)hipMemsetAsync(ctx->log_count, 0, sizeof(unsigned long long), ctx->stream);
    ctx->log_nonempty = false;
  }
  e = hipStreamSynchronize(ctx->stream);
  if (rc == LADDS_OK && e != hipSuccess) rc = ctx->hipfail(e, "tuner restore");
  snap.release();
  *us = best / kTuneSteps;
  return rc;
}

int ladds_autotune(ladds_ctx *ctx, double dt_interp, int repetitions, ladds_tuning *out) {
  if (!ctx || repetitions < 1) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (ctx->n < 2) return ctx->fail(LADDS_E_STATE, "set objects before autotuning");
  if (ctx->variant != kVariantPlaneTiles) return ctx->fail(LADDS_E_STATE, "autotuning covers the plane-tile scan");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // the tuner's detections share the device pair log with ladds_simulate: refuse to drop
  // an unfetched log
  if (ctx->log_nonempty) return ctx->fail(LADDS_E_STATE, "pair log not empty: fetch and clear it before ladds_autotune");
  const size_t n = ctx->n;
  std::vector<double> rx(n);
  HIPCHK(ctx, hipMemcpy(rx.data(), ctx->d.rx, sizeof(double) * n, hipMemcpyDeviceToHost));
  // start from the untuned default
  ctx->tune_cell_km = 0.0;
  ctx->tune_pb = -1;
  ctx->tune_threads = 0;
  ctx->tune_tile = 0;
  ctx->tune_chunks = 0;
  ctx->tune_fx = 0.0;
  ctx->tune_wgs = 0;
  ctx->tune_order = 0;
  ctx->tune_ens_plane = 0;
  ctx->tune_rows = 0;
  ctx->tune_bcells = 0;
  ctx->tune_persistent = 0;
  ctx->tune_pair_streams = 0;
  int rc = setup_grid(ctx, n, rx.data());
  if (rc != LADDS_OK) return rc;
  const double base_cell = 1.0 / ctx->grid.inv_cell;
  const bool ensemble = ctx->tiles.nscen > 1;
  // current point of the search: {plane bits (-1: sized from the objects per plane), threads,
  // tile, chunks, workgroups per CU, dispatch order, ensemble objects per plane (0: 64), row
  // height and bucket width in cells, persistent pair kernel, pair-scan streams (0: by form)}
  double cur_cell = base_cell, cur_fx = 1.0;
  int cur[11] = {-1, (int)ctx->tiles.pair_threads, (int)ctx->tiles.tile_size, (int)ctx->tiles.nchunk, 0, 0, 0, 1, 1, 0, 0};
  double default_us = 0.0;
  // the objective: one detection, or (ladds_config.tune_simulate) a burst of simulate steps
  auto measure = [&](double *t) {
    return ctx->cfg.tune_simulate ? time_simulate(ctx, repetitions, t)
                                  : time_detection(ctx, dt_interp, repetitions, t);
  };
  rc = measure(&default_us);
  if (rc != LADDS_OK) return rc;
  double best_us = default_us;
  int tried = 1;
  auto apply = [&](double cell, double fx, const int *v) {
    ctx->tune_cell_km = cell;
    ctx->tune_fx = fx;
    ctx->tune_pb = v[0];
    ctx->tune_threads = v[1];
    ctx->tune_tile = v[2];
    ctx->tune_chunks = v[3];
    ctx->tune_wgs = v[4];
    ctx->tune_order = v[5];
    ctx->tune_ens_plane = v[6];
    ctx->tune_rows = v[7];
    ctx->tune_bcells = v[8];
    ctx->tune_persistent = v[9];
    ctx->tune_pair_streams = v[10];
    return setup_grid(ctx, n, rx.data());
  };
  // Coordinate descent, one dimension at a time, keeping every improvement: chunks per
  // plane (load balance of dense planes), workgroup size, bucket shape, cell size, plane
  // shape, buckets per plane, tile, LDS share (workgroups per CU), dispatch order, persistent
  // pair kernel, objects per plane (ensembles).  Ensembles take a second pass: their plane size, workgroup size
  // and buckets interact (measured on 256 scenarios: 256-object planes need 256 threads and
  // 512 buckets).
  const int chunk_opts[] = {1, 2, 4, 8, 16, 32};
  const int thread_opts[] = {128, 256, 384, 512, 1024};
  const double cell_opts[] = {1.0, 1.25, 1.5, 2.0};
  const int tile_opts[] = {4096, 8192, 16384};
  const double fx_opts[] = {1.0, 1.5, 2.0, 3.0};
  const int wgs_opts[] = {0, 2, 3, 4, 5};
  const int ens_opts[] = {32, 64, 128, 256, 512};
  struct Cand {
    double cell, fx;
    std::array<int, 11> v;
  };
  // dimensions in search order (9: bucket shape, 10: persistent pair kernel), one pass (two
  // for ensembles), then 11: cell size and buckets per plane together -- one at a time, a
  // cell size that ties at the current bucket count is kept although it loses with more
  // buckets (measured on the 256-scenario ensemble: 12.6 km cells with 512 buckets per
  // plane 6% faster than the 10.1 or 15.2 km cells with 128 that the passes stopped at);
  // last, with the simulate objective only, 12: one or two pair-scan streams; then 1 and 9
  // again
  const int dims[] = {0, 1, 9, 2, 3, 4, 5, 6, 7, 10, 8};
  constexpr int kDims = (int)(sizeof(dims) / sizeof(dims[0]));
  std::vector<int> order(dims, dims + kDims);
  if (ensemble) order.insert(order.end(), dims, dims + kDims);
  order.push_back(11);
  if (ctx->cfg.tune_simulate) order.push_back(12);  // (a single detection cannot see it)
  // workgroup size and bucket shape once more, after the shapes they interact with: two
  // full runs at 1M objects (split, simulate objective) stopped at 256 threads with 2x1
  // buckets (66.7 us per step) and at 512 threads with 2x2 (68.3 us)
  order.push_back(1);
  order.push_back(9);
  for (const int dim : order) {
    if (dim == 8 && !ensemble) continue;
    std::vector<Cand> cands;
    std::array<int, 11> c{cur[0], cur[1], cur[2], cur[3], cur[4], cur[5], cur[6], cur[7], cur[8], cur[9], cur[10]};
    if (dim == 0)
      for (int k : chunk_opts) { c[3] = k; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 1)
      for (int t : thread_opts) { c[1] = t; cands.push_back({cur_cell, cur_fx, c}); }
    // cell size and plane shape change the objects per bucket: each option is tried with the
    // current buckets per plane and with the count sized from it again (they interact: on the
    // 256-scenario ensemble a fixed count kept the search in a 6% slower basin)
    std::array<int, 11> c_auto = c;
    c_auto[0] = -1;
    if (dim == 2)
      for (double f : cell_opts) {
        cands.push_back({base_cell * f, cur_fx, c});
        if (c[0] != -1) cands.push_back({base_cell * f, cur_fx, c_auto});
      }
    if (dim == 3)
      for (double f : fx_opts) {
        cands.push_back({cur_cell, f, c});
        if (c[0] != -1) cands.push_back({cur_cell, f, c_auto});
      }
    if (dim == 4) {
      // buckets per plane around the count the current plane size gives
      rc = apply(cur_cell, cur_fx, cur);
      if (rc != LADDS_OK) return rc;
      const int pb_now = (int)(ctx->grid.by + ctx->grid.bz);
      for (int d = -1; d <= 2; ++d)
        if (pb_now + d >= 4 && pb_now + d <= 12) { c[0] = pb_now + d; cands.push_back({cur_cell, cur_fx, c}); }
    }
    if (dim == 5)
      for (int t : tile_opts) { c[2] = t; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 6)
      for (int w : wgs_opts) { c[4] = w; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 7)
      for (int o = 0; o < 2; ++o) { c[5] = o; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 8)  // ensembles: objects per plane, buckets sized from it again
      for (int e : ens_opts) { c[6] = e; c[0] = -1; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 10)
      for (int p = 0; p < 2; ++p) { c[9] = p; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 12)  // pair-scan streams of the overlapped loop
      for (int p = 1; p <= 2; ++p) { c[10] = p; cands.push_back({cur_cell, cur_fx, c}); }
    if (dim == 11) {
      rc = apply(cur_cell, cur_fx, cur);
      if (rc != LADDS_OK) return rc;
      const int pb_now = (int)(ctx->grid.by + ctx->grid.bz);
      for (double f : cell_opts)
        for (int d = -1; d <= 2; ++d)
          if (pb_now + d >= 4 && pb_now + d <= 12) {
            c[0] = pb_now + d;
            cands.push_back({base_cell * f, cur_fx, c});
          }
    }
    if (dim == 9)  // bucket shape in cells (y x z): 1 x 1, 2 x 1, 2 x 2
      for (int k = 0; k < 3; ++k) {
        c[7] = k ? 2 : 1;
        c[8] = k == 2 ? 2 : 1;
        cands.push_back({cur_cell, cur_fx, c});
      }
    for (const Cand &cand : cands) {
      if (cand.cell == cur_cell && cand.fx == cur_fx && std::equal(cand.v.begin(), cand.v.end(), cur)) continue;
      rc = apply(cand.cell, cand.fx, cand.v.data());
      if (rc != LADDS_OK) return rc;
      double us = 0.0;
      rc = measure(&us);
      if (rc != LADDS_OK) return rc;
      ++tried;
      // keep a candidate only when it is faster by more than the timing noise: near-ties
      // (the 256-scenario ensemble has shapes within 0.2% in detection time that differ by 6%
      // in the overlapped loop) then stay with the earlier, default-side choice
      if (us < best_us * (1.0 - kTuneMargin)) {
        best_us = us;
        cur_cell = cand.cell;
        cur_fx = cand.fx;
        for (int k = 0; k < 11; ++k) cur[k] = cand.v[k];
      }
    }
  }
  rc = apply(cur_cell, cur_fx, cur);
  if (rc != LADDS_OK) return rc;
  HIPCHK(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(ScanCounters), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->log_count, 0, sizeof(unsigned long long), ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (out) {
    out->cell_size_km = 1.0 / ctx->grid.inv_cell;
    out->plane_width_km = 1.0 / ctx->grid.inv_cell_x;
    out->plane_bits = (int32_t)(ctx->grid.by + ctx->grid.bz);
    out->pair_threads = (int32_t)ctx->tiles.pair_threads;
    out->tile_size = (int32_t)ctx->tiles.tile_size;
    out->chunks = (int32_t)ctx->tiles.nchunk;
    out->workgroups_per_cu = (int32_t)ctx->tune_wgs;
    out->plane_order = (int32_t)ctx->tune_order;
    out->ensemble_plane_objects = (int32_t)ctx->tune_ens_plane;
    out->row_cells = ctx->tiles.half_rows ? 2 : 1;
    out->bucket_cells = ctx->tiles.half_buckets ? 2 : 1;
    out->persistent = ctx->tiles.persistent_groups ? 1 : 0;
    out->pair_streams = pair_streams(ctx);
    out->scan_us = best_us;
    out->default_us = default_us;
    out->tried = tried;
  }
  return LADDS_OK;
}

}  // extern "C"

namespace {

int ensure_xbuf(ladds_ctx *ctx, size_t bytes) {
  if (!ctx->xcount) HIPCHK(ctx, dalloc(&ctx->xcount, 2));
  if (bytes <= ctx->xbuf_bytes) return LADDS_OK;
  hipFree(ctx->xbuf);
  ctx->xbuf = nullptr;
  ctx->xbuf_bytes = 0;
  HIPCHK(ctx, hipMalloc(&ctx->xbuf, bytes));
  ctx->xbuf_bytes = bytes;
  return LADDS_OK;
}

}  // namespace

extern "C" {

int ladds_set_tuning(ladds_ctx *ctx, const ladds_tuning *t) {
  if (!ctx || !t) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (ctx->variant != kVariantPlaneTiles) return ctx->fail(LADDS_E_STATE, "tuning covers the plane-tile scan");
  const double radius = search_radius(ctx->form, ctx->cfg.cutoff_km);
  const bool ok = t->cell_size_km >= radius * (1.0 + 1e-9) && t->plane_width_km >= t->cell_size_km &&
                  t->plane_bits >= 4 && t->plane_bits <= 12 &&
                  valid_pair_threads(t->pair_threads) &&
                  (t->tile_size == 4096 || t->tile_size == 8192 || t->tile_size == 16384) && t->chunks >= 1 &&
                  t->chunks <= 32 && t->workgroups_per_cu >= 0 && t->workgroups_per_cu <= 8 &&
                  (t->plane_order == 0 || t->plane_order == 1) && t->ensemble_plane_objects >= 0 &&
                  t->ensemble_plane_objects <= 65536 && t->row_cells >= 0 && t->row_cells <= 2 &&
                  t->bucket_cells >= 0 && t->bucket_cells <= 2 && (t->persistent == 0 || t->persistent == 1) &&
                  t->pair_streams >= 0 && t->pair_streams <= 2;
  if (!ok) return ctx->fail(LADDS_E_INVALID, "tuning values outside the tuner's range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ctx->tune_cell_km = t->cell_size_km;
  ctx->tune_fx = t->plane_width_km / t->cell_size_km;
  ctx->tune_pb = t->plane_bits;
  ctx->tune_threads = t->pair_threads;
  ctx->tune_tile = t->tile_size;
  ctx->tune_chunks = t->chunks;
  ctx->tune_wgs = t->workgroups_per_cu;
  ctx->tune_order = t->plane_order;
  ctx->tune_ens_plane = t->ensemble_plane_objects;
  ctx->tune_rows = t->row_cells;
  ctx->tune_bcells = t->bucket_cells;
  ctx->tune_persistent = t->persistent;
  ctx->tune_pair_streams = t->pair_streams;
  if (ctx->n == 0) return LADDS_OK;  // applied at the next ladds_set_objects
  std::vector<double> rx(ctx->n);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(rx.data(), ctx->d.rx, sizeof(double) * ctx->n, hipMemcpyDeviceToHost));
  return setup_grid(ctx, ctx->n, rx.data());
}

int ladds_conjunctions(ladds_ctx *ctx, double dt_interp, ladds_pair *hits, size_t cap, size_t *n, ladds_pair *evaded,
                       size_t ecap, size_t *ne) {
  if (!ctx || !n || !ne || (cap && !hits) || (ecap && !evaded)) return LADDS_E_INVALID;
  *n = *ne = 0;
  if (ctx->n < 2) return LADDS_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // the device log is shared with ladds_simulate: preserve its contents
  if (ctx->log_nonempty) return ctx->fail(LADDS_E_STATE, "pair log not empty: fetch and clear it before ladds_conjunctions");
  // One synchronisation: the counters and the first entries of the log (pinned staging)
  // come back together; more entries than staged, or than the log holds, cost a second.
  uint64_t entries = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    HIPCHK(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(DeviceReadback), ctx->stream));  // with the log count
    int rc = launch_detection(ctx, ctx->keys_valid, dt_interp, false, 0);
    if (rc != LADDS_OK) return rc;
    if (!ctx->h_log) {
      HIPCHK(ctx, hipHostMalloc(reinterpret_cast<void **>(&ctx->h_log), sizeof(LoggedPair) * 4096, 0));
      ctx->h_log_cap = 4096;
    }
    // the first 256 entries come back with the counters (a call that finds more copies the
    // rest after the first synchronisation)
    const size_t staged = std::min<size_t>({ctx->h_log_cap, ctx->log_cap, 256});
    HIPCHK(ctx, hipMemcpyAsync(&ctx->h_rb->sc, ctx->counters, sizeof(DeviceReadback), hipMemcpyDeviceToHost,
                               ctx->stream));  // counters and log count
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_log, ctx->log, sizeof(LoggedPair) * staged, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->log_count, 0, sizeof(unsigned long long), ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pending = false;
    if (ctx->h_rb->sc.overflow) return ctx->fail(LADDS_E_OVERFLOW, "conjunction scan scratch arena exhausted");
    entries = ctx->h_rb->log_count;
    if (entries > ctx->log_cap) {  // the log was too small: grow it and scan again
      rc = ensure_log(ctx, entries);
      if (rc != LADDS_OK) return rc;
      continue;
    }
    if (entries > staged) {  // more than staged: grow the pinned block if needed, copy the rest
      if (entries > ctx->h_log_cap) {
        size_t want = ctx->h_log_cap;
        while (want < entries) want *= 2;
        LoggedPair *bigger = nullptr;
        HIPCHK(ctx, hipHostMalloc(reinterpret_cast<void **>(&bigger), sizeof(LoggedPair) * want, 0));
        std::memcpy(bigger, ctx->h_log, sizeof(LoggedPair) * staged);
        hipHostFree(ctx->h_log);
        ctx->h_log = bigger;
        ctx->h_log_cap = want;
      }
      HIPCHK(ctx, hipMemcpyAsync(ctx->h_log + staged, ctx->log + staged, sizeof(LoggedPair) * (entries - staged),
                                 hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    break;
  }
  size_t nh = 0, nev = 0;
  for (uint64_t k = 0; k < entries; ++k) {
    const LoggedPair &lp = ctx->h_log[k];
    if (lp.evaded) {
      if (nev < ecap) std::memcpy(&evaded[nev], &lp.pair, sizeof(ladds_pair));
      ++nev;
    } else {
      if (nh < cap) std::memcpy(&hits[nh], &lp.pair, sizeof(ladds_pair));
      ++nh;
    }
  }
  *n = nh;
  *ne = nev;
  if (nh > cap || nev > ecap) return ctx->fail(LADDS_E_OVERFLOW, "pair output capacity too small");
  return LADDS_OK;
}

int ladds_gather_state(ladds_ctx *ctx, const uint32_t *idx, size_t count, double *r, double *v) {
  if (!ctx || (count && (!idx || !r || !v))) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (count == 0) return LADDS_OK;
  for (size_t k = 0; k < count; ++k)
    if (idx[k] >= ctx->n) return ctx->fail(LADDS_E_INVALID, "object index out of range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t in_b = sizeof(uint32_t) * count, out_b = sizeof(double) * 6 * count;
  const size_t bytes = (in_b + 255) / 256 * 256 + out_b;
  int rc = ensure_xbuf(ctx, bytes);
  if (rc != LADDS_OK) return rc;
  if (bytes > ctx->h_x_bytes) {
    if (ctx->h_x) hipHostFree(ctx->h_x);
    ctx->h_x = nullptr;
    ctx->h_x_bytes = 0;
    HIPCHK(ctx, hipHostMalloc(&ctx->h_x, bytes, 0));
    ctx->h_x_bytes = bytes;
  }
  char *hx = static_cast<char *>(ctx->h_x), *dx = static_cast<char *>(ctx->xbuf);
  const size_t out_off = (in_b + 255) / 256 * 256;
  std::memcpy(hx, idx, in_b);
  HIPCHK(ctx, hipMemcpyAsync(dx, hx, in_b, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, launch_gather_state(ctx->d, reinterpret_cast<const uint32_t *>(dx), count,
                                  reinterpret_cast<double *>(dx + out_off), ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(hx + out_off, dx + out_off, out_b, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const double *o = reinterpret_cast<const double *>(hx + out_off);
  for (size_t k = 0; k < count; ++k)
    for (int c = 0; c < 3; ++c) {
      r[3 * k + c] = o[6 * k + c];
      v[3 * k + c] = o[6 * k + 3 + c];
    }
  return LADDS_OK;
}

int ladds_conjunction_digest(ladds_ctx *ctx, double dt_interp, uint64_t *n, uint64_t digest[2], uint64_t *ne,
                             uint64_t edigest[2]) {
  if (!ctx || !n || !digest || !ne || !edigest) return LADDS_E_INVALID;
  *n = *ne = 0;
  digest[0] = digest[1] = edigest[0] = edigest[1] = 0;
  if (ctx->n < 2) return LADDS_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ScanCounters sc{};
  uint64_t entries = 0;
  unsigned long long saved = 0;
  HIPCHK(ctx, hipMemcpyAsync(&saved, ctx->log_count, sizeof(saved), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const int rc = run_detection_sync(ctx, dt_interp, true, &sc, &entries);
  HIPCHK(ctx, hipMemcpyAsync(ctx->log_count, &saved, sizeof(saved), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (rc != LADDS_OK) return rc;
  *n = sc.hits;
  *ne = sc.evaded;
  digest[0] = sc.digest[0];
  digest[1] = sc.digest[1];
  edigest[0] = sc.edigest[0];
  edigest[1] = sc.edigest[1];
  return LADDS_OK;
}

int ladds_simulate(ladds_ctx *ctx, uint64_t first_iteration, int nsteps, int tpcd, double min_altitude_km,
                   ladds_run_stats *stats) {
  if (!ctx || nsteps < 0 || tpcd < 1) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const bool digest = ctx->cfg.pair_output == LADDS_PAIRS_DIGEST;
  const double burnup_r = min_altitude_km > 0.0 ? min_altitude_km + kREarth : 0.0;
  const double burnup_r2 = burnup_r * burnup_r;  // Simulation.cpp:576,583
  {
    const int rc = alloc_env(ctx);
    if (rc != LADDS_OK) return rc;
  }
  // Step graphs (ladds_config.step_graph, detection every step): see capture_step_graph;
  // they alternate buffers 0 and 1.  Otherwise the eager loop cycles through three or four
  // buffers, overlapping propagation (stream), partition (stream3) and pair scans (stream2,
  // and stream4 with pair_streams 2) of consecutive steps: see overlapped_iteration.
  const bool use_graph = ctx->cfg.step_graph != 0 && tpcd == 1 && ctx->variant == kVariantPlaneTiles &&
                         !ctx->profiling && ctx->n > 0 && nsteps > 0;
  const bool use_overlap = !use_graph && ctx->variant == kVariantPlaneTiles && ctx->n > 0 && nsteps > 0;
  // The eager loop's first iteration runs on stream2 (overlapped_iteration's `fill`), and so
  // does what precedes it here (counter reset, spawn state, environments), after anything
  // still queued on the context's stream.
  hipStream_t head = ctx->stream;
  if (use_overlap) {
    head = ctx->stream2;
    const hipError_t q = hipStreamQuery(ctx->stream);  // (a busy stream's answer is no error)
    if (q == hipErrorNotReady) {
      const int rc = follow(ctx, ctx->ev_join, ctx->stream, head);
      if (rc != LADDS_OK) return rc;
    } else {
      HIPCHK(ctx, q);
    }
  }
  HIPCHK(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(ScanCounters), head));
  // Spawn protection on the device while objects with a parent id exist: every detection
  // accounts the iterations since the previous one (span), trailing iterations at the end.
  const bool spawn = ctx->protected_objects && ctx->n > 0 && nsteps > 0;
  uint64_t next_acc = first_iteration;  // first iteration whose start is not accounted yet
  if (spawn) {
    if (!ctx->d_spawn) HIPCHK(ctx, dalloc(&ctx->d_spawn, 1));
    ctx->h_rb->spawn = SpawnState{ctx->quiet, 0ull, 0ull, 0ull};
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_spawn, &ctx->h_rb->spawn, sizeof(SpawnState), hipMemcpyHostToDevice, head));
  }
  int cur = 0;                              // state buffer holding the latest state
  bool pending[4] = {false, false, false, false};  // a pair scan reading buffer b may still run
  bool fill = use_overlap;                  // the next overlapped iteration is the call's first
  PendingScan scan;                         // pair scan still to queue (overlapped loop)
  const bool two_streams = !spawn && pair_streams(ctx) == 2;
  // With two pair-scan streams a buffer's pair scan can still run beside the next step's,
  // so the propagation that reuses it comes one step later: four buffers.  Measured at 1M
  // objects (two streams): FP64 96.6 -> 93.3 us per step; split 69.7 -> 69.6 (20-step calls
  // 77.3 -> 77.7, within the spread of repeated runs).
  const int nbuf = two_streams ? 4 : 3;
  bool alt = false, used4 = false, last4 = false;  // next scan on stream4; stream4 used; last scan there
  if (use_overlap || use_graph) {
    const int rc = ensure_overlap_buffers(ctx, use_graph ? 2 : nbuf);
    if (rc != LADDS_OK) return rc;
  }
  hipGraphExec_t graphs[2] = {nullptr, nullptr};
  auto drop_graphs = [&] {
    for (hipGraphExec_t &g : graphs)
      if (g) hipGraphExecDestroy(g);
  };
  if (use_graph) {
    if (!ctx->d_step) HIPCHK(ctx, dalloc(&ctx->d_step, 1));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < 2; ++b) {
      const int rc = capture_step_graph(ctx, b, burnup_r2, digest, &graphs[b]);
      if (rc != LADDS_OK) {
        drop_graphs();
        return rc;
      }
    }
  }
  for (int seg0 = 0; seg0 < nsteps && ctx->n > 0; seg0 += kEnvSegment) {
    const int seg = std::min(kEnvSegment, nsteps - seg0);
    // environments of the whole segment: the staged segment when it covers these steps (an
    // earlier call's), else staged here in stream order after the previous segment's steps
    size_t env_off = 0;
    {
      const int rc = env_window(ctx, seg, &env_off, seg0 == 0 ? head : ctx->stream);
      if (rc != LADDS_OK) {
        drop_graphs();
        return rc;
      }
    }
    if (use_graph) {
      // prologue: step 0 of the segment propagated and partitioned eagerly; steps j = 0..seg-2
      // replay graph(buffer of step j): pair scan of step j beside step j+1's propagation;
      // epilogue: the pair scan of the segment's last step.
      const uint64_t base = first_iteration + (uint64_t)seg0;
      int rc = LADDS_OK;
      // the device step indexes the environments (from env_off) and labels the iteration
      // (iter_base + step, modulo 2^64)
      hipError_t e = launch_set_step(ctx->d_step, (uint32_t)env_off, base - (uint64_t)env_off, ctx->stream);
      if (e == hipSuccess) rc = launch_prop_part(ctx, cur, env_off, 1, true, burnup_r2, nullptr, 0);
      cur = 1 - cur;
      for (int j = 0; j + 1 < seg && rc == LADDS_OK && e == hipSuccess; ++j) {
        e = hipGraphLaunch(graphs[cur], ctx->stream);
        cur = 1 - cur;
      }
      if (rc == LADDS_OK && e == hipSuccess)
        rc = launch_pairs_of(ctx, cur, ctx->cfg.dt, digest, base + (uint64_t)(seg - 1), nullptr, ctx->stream,
                             spawn ? 1 : 0);
      if (rc == LADDS_OK && e != hipSuccess) rc = ctx->hipfail(e, "step graph launch");
      if (rc != LADDS_OK) {
        drop_graphs();
        return rc;
      }
      next_acc = base + (uint64_t)seg;
      ctx->step_count += (uint64_t)seg;
      continue;
    }
    int s = 0;
    while (s < seg) {
      // integrate iterations up to (and including) the next detection iteration
      int k = 1;
      uint64_t it = first_iteration + (uint64_t)(seg0 + s);
      while (s + k < seg && it % (uint64_t)tpcd != 0) {
        ++k;
        ++it;
      }
      const bool detect = (it % (uint64_t)tpcd) == 0;
      const uint64_t span = (spawn && detect) ? it - next_acc + 1 : 0;
      if (detect) next_acc = spawn ? it + 1 : next_acc;
      int rc;
      if (use_overlap) {
        const int nxt = (cur + 1) % nbuf;
        // the prologue ran on stream2: order this propagation after it
        rc = (fill && !detect) ? follow(ctx, ctx->ev_join, ctx->stream2, ctx->stream) : LADDS_OK;
        if (rc == LADDS_OK)
          rc = overlapped_iteration(ctx, nbuf, cur, pending[nxt], env_off + (size_t)s, k, detect, burnup_r2,
                                    ctx->cfg.dt * (double)tpcd, digest, it, span, fill);
        // the previous iteration's pair scan, then this one's (the first iteration's ran)
        if (rc == LADDS_OK && scan.valid) {
          rc = issue_scan(ctx, scan, digest, alt ? ctx->stream4 : ctx->stream2);
          used4 |= alt;
          last4 = alt;
          alt = two_streams && !alt;
        }
        scan.valid = rc == LADDS_OK && detect && !fill;
        scan.b = nxt;
        scan.dt_interp = ctx->cfg.dt * (double)tpcd;
        scan.iteration = it;
        scan.span = span;
        fill = false;
        if (rc == LADDS_OK) {
          pending[nxt] = detect;
          cur = nxt;
        }
      } else {
        rc = launch_propagate_steps(ctx, env_off + (size_t)s, k, detect, burnup_r2);
        if (rc == LADDS_OK && detect)
          rc = launch_detection(ctx, true, ctx->cfg.dt * (double)tpcd, digest, it, nullptr, span);
      }
      if (rc != LADDS_OK) return rc;
      s += k;
    }
    ctx->step_count += (uint64_t)seg;
  }
  if (scan.valid) {  // the last detection's pair scan
    const int rc = issue_scan(ctx, scan, digest, alt ? ctx->stream4 : ctx->stream2);
    if (rc != LADDS_OK) return rc;
    used4 |= alt;
    last4 = alt;
  }
  // join the pair-scan streams, account trailing iterations, read the counters back: one
  // synchronisation.  When the eager loop's last iteration detected, its pair scan follows
  // every launch of the call but the other pair stream's (its partition followed its
  // propagation, which followed the others on the context's stream): the readback goes on
  // that pair stream, after the other one.
  const uint64_t end_it = first_iteration + (uint64_t)nsteps;
  const bool trailing = spawn && end_it > next_acc;
  hipStream_t tail = ctx->stream;
  if (use_overlap && pending[cur] && !trailing) {
    tail = last4 ? ctx->stream4 : ctx->stream2;
    if (used4) {
      const int rc = follow(ctx, ctx->ev_join4, last4 ? ctx->stream2 : ctx->stream4, tail);
      if (rc != LADDS_OK) return rc;
    }
  } else if (use_overlap || use_graph) {
    int rc = follow(ctx, ctx->ev_join, ctx->stream2, ctx->stream);
    if (rc == LADDS_OK && used4) rc = follow(ctx, ctx->ev_join4, ctx->stream4, ctx->stream);
    if (rc != LADDS_OK) return rc;
  }
  if (trailing) {
    HIPCHK(ctx, spawn_pre(ctx, end_it - next_acc, ctx->stream));
    HIPCHK(ctx, spawn_post(ctx, end_it - next_acc, false, ctx->stream));
  }
  static_assert(sizeof(DeviceReadback) % 8 == 0, "whole words");
  HIPCHK(ctx, launch_publish(reinterpret_cast<const unsigned long long *>(ctx->counters),
                             reinterpret_cast<unsigned long long *>(&ctx->d_rb_host->sc),
                             (int)(sizeof(DeviceReadback) / 8), tail));
  if (spawn)
    HIPCHK(ctx, hipMemcpyAsync(&ctx->h_rb->spawn, ctx->d_spawn, sizeof(SpawnState), hipMemcpyDeviceToHost, tail));
  HIPCHK(ctx, hipStreamSynchronize(tail));
  drop_graphs();
  if ((use_overlap || use_graph) && cur != 0) {  // the final state is in another buffer: make it the context's own
    PipeBuffer &pb = ctx->pipe[cur];
    std::swap(ctx->d.rx, pb.r[0]);
    std::swap(ctx->d.ry, pb.r[1]);
    std::swap(ctx->d.rz, pb.r[2]);
    std::swap(ctx->d.vx, pb.v[0]);
    std::swap(ctx->d.vy, pb.v[1]);
    std::swap(ctx->d.vz, pb.v[2]);
    std::swap(ctx->d.activity, pb.act);
  }
  const ScanCounters sc = ctx->h_rb->sc;
  const unsigned long long cnt = ctx->h_rb->log_count;
  ctx->log_nonempty = cnt != 0;
  ctx->keys_valid = false;
  // spawn-protection count at the start of the next iteration (Simulation.cpp:399-402, 438)
  if (spawn) {
    ctx->quiet = ctx->h_rb->spawn.quiet;
    if (ctx->h_rb->spawn.cleared) ctx->protected_objects = false;
  } else if (sc.last_hit) {
    ctx->quiet = end_it - (sc.last_hit - 1ull);
  } else {
    ctx->quiet += (uint64_t)nsteps;
  }
  if (sc.overflow) return ctx->fail(LADDS_E_OVERFLOW, "conjunction scan scratch arena exhausted");
  // burnt-up objects were flagged by the propagation; compaction only when there are some
  if (min_altitude_km > 0.0 && sc.burnups > 0) {
    const int rc = compact_if_needed(ctx);
    if (rc != LADDS_OK) return rc;
  }
  if (stats) {
    stats->collisions = sc.hits;
    stats->evasions = sc.evaded;
    stats->logged = std::min<uint64_t>(cnt, ctx->log_cap);
    stats->candidates = sc.candidates;
    stats->digest[0] = sc.digest[0];
    stats->digest[1] = sc.digest[1];
    stats->edigest[0] = sc.edigest[0];
    stats->edigest[1] = sc.edigest[1];
    stats->remaining = ctx->n;
  }
  return LADDS_OK;
}

int ladds_remove_spawn_protection(ladds_ctx *ctx) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_clear_parents(nullptr, 1, ctx->d.parent, ctx->n, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->protected_objects = false;
  return LADDS_OK;
}

int ladds_fetch_log(ladds_ctx *ctx, ladds_logged_pair *out, size_t cap, size_t *n) {
  if (!ctx || !n || (cap && !out)) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned long long cnt = 0;
  HIPCHK(ctx, hipMemcpyAsync(&cnt, ctx->log_count, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const uint64_t stored = std::min<uint64_t>(cnt, ctx->log_cap);
  *n = (size_t)cnt;
  const size_t m = std::min<size_t>(stored, cap);
  if (m) HIPCHK(ctx, hipMemcpyAsync(out, ctx->log, sizeof(LoggedPair) * m, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (cnt > ctx->log_cap) return ctx->fail(LADDS_E_OVERFLOW, "device pair log overflowed (log_capacity)");
  if (cnt > cap) return ctx->fail(LADDS_E_OVERFLOW, "output capacity too small");
  return LADDS_OK;
}

}  // extern "C"


extern "C" {

int ladds_particle_records(ladds_ctx *ctx, float *pos, float *vel, uint32_t *ids) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  const size_t n = ctx->n;
  if (n && (!pos || !vel || !ids)) return LADDS_E_INVALID;
  if (n == 0) return LADDS_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t fb = sizeof(float) * 3 * n;
  int rc = ensure_xbuf(ctx, 2 * fb + sizeof(uint32_t) * n);
  if (rc != LADDS_OK) return rc;
  float *dpos = static_cast<float *>(ctx->xbuf), *dvel = dpos + 3 * n;
  uint32_t *dids = reinterpret_cast<uint32_t *>(dvel + 3 * n);
  HIPCHK(ctx, timed(ctx, "particle_records",
                    [&] { return launch_particle_records(ctx->d, n, dpos, dvel, dids, ctx->stream); }));
  HIPCHK(ctx, hipMemcpyAsync(pos, dpos, fb, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(vel, dvel, fb, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ids, dids, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return LADDS_OK;
}

int ladds_log_records(ladds_ctx *ctx, uint64_t iteration, ladds_pair_record *hits, size_t cap, size_t *n,
                      ladds_pair_record *evaded, size_t ecap, size_t *ne) {
  if (!ctx || !n || !ne || (cap && !hits) || (ecap && !evaded)) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t rb = sizeof(PairRecord32);
  int rc = ensure_xbuf(ctx, std::max<size_t>(rb * (cap + ecap), 64));
  if (rc != LADDS_OK) return rc;
  PairRecord32 *dh = static_cast<PairRecord32 *>(ctx->xbuf), *de = dh + cap;
  HIPCHK(ctx, hipMemsetAsync(ctx->xcount, 0, 2 * sizeof(unsigned long long), ctx->stream));
  HIPCHK(ctx, timed(ctx, "log_records", [&] {
    return launch_log_records(ctx->log, ctx->log_count, ctx->log_cap, iteration, dh, cap, de, ecap, ctx->xcount,
                              ctx->stream);
  }));
  unsigned long long cnt[2] = {0, 0};
  HIPCHK(ctx, hipMemcpyAsync(cnt, ctx->xcount, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *n = (size_t)cnt[0];
  *ne = (size_t)cnt[1];
  const size_t mh = std::min<size_t>(cnt[0], cap), me = std::min<size_t>(cnt[1], ecap);
  if (mh) HIPCHK(ctx, hipMemcpyAsync(hits, dh, rb * mh, hipMemcpyDeviceToHost, ctx->stream));
  if (me) HIPCHK(ctx, hipMemcpyAsync(evaded, de, rb * me, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (cnt[0] > cap || cnt[1] > ecap) return ctx->fail(LADDS_E_OVERFLOW, "output capacity too small");
  return LADDS_OK;
}

int ladds_clear_log(ladds_ctx *ctx) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMemsetAsync(ctx->log_count, 0, sizeof(unsigned long long), ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->log_nonempty = false;
  return LADDS_OK;
}

int ladds_delete_burnups(ladds_ctx *ctx, double min_altitude_km, size_t *removed) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t before = ctx->n;
  if (before) {
    // mark, then compact: a zero-step propagation applies only the burn-up test
    const double r = min_altitude_km + kREarth;
    std::vector<double> hx(before), hy(before), hz(before);
    std::vector<int32_t> act(before);
    HIPCHK(ctx, hipMemcpyAsync(hx.data(), ctx->d.rx, 8 * before, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(hy.data(), ctx->d.ry, 8 * before, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(hz.data(), ctx->d.rz, 8 * before, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(act.data(), ctx->d.activity, 4 * before, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const double r2 = r * r;
    for (size_t i = 0; i < before; ++i) {
      double s = 0.0;
      s += hx[i] * hx[i];
      s += hy[i] * hy[i];
      s += hz[i] * hz[i];
      if (s < r2) act[i] = -1;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d.activity, act.data(), 4 * before, hipMemcpyHostToDevice, ctx->stream));
    const int rc = compact_if_needed(ctx);
    if (rc != LADDS_OK) return rc;
  }
  if (removed) *removed = before - ctx->n;
  return LADDS_OK;
}

int ladds_delete_objects(ladds_ctx *ctx, const uint64_t *ids, size_t count, size_t *removed) {
  if (!ctx || (count && !ids)) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t before = ctx->n;
  if (before && count) {
    std::unordered_set<uint64_t> doomed(ids, ids + count);
    std::vector<uint64_t> id(before);
    std::vector<int32_t> act(before);
    HIPCHK(ctx, hipMemcpyAsync(id.data(), ctx->d.id, 8 * before, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(act.data(), ctx->d.activity, 4 * before, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    bool any = false;
    for (size_t i = 0; i < before; ++i)
      if (doomed.count(id[i])) {
        act[i] = -1;
        any = true;
      }
    if (any) {
      HIPCHK(ctx, hipMemcpyAsync(ctx->d.activity, act.data(), 4 * before, hipMemcpyHostToDevice, ctx->stream));
      const int rc = compact_if_needed(ctx);
      if (rc != LADDS_OK) return rc;
    }
  }
  if (removed) *removed = before - ctx->n;
  return LADDS_OK;
}

int ladds_set_profiling(ladds_ctx *ctx, int enable) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  resolve_profiles(ctx);
  ctx->profiling = enable != 0;
  for (auto &kv : ctx->prof) {
    kv.second.launches = 0;
    kv.second.ms = 0.0;
  }
  return LADDS_OK;
}

int ladds_kernel_stats(ladds_ctx *ctx, const char *name, uint64_t *launches, double *total_ms) {
  if (!ctx || !name || !launches || !total_ms) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  bool known = false;
  for (const char *k : kKernelNames) known |= std::strcmp(k, name) == 0;
  if (!known) return ctx->fail(LADDS_E_INVALID, std::string("unknown kernel name ") + name);
  resolve_profiles(ctx);
  auto it = ctx->prof.find(name);
  *launches = it == ctx->prof.end() ? 0 : it->second.launches;
  *total_ms = it == ctx->prof.end() ? 0.0 : it->second.ms;
  return LADDS_OK;
}

int ladds_synchronize(ladds_ctx *ctx) {
  if (!ctx) return LADDS_E_INVALID;
  LADDS_SETTLE(ctx);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return LADDS_OK;
}

void *ladds_stream(ladds_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

}  // extern "C"

/*
 * ladds_amd.h -- C ABI of the MI355X (gfx950) propagate + conjunction engine.
 *
 * This is the drop-in boundary for LADDS's per-timestep hot path.  Every entry
 * point lists the reference interface it replaces (paths relative to the LADDS
 * tree; OrbitPropagator and AutoPas are the pinned upstream dependencies that
 * LADDS fetches at build time):
 *
 *   YoshidaIntegrator<AutoPas_t>(autopas, accumulator, deltaT)    src/ladds/Simulation.cpp:145
 *   YoshidaIntegrator::integrate(false)                             src/ladds/Simulation.cpp:297
 *   AccelerationAccumulator<AutoPas_t>(flags[8], autopas, 0.0, w)   src/ladds/Simulation.cpp:127-143
 *   AutoPas<Particle>::iteratePairwise(&CollisionFunctor)            src/ladds/Simulation.cpp:220
 *   CollisionFunctor(cutoff, dt, cdf, minDetRadius, evasionCutoff)  src/ladds/CollisionFunctor.cpp:18-30
 *   CollisionFunctor::getCollisions / getEvadedCollisions           src/ladds/CollisionFunctor.cpp:32-38
 *   Simulation::simulationLoop (integrate + collision detection)    src/ladds/Simulation.cpp:287-415
 *   Simulation::deleteBurnUps                                       src/ladds/Simulation.cpp:561-596
 *
 * Conventions
 *   - Units follow LADDS: km, km/s, s; radius in m; aom in km^2/kg; bc_inv in m^2/kg.
 *   - Time t is seconds since J2000.0; the accumulator start time of the reference
 *     (Simulation.cpp:143, 0.0) is ladds_config.t0.
 *   - All entry points return LADDS_OK (0) or a negative status.  No C++ exception
 *     crosses this boundary; ladds_last_error() holds the message.
 *   - A context owns its device memory and one HIP stream, and is bound to one host
 *     thread at a time (not re-entrant), as the reference's single-threaded driver.
 *   - Pairs carry ids and object indices (positions in the array last passed to
 *     ladds_set_objects, after any burn-up compaction) instead of Particle*; a host
 *     adapter maps indices back to its own objects.
 */
#ifndef LADDS_AMD_H
#define LADDS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LADDS_ABI_VERSION 1

#if defined(__GNUC__)
#define LADDS_API __attribute__((visibility("default")))
#else
#define LADDS_API
#endif

/* ---- status codes ---- */
#define LADDS_OK 0
#define LADDS_E_INVALID (-1)   /* bad argument                                       */
#define LADDS_E_HIP (-2)       /* HIP runtime error                                  */
#define LADDS_E_OVERFLOW (-3)  /* output capacity too small; *n holds required count */
#define LADDS_E_NOMEM (-4)     /* device allocation failed                           */
#define LADDS_E_STATE (-5)     /* call out of order (e.g. no objects set)            */
#define LADDS_E_NODEVICE (-6)  /* no usable gfx950 device                            */

/* ---- force components, bit order of the flag array at src/ladds/Simulation.cpp:127-135 ---- */
#define LADDS_KEP (1u << 0)
#define LADDS_J2 (1u << 1)
#define LADDS_C22 (1u << 2)
#define LADDS_S22 (1u << 3)
#define LADDS_SOL (1u << 4)
#define LADDS_LUN (1u << 5)
#define LADDS_SRP (1u << 6)
#define LADDS_DRAG (1u << 7)
#define LADDS_ALL_COMPONENTS 0xFFu

/* ---- precision forms ---- */
#define LADDS_FP64 0       /* every stage in FP64                                                   */
#define LADDS_SPLIT 1      /* FP64: state, Kepler, J2, C22/S22, drag, closest approach;              */
                           /* FP32: Sun, Moon, SRP and the conservative conjunction prefilter       */
#define LADDS_SPLIT_ALL32 2 /* as LADDS_SPLIT with J2 and C22/S22 also in FP32 (larger error)       */

/* ---- conjunction threshold modes ---- */
#define LADDS_THRESHOLD_RADIUS_SCALED 0 /* reference: d^2 <= ((R_i+R_j) cdf / 1000)^2, CollisionFunctor.cpp:85,112 */
#define LADDS_THRESHOLD_ABSOLUTE 1      /* d^2 <= threshold_km^2 (no reference counterpart) */

/* ---- ActivityState, src/ladds/particle/Particle.h:37-53 ---- */
#define LADDS_PASSIVE 0
#define LADDS_EVASIVE 1
#define LADDS_EVASIVE_PRESERVING 2

/* ---- pair-scan output modes ---- */
#define LADDS_PAIRS_LIST 0   /* write pair records                                  */
#define LADDS_PAIRS_DIGEST 1 /* count + order-independent 128-bit digest only       */

typedef struct ladds_ctx ladds_ctx;

typedef struct ladds_config {
  double dt;                         /* sim/deltaT [s]                                        */
  double t0;                         /* accumulator start time [s since J2000]                */
  uint32_t components;               /* sim/prop/use{KEP..DRAG}Component bits                  */
  int32_t precision;                 /* LADDS_FP64 | LADDS_SPLIT | LADDS_SPLIT_ALL32          */
  double cutoff_km;                  /* autopas/cutoff                                        */
  double collision_distance_factor;  /* sim/collisionDistanceFactor                           */
  double min_detection_radius_m;     /* sim/minDetectionRadius                                */
  double evasion_cutoff_km;          /* sim/evasionTrackingCutoffInKM                         */
  int32_t threshold_mode;            /* LADDS_THRESHOLD_*                                     */
  double threshold_km;               /* absolute mode threshold                               */
  double cell_size_km;               /* 0: 2 * cutoff (minimum allowed)                        */
  int32_t scan_variant;              /* 0: auto/autotuned, 1: global-memory scan, 2: LDS tiles */
  int32_t block_size;                /* plane-tile pair kernel workgroup: 128/256/384/512/1024, else auto */
  int32_t device;                    /* HIP device ordinal                                    */
  int32_t pair_output;               /* ladds_simulate: LADDS_PAIRS_LIST | LADDS_PAIRS_DIGEST  */
  uint64_t log_capacity;             /* pair log entries kept on device by ladds_simulate     */
  int32_t step_graph;                /* ladds_simulate with tpcd 1: replay each step from one  */
                                     /* captured HIP graph (plane-tile scan, no profiling)     */
  int32_t scan_chunks;               /* plane-tile scan: most workgroups per plane, 1-32 (0: auto) */
  int32_t tune_simulate;             /* ladds_autotune times each shape over 48 ladds_simulate    */
                                     /* steps (state, time, pair log and counters restored after), */
                                     /* as AutoPas times real iterations; 0: one detection alone   */
} ladds_config;

/* Object arrays, structure-of-arrays, host memory; copied in (caller keeps ownership). */
typedef struct ladds_objects {
  size_t n;
  const double *rx, *ry, *rz;   /* position [km]                                  */
  const double *vx, *vy, *vz;   /* velocity [km/s]                                */
  const double *radius_m;       /* Particle::_radius [m]                          */
  const double *aom;            /* Particle::_aom [km^2/kg]   (NULL: 0)           */
  const double *bc_inv;         /* Particle::_bc_inv [m^2/kg] (NULL: 0)           */
  const double *cr;             /* radiation pressure coefficient (NULL: 1.0)     */
  const int32_t *activity;      /* ActivityState (NULL: passive)                  */
  const uint64_t *id;           /* Particle id (NULL: index)                      */
  const uint64_t *parent;       /* parentIdentifier (NULL: UINT64_MAX)            */
} ladds_objects;

/* One conjunction: ids ordered as CollisionFunctor.cpp:94-128 (smaller id first). */
typedef struct ladds_pair {
  uint64_t id_a, id_b;
  uint32_t idx_a, idx_b;
  double d2;      /* squared closest-approach distance [km^2] */
  double mid[3];  /* collision point p2 + dr/2 [km]          */
} ladds_pair;     /* 56 bytes */

/* A pair recorded by ladds_simulate, tagged with its iteration. */
typedef struct ladds_logged_pair {
  ladds_pair pair;
  uint64_t iteration;
  uint32_t evaded;   /* 0 collision, 1 evaded conjunction      */
  uint32_t scenario; /* ensemble scenario (0 without scenarios) */
} ladds_logged_pair; /* 72 bytes */

/* HDF5Definitions::CollisionInfo / EvasionInfo (src/ladds/io/hdf5/HDF5Definitions.h:41-64):
 * the record HDF5Writer::writeConjunctions stores per pair. */
typedef struct {
  uint32_t id_a, id_b;
  float d2;
} ladds_pair_record; /* 12 bytes */

typedef struct ladds_run_stats {
  uint64_t collisions;      /* total collisions found (also beyond log capacity) */
  uint64_t evasions;        /* total evaded conjunctions                          */
  uint64_t logged;          /* entries stored in the device log                  */
  uint64_t candidates;      /* pairs that passed the cutoff prefilter             */
  uint64_t digest[2];       /* order-independent digest of collision id pairs     */
  uint64_t edigest[2];      /* same for evaded conjunctions                        */
  uint64_t remaining;       /* objects left after burn-up compaction               */
} ladds_run_stats;

/* ---- library ---- */
LADDS_API const char *ladds_version(void);
LADDS_API int ladds_abi_version(void);
LADDS_API int ladds_device_count(int *count);
LADDS_API void ladds_config_default(ladds_config *cfg);

/* ---- context: replaces the (csvWriter, accumulator, integrator) tuple built by
 *      Simulation::initIntegrator (Simulation.cpp:122-148) plus the AutoPas container
 *      built by Simulation::initAutoPas (Simulation.cpp:59-120) ---- */
LADDS_API int ladds_create(const ladds_config *cfg, ladds_ctx **out);
LADDS_API int ladds_destroy(ladds_ctx *ctx);
/* Message of the context's last failure; with NULL, that of the last failed
 * ladds_conjunction_digest_noalloc call on this thread. */
LADDS_API const char *ladds_last_error(const ladds_ctx *ctx);
/* Update the parameters the integrator, accumulator and CollisionFunctor own (dt, t0,
 * components, functor thresholds, pair output, log capacity) -- the reference sets them in
 * the constructors of YoshidaIntegrator (Simulation.cpp:145), AccelerationAccumulator
 * (:142-143) and CollisionFunctor (CollisionFunctor.cpp:18-30).  Changing t0 or dt restarts
 * the step count at t0.  Fields fixed at ladds_create (precision, cutoff_km, cell_size_km,
 * scan_variant, device) must be unchanged once objects are set (LADDS_E_STATE otherwise). */
LADDS_API int ladds_configure(ladds_ctx *ctx, const ladds_config *cfg);

/* AutoPas::addParticle for a whole population (SatelliteLoader.cpp:17-84). Replaces all objects. */
LADDS_API int ladds_set_objects(ladds_ctx *ctx, const ladds_objects *objs);
LADDS_API size_t ladds_num_objects(const ladds_ctx *ctx);
/* Breakup hook, part 1: append objects to the current population, keeping its state --
 * BreakupWrapper.cpp:38-43 (particle/) adds each fragment with
 * autopas.addParticle(..., parentID).  Fragments of one breakup share a parent id, and the
 * predicate skips pairs with equal parent ids (CollisionFunctor.cpp:65-69).  Ids default
 * to the object's index.  Not for ensembles (LADDS_E_STATE). */
LADDS_API int ladds_add_objects(ladds_ctx *ctx, const ladds_objects *objs);
/* Breakup hook, part 2: remove the objects with the given ids (order of the rest kept) --
 * autopas.deleteParticle of the two colliding objects, BreakupWrapper.cpp:23-25. */
LADDS_API int ladds_delete_objects(ladds_ctx *ctx, const uint64_t *ids, size_t count, size_t *removed);

/* Batched ensembles (no LADDS counterpart; LADDS runs one scenario per process): the
 * objects last set are the concatenation of independent scenarios, scenario s occupying
 * indices [offsets[s], offsets[s+1]) (nscenarios + 1 entries, offsets[0] = 0,
 * offsets[nscenarios] = ladds_num_objects()).  Each scenario gets its own cell grid;
 * objects of different scenarios never form pairs; logged pairs carry their scenario.
 * Burn-up compaction keeps the scenarios (ladds_get_scenarios gives the new offsets).
 * ladds_set_objects resets to a single scenario.  Needs the plane-tile scan. */
LADDS_API int ladds_set_scenarios(ladds_ctx *ctx, size_t nscenarios, const uint64_t *offsets);
LADDS_API size_t ladds_num_scenarios(const ladds_ctx *ctx);
LADDS_API int ladds_get_scenarios(const ladds_ctx *ctx, uint64_t *offsets);
/* Read back state (caller-owned arrays of ladds_num_objects() entries; NULL skips). */
LADDS_API int ladds_get_state(ladds_ctx *ctx, double *rx, double *ry, double *rz, double *vx, double *vy, double *vz);
/* Ids of the current objects, in index order (after compaction). */
LADDS_API int ladds_get_ids(ladds_ctx *ctx, uint64_t *id);
LADDS_API double ladds_get_time(const ladds_ctx *ctx);
LADDS_API int ladds_set_time(ladds_ctx *ctx, double t);

/* YoshidaIntegrator::integrate(false), nsteps times (Simulation.cpp:297).  Enqueues the
 * steps on the context's stream and returns without waiting: a detection that follows
 * (ladds_conjunctions, ladds_conjunction_digest) is ordered after them on that stream, and
 * every call that reads or writes object data from the host waits for them first (as does
 * ladds_synchronize).  The Sun/Moon/Earth-rotation environment of 8,192 steps is staged at
 * once and reused by the calls that follow. */
LADDS_API int ladds_propagate(ladds_ctx *ctx, int nsteps);

/* Simulation::collisionDetection (Simulation.cpp:211-222): AutoPas::iteratePairwise with a
 * CollisionFunctor(cutoff, dt_interp, cdf, minDetectionRadius, evasionCutoff) on the current
 * state.  Unordered.  On overflow returns LADDS_E_OVERFLOW with *n / *ne = required counts. */
LADDS_API int ladds_conjunctions(ladds_ctx *ctx, double dt_interp, ladds_pair *hits, size_t cap, size_t *n,
                       ladds_pair *evaded, size_t ecap, size_t *ne);

/* State of the objects at device indices idx[0..count) (ladds_pair.idx_a / idx_b), as
 * r[3 * k + c] and v[3 * k + c] -- what CollisionFunctor's (Particle*, Particle*, ...) tuples
 * point at (CollisionFunctor.h:95-108), without downloading the whole population. */
LADDS_API int ladds_gather_state(ladds_ctx *ctx, const uint32_t *idx, size_t count, double *r, double *v);

/* Count + order-independent 128-bit digest of the collision and evasion sets on the current
 * state, without materialising pairs (for populations whose pair sets do not fit memory). */
LADDS_API int ladds_conjunction_digest(ladds_ctx *ctx, double dt_interp, uint64_t *n, uint64_t digest[2], uint64_t *ne,
                             uint64_t edigest[2]);

/* Simulation::simulationLoop hot path (Simulation.cpp:287-415) for iterations
 * [first_iteration, first_iteration + nsteps): integrate; deleteBurnUps if
 * min_altitude_km > 0 (Simulation.cpp:301-303); collision detection when
 * iteration % tpcd == 0 with window dt*tpcd (Simulation.cpp:389-397).  Pairs are
 * appended to the device log (ladds_fetch_log); totals in *stats (may be NULL).
 * Spawn protection as the loop keeps it (Simulation.cpp:287-293, 399-402, 438): the
 * context counts iterations since the last one with collisions (from 0 at
 * ladds_create, continued across calls and uploads); at the start of an iteration where the
 * count is 100, every parent id is reset (removeParticleSpawnProtection, :470-473). */
LADDS_API int ladds_simulate(ladds_ctx *ctx, uint64_t first_iteration, int nsteps, int tpcd, double min_altitude_km,
                   ladds_run_stats *stats);
/* Simulation::removeParticleSpawnProtection (Simulation.cpp:470-473): every parent id set
 * to "none" (UINT64_MAX), so fragments of one breakup can conjunct again -- for callers
 * that drive the loop themselves (integrate + ladds_conjunctions). */
LADDS_API int ladds_remove_spawn_protection(ladds_ctx *ctx);
LADDS_API int ladds_fetch_log(ladds_ctx *ctx, ladds_logged_pair *out, size_t cap, size_t *n);
LADDS_API int ladds_clear_log(ladds_ctx *ctx);

/* Output records, packed on the device so that only narrow data crosses PCIe.
 * ladds_particle_records: the payload of HDF5Writer::writeParticles
 * (io/hdf5/HDF5Writer.cpp:89-149): float32 xyz positions and velocities (3 * n floats
 * each, object order) and uint32 ids of the current objects.
 * ladds_log_records: the payload of HDF5Writer::writeConjunctions / writeEvasions
 * (HDF5Writer.cpp:151-187) for one iteration of the device log, unordered; on overflow
 * LADDS_E_OVERFLOW with *n / *ne = required counts.  Casts as the reference's
 * static_casts: double -> float to nearest, ids keep their low 32 bits. */
LADDS_API int ladds_particle_records(ladds_ctx *ctx, float *pos, float *vel, uint32_t *ids);
LADDS_API int ladds_log_records(ladds_ctx *ctx, uint64_t iteration, ladds_pair_record *hits, size_t cap, size_t *n,
                                ladds_pair_record *evaded, size_t ecap, size_t *ne);

/* Simulation::deleteBurnUps (Simulation.cpp:561-596): remove objects with |r| < R_E + h_min,
 * compacting the arrays on the device (order preserved). */
LADDS_API int ladds_delete_burnups(ladds_ctx *ctx, double min_altitude_km, size_t *removed);

/* ---- autotuning ----
 * The counterpart of AutoPas's tuning phase (autopas/tuning, driven by the
 * iteratePairwise calls of Simulation.cpp:211-222, whose return value reports
 * "still tuning"): time the plane-tile conjunction scan on the current state for
 * candidate workgroups per plane (1-32, for planes dense with clusters), pair-kernel
 * workgroup sizes, cell sizes (1, 1.25, 1.5, 2 x the search radius), plane widths (1, 1.5,
 * 2, 3 x the cell size), plane shapes (log2 buckets per x-plane around the
 * population-derived default), partition tile sizes (4096, 8192, 16384 objects) and the
 * pair kernel's LDS per plane pair (sized for 2-5 workgroups per CU, or automatic), one
 * dimension at a time (coordinate descent), and keep the fastest for this context.  Pair
 * results do not depend on the choice. */
typedef struct ladds_tuning {
  double cell_size_km;   /* chosen cell side [km]                               */
  int32_t plane_bits;    /* chosen log2 buckets per x-plane                     */
  int32_t pair_threads;  /* chosen pair-kernel workgroup size                   */
  double scan_us;        /* device time of one detection with the choice [us]   */
                         /* (ladds_config.tune_simulate: wall time per step)    */
  double default_us;     /* the same for the untuned default [us]               */
  int32_t tried;         /* configurations timed                                */
  int32_t tile_size;     /* chosen objects per partition tile                   */
  int32_t chunks;        /* chosen workgroups per plane (dense planes)          */
  int32_t workgroups_per_cu; /* pair-kernel workgroups per CU its LDS is sized for */
  double plane_width_km; /* chosen cell length along x (plane width) [km]       */
  int32_t plane_order;   /* pair-kernel dispatch: 0 plane order, 1 heaviest planes first */
  int32_t ensemble_plane_objects; /* ensembles: objects per x-plane (0: the default 64)   */
  int32_t row_cells;     /* bucket row height along y in cells: 1 (an object scans three */
                         /* rows of the next plane) or 2 (two rows, on the side of the   */
                         /* half of its row it lies in); 0 is read as 1                  */
  int32_t bucket_cells;  /* bucket width along z in cells: 1 (three buckets of a row) or  */
                         /* 2 (two, on the side of the half of its bucket); 0 is read as 1 */
  int32_t persistent;    /* pair kernel: 0 one workgroup per plane pair; 1 as many         */
                         /* workgroups as the CUs hold, taking plane pairs from a queue   */
                         /* per XCD until none is left                                     */
  int32_t pair_streams;  /* ladds_simulate's pair scans: 1 stream (each starts after the  */
                         /* previous ends) or 2 alternating (one starts while the other  */
                         /* drains; 1 while spawn protection is active); 0: 2 for the    */
                         /* split forms, 1 for FP64.  Tuned only with tune_simulate       */
} ladds_tuning;

LADDS_API int ladds_autotune(ladds_ctx *ctx, double dt_interp, int repetitions, ladds_tuning *out);
/* Apply a shape that ladds_autotune chose earlier (cell_size_km, plane_width_km,
 * plane_bits, pair_threads, tile_size, chunks, workgroups_per_cu, plane_order,
 * ensemble_plane_objects, row_cells, bucket_cells, persistent, pair_streams; the timing fields are
 * ignored) without timing anything -- the counterpart of running LADDS with the
 * calibrated AutoPas configuration that tuning mode writes out (Simulation.cpp:404-407,
 * dumpCalibratedConfig).  A value outside the tuner's range is LADDS_E_INVALID. */
LADDS_API int ladds_set_tuning(ladds_ctx *ctx, const ladds_tuning *tuning);

/* ---- measurement ---- */
/* Record HIP events around every kernel launch; ladds_kernel_stats reports per kernel kind. */
LADDS_API int ladds_set_profiling(ladds_ctx *ctx, int enable);
/* name: "propagate", "bin_reduce", "bin_scan", "bin_scatter", "pair_scan"; returns launches and
 * summed device milliseconds since profiling was enabled (synchronises). */
LADDS_API int ladds_kernel_stats(ladds_ctx *ctx, const char *name, uint64_t *launches, double *total_ms);
LADDS_API int ladds_synchronize(ladds_ctx *ctx);
/* HIP stream the context launches on (hipStream_t as void*). */
LADDS_API void *ladds_stream(ladds_ctx *ctx);

/* ---- differentiable propagation (FP64): d(final state)/d(initial state, bc_inv) ----
 * No LADDS counterpart: the reference propagates forward only (YoshidaIntegrator::integrate,
 * Simulation.cpp:297).  The forward pass is the FP64 step of ladds_propagate; the backward
 * pass is a hand-written adjoint of the Yoshida-4 step with checkpoints every
 * checkpoint_every steps.  All array arguments are DEVICE pointers on the plan's device;
 * vectors are [3][n] blocks (x row, y row, z row); stream is a hipStream_t (NULL: default).
 * aom, bc_inv, cr, activity may be NULL (0, 0, 1, passive); g_rT / g_vT may be NULL (zero);
 * g_bc may be NULL (not computed).  The caller owns the checkpoint buffer
 * (ladds_adjoint_checkpoint_doubles() doubles) between forward and backward; the plan owns
 * a workspace, so calls on one plan must not overlap. */
typedef struct ladds_adjoint ladds_adjoint;
typedef struct ladds_adjoint_config {
  uint64_t n;               /* objects                                      */
  int32_t nsteps;           /* Yoshida steps of one forward pass            */
  int32_t checkpoint_every; /* steps between stored states (0: 64)          */
  double dt;                /* step [s]                                     */
  double t0;                /* start time [s since J2000]                   */
  uint32_t components;      /* LADDS_KEP..LADDS_DRAG bits                   */
  int32_t device;           /* HIP device ordinal                           */
} ladds_adjoint_config;

LADDS_API int ladds_adjoint_create(const ladds_adjoint_config *cfg, ladds_adjoint **out);
LADDS_API int ladds_adjoint_destroy(ladds_adjoint *plan);
LADDS_API const char *ladds_adjoint_last_error(const ladds_adjoint *plan);
LADDS_API uint64_t ladds_adjoint_checkpoint_doubles(const ladds_adjoint *plan);
LADDS_API int ladds_adjoint_forward(ladds_adjoint *plan, const double *r0, const double *v0, const double *aom,
                                    const double *bc_inv, const double *cr, const int32_t *activity, double *rT,
                                    double *vT, double *checkpoints, void *stream);
LADDS_API int ladds_adjoint_backward(ladds_adjoint *plan, const double *aom, const double *bc_inv, const double *cr,
                                     const int32_t *activity, const double *checkpoints, const double *g_rT,
                                     const double *g_vT, double *g_r0, double *g_v0, double *g_bc, void *stream);

/* ---- conjunction digest without allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the count
 * and digest of ladds_conjunction_digest for objects in the caller's device memory, with
 * the scan's scratch (cell keys, sorted records, tile offsets, arena, counters) in one
 * static device array of each GPU, sized for at most LADDS_NOALLOC_MAX_OBJECTS objects
 * (about 6.5 MB, held until the process ends).  Calls no hipMalloc / hipFree.  Uses
 * cfg->cutoff_km, precision, threshold_mode, threshold_km, collision_distance_factor,
 * min_detection_radius_m, evasion_cutoff_km and device; runs on `stream` (a hipStream_t,
 * NULL: the default stream) and waits for it.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one device race.  More objects than the maximum:
 * LADDS_E_INVALID, message from ladds_last_error(NULL). */
#define LADDS_NOALLOC_MAX_OBJECTS 65536
typedef struct ladds_device_objects {
  size_t n;
  const double *rx, *ry, *rz;   /* device pointers, as ladds_objects       */
  const double *vx, *vy, *vz;
  const double *radius_m;
  const int32_t *activity;
  const uint64_t *id, *parent;
} ladds_device_objects;
LADDS_API int ladds_conjunction_digest_noalloc(const ladds_device_objects *objs, const ladds_config *cfg,
                                               double dt_interp, uint64_t *n, uint64_t digest[2], uint64_t *ne,
                                               uint64_t edigest[2], void *stream);

/* ---- seeded synthetic populations (host only, no device needed) ----
 * workload 0: Keplerian elements as the 16,384-object case (perigee altitude U[200,2000] km,
 *             e U[0,0.02], inclination mixture, radius log-U[0.05,5] m, bc_inv from c_D = 2.2);
 * workload 1: as 0 plus a 1-cm debris tail (90% radius log-U[0.005,0.1] m), bc_inv U[0.005,0.05],
 *             Cr U[1.0,1.5];
 * workload 2: 60% constellation shells replicated with +-50 m jitter, 40% as workload 0.
 * planted_pairs > 0 appends co-orbital / crossing pairs 0.3-0.9 km apart (for tests). */
LADDS_API int ladds_generate_population(int workload, uint64_t seed, size_t n, size_t planted_pairs, double *rx, double *ry,
                              double *rz, double *vx, double *vy, double *vz, double *radius_m, double *mass_kg,
                              double *aom, double *bc_inv, double *cr);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oracle.h -- CPU restatement of the LADDS propagate + conjunction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  Nothing in the product
 * (ladds_amd/) links, loads or calls it.
 *
 * Plain C11, FP64 throughout, compiled with -ffp-contract=off and without
 * fast-math so every operation is a single IEEE-754 rounding, matching the
 * reference build (no FMA contraction).  -O3 -march=x86-64-v3 vectorises without
 * changing any result: without fast-math the compiler may not reassociate.
 *
 * What each part follows (reference paths relative to the LADDS tree):
 *   or_distance_by_linear_interpolation  src/ladds/utils/DistanceApproximation.cpp:13-58
 *   or_collision_functor                 src/ladds/CollisionFunctor.cpp:40-131 (+ ctor :18-30)
 *   or_pairs_bruteforce / or_pairs_grid  AutoPas::iteratePairwise + Newton3 traversal as used at
 *                                        src/ladds/Simulation.cpp:211-222; pinned by the container-
 *                                        independence invariant tests/testladds/CollisionFunctorIntegrationTest.cpp:18-135
 *   or_yoshida_steps                     YoshidaIntegrator::integrate (OrbitPropagator @ b90f3bae, not in
 *                                        the tree; used at src/ladds/Simulation.cpp:145,297).  Restated from
 *                                        H. Yoshida, Phys. Lett. A 150 (1990) 262, 4th-order composition.
 *   or_yoshida_vjp                       gradient of or_yoshida_steps by complex-step differentiation
 *                                        of the same arithmetic (gradient.c): the reference for the
 *                                        engine's adjoint, which has no LADDS counterpart
 *   or_acceleration                      AccelerationAccumulator (OrbitPropagator @ b90f3bae; flags built at
 *                                        src/ladds/Simulation.cpp:127-135).  Components restated from
 *                                        Montenbruck & Gill, Satellite Orbits (2000) Sec. 3.2-3.5 and the
 *                                        exponential atmosphere of Vallado, Fundamentals of Astrodynamics
 *                                        and Applications (4th ed. 2013) Table 8-4.
 *
 * Parity status (what pins each function) is listed per function below.
 */
#ifndef LADDS_ORACLE_H
#define LADDS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- component bits, in the order of src/ladds/Simulation.cpp:127-135 ---- */
enum {
  OR_KEP = 1u << 0,
  OR_J2 = 1u << 1,
  OR_C22 = 1u << 2,
  OR_S22 = 1u << 3,
  OR_SOL = 1u << 4,
  OR_LUN = 1u << 5,
  OR_SRP = 1u << 6,
  OR_DRAG = 1u << 7
};

/* ActivityState, src/ladds/particle/Particle.h:37-53 */
enum { OR_PASSIVE = 0, OR_EVASIVE = 1, OR_EVASIVE_PRESERVING = 2 };

/* ---- constants actually used (exported so tests can read them) ---- */
double or_constant(const char *name);

/* ---- ephemerides and environment (uniform per sub-step) ----
 * t: seconds since J2000.0 (2000-01-01 12:00 TT), TT ~ UT1 assumed.
 * parity unpinned against OrbitPropagator (its ephemeris is not in the tree); the Sun is
 * pinned by the J2000 almanac position, the Moon by a published worked example (Meeus
 * Example 47.a) and by physical invariants in the tests. */
void or_sun_position(double t, double out[3]);  /* km, EME2000 (M&G eq. 3.46-3.47) */
void or_moon_position(double t, double out[3]); /* km, EME2000 (M&G eq. 3.48-3.50) */
/* The lunar series with caller-supplied amplitudes (longitude and latitude in arcsec,
 * distance in km; or_moon_series_default gives the published ones).  Lets the tests plant
 * a mistyped term and show that the pins catch it.  or_moon_position is pinned by the
 * worked example of Meeus, Astronomical Algorithms (2nd ed. 1998) Example 47.a. */
#define OR_MOON_NLAM 14
#define OR_MOON_NBETA 8
#define OR_MOON_NR 9
void or_moon_series_default(double lam[OR_MOON_NLAM], double beta[OR_MOON_NBETA], double r[OR_MOON_NR]);
void or_moon_position_series(double t, const double *lam, const double *beta, const double *r, double out[3]);
double or_earth_rotation_angle(double t);       /* rad, GMST(J2000) + omega_E t, reduced to [0, 2pi) */
double or_density(double h_km);                 /* kg/m^3, Vallado Table 8-4 exponential model */

/* per-object parameters for the force model */
typedef struct {
  double aom;    /* area over mass [km^2/kg]     Particle.h:72 */
  double bc_inv; /* c_D A / m      [m^2/kg]      Particle.cpp:13-27 */
  double cr;     /* radiation pressure coefficient (no reference field) */
  int32_t activity;
} or_params;

/* acceleration of one object at time t [km/s^2]; comp selects one component bit only
 * (or the full mask).  evasivePreserving objects get KEP only (Particle.h:49-52). */
void or_acceleration(const double r[3], const double v[3], double t, unsigned mask, const or_params *p,
                     double a[3]);

/* Yoshida-4 coefficients c[4], d[3] (Yoshida 1990) */
void or_yoshida_coefficients(double c[4], double d[3]);

/* nsteps Yoshida-4 steps over SoA arrays, in place; step k starts at t0 + k dt.
 * OpenMP over objects; per-object arithmetic is thread-count independent. */
void or_yoshida_steps(size_t n, double *rx, double *ry, double *rz, double *vx, double *vy, double *vz,
                      const double *aom, const double *bc_inv, const double *cr, const int32_t *activity,
                      double t0, double dt, unsigned mask, int nsteps, int nthreads);

/* Vector-Jacobian product of or_yoshida_steps: with (r_T, v_T) the state after nsteps
 * steps from (r, v), writes o_x = sum_k g_k d(state_T)_k / dx for x in r0, v0, bc_inv
 * (o_bc may be NULL).  Complex-step derivatives (h = 1e-20), 7 complex propagations per
 * object; pinned against central finite differences of or_yoshida_steps and the closed
 * form of the straight line (all components off). */
void or_yoshida_vjp(size_t n, const double *rx, const double *ry, const double *rz, const double *vx,
                    const double *vy, const double *vz, const double *aom, const double *bc_inv, const double *cr,
                    const int32_t *activity, double t0, double dt, unsigned mask, int nsteps, const double *g_rx,
                    const double *g_ry, const double *g_rz, const double *g_vx, const double *g_vy,
                    const double *g_vz, double *o_rx, double *o_ry, double *o_rz, double *o_vx, double *o_vy,
                    double *o_vz, double *o_bc, int nthreads);

/* ---- conjunction predicate ---- */
void or_distance_by_linear_interpolation(const double ri[3], const double vi[3], const double rj[3],
                                         const double vj[3], double dt, double dr[3], double p2[3]);

typedef struct {
  double cutoff_km;                 /* autopas/cutoff                       */
  double dt;                        /* interpolation window                 */
  double collision_distance_factor; /* sim/collisionDistanceFactor          */
  double min_detection_radius_m;    /* sim/minDetectionRadius               */
  double evasion_cutoff_km;         /* sim/evasionTrackingCutoffInKM        */
  int threshold_mode;               /* 0 radius-scaled (reference), 1 absolute */
  double threshold_km;              /* absolute mode only                   */
} or_functor_cfg;

typedef struct {
  uint64_t id_a, id_b; /* smaller id first, as CollisionFunctor.cpp:94-128 */
  uint32_t idx_a, idx_b;
  double d2;
  double mid[3];
} or_pair; /* 56 B, same layout as ladds_pair */

typedef struct {
  size_t n;
  const double *rx, *ry, *rz, *vx, *vy, *vz;
  const double *radius_m;
  const int32_t *activity;
  const uint64_t *id, *parent;
} or_objects;

/* AoSFunctor(i, j): returns 0 (nothing), 1 (collision), 2 (evaded); fills *out for 1/2. */
int or_collision_functor(const or_objects *o, size_t i, size_t j, const or_functor_cfg *cfg, or_pair *out);

/* All unordered pairs, AoSFunctor called with (lower index, higher index).
 * Returns total counts; writes at most cap entries to each list.  Output unsorted for grid. */
void or_pairs_bruteforce(const or_objects *o, const or_functor_cfg *cfg, or_pair *hits, size_t cap, size_t *n,
                         or_pair *evaded, size_t ecap, size_t *ne);
void or_pairs_grid(const or_objects *o, const or_functor_cfg *cfg, or_pair *hits, size_t cap, size_t *n,
                   or_pair *evaded, size_t ecap, size_t *ne, int nthreads);

/* Count and order-independent digest (or_pairset_digest's definition) of the collisions
 * and of the evasions that or_pairs_grid reports, without storing the pairs: the check
 * for populations whose pair sets are too large to hold (dense constellation clusters).
 * Same traversal as or_pairs_grid, so it inherits that function's pins. */
void or_pairs_grid_digest(const or_objects *o, const or_functor_cfg *cfg, uint64_t *n, uint64_t digest[2],
                          uint64_t *ne, uint64_t edigest[2], int nthreads);

/* sort pairs by (id_a, id_b, idx_a, idx_b) */
void or_sort_pairs(or_pair *p, size_t n);

/* order-independent 128-bit digest over (id_a, id_b) pairs, same definition as ladds_pairset_digest */
void or_pairset_digest(const or_pair *p, size_t n, uint64_t digest[2]);

#ifdef __cplusplus
}
#endif
#endif

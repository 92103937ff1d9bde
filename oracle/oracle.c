/*
 * oracle.c -- CPU restatement of the LADDS propagate + conjunction hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  See oracle.h
 * for what each function follows and what pins it.  Build: oracle/Makefile
 * (gcc -O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math -fopenmp).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double or_constant(const char *name) {
  static const struct {
    const char *n;
    double v;
  } tab[] = {{"mu", MU_EARTH},       {"r_earth", R_EARTH},     {"j2", J2_EARTH},
             {"c22", C22_EARTH},     {"s22", S22_EARTH},       {"gm_sun", GM_SUN},
             {"gm_moon", GM_MOON},   {"au", AU_KM},            {"p_sun", P_SUN},
             {"omega_earth", OMEGA_EARTH}, {"theta_j2000", THETA_J2000}, {"obliquity", OBLIQUITY}};
  for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i)
    if (strcmp(tab[i].n, name) == 0) return tab[i].v;
  return NAN;
}

/* the force model and the object step, instantiated for double */
#define OR_T double
#define OR_SQRT sqrt
#define OR_EXP exp
#define OR_RE(x) (x)
#define OR_FN(name) real_##name
#include "force_model.inc"

double or_density(double h) { return real_density(h); }

/* ------------------------------------------------------------------------- */
/* ephemerides                                                                */
/* ------------------------------------------------------------------------- */
void or_sun_position(double t, double out[3]) {
  /* Montenbruck & Gill (2000) eq. 3.46-3.47: mean anomaly, ecliptic longitude
   * (Omega+omega = 282.9400 deg), distance; ecliptic latitude 0. */
  const double T = t / SEC_PER_CENTURY;
  const double M = (357.5256 + 35999.049 * T) * DEG;
  const double L = 282.9400 * DEG + M + (6892.0 * sin(M) + 72.0 * sin(2.0 * M)) * ARCSEC;
  const double r = (149.619 - 2.499 * cos(M) - 0.021 * cos(2.0 * M)) * 1.0e6;
  const double ce = cos(OBLIQUITY), se = sin(OBLIQUITY);
  out[0] = r * cos(L);
  out[1] = r * sin(L) * ce;
  out[2] = r * sin(L) * se;
}

/* Amplitudes of the low-precision lunar series, Montenbruck & Gill (2000) eq. 3.48-3.50,
 * in the order the terms are summed in or_moon_position_series: longitude and latitude
 * in arcsec, distance in km (the mean distance first). */
static const double MOON_LAM[OR_MOON_NLAM] = {22640.0, 769.0, -4586.0, 2370.0, -668.0, -412.0, -212.0,
                                              -206.0,  192.0, -165.0,  148.0,  -125.0, -110.0, -55.0};
static const double MOON_BETA[OR_MOON_NBETA] = {18520.0, -526.0, 44.0, -31.0, -25.0, -23.0, 21.0, 11.0};
static const double MOON_R[OR_MOON_NR] = {385000.0, -20905.0, -3699.0, -2956.0, -570.0,
                                          246.0,    -205.0,   -171.0,  -152.0};

void or_moon_series_default(double lam[OR_MOON_NLAM], double beta[OR_MOON_NBETA], double r[OR_MOON_NR]) {
  memcpy(lam, MOON_LAM, sizeof(MOON_LAM));
  memcpy(beta, MOON_BETA, sizeof(MOON_BETA));
  memcpy(r, MOON_R, sizeof(MOON_R));
}

void or_moon_position_series(double t, const double *A, const double *B, const double *R, double out[3]) {
  /* Montenbruck & Gill (2000) eq. 3.48-3.50: fundamental arguments (mean longitude referred
   * to the J2000 equinox by the -1.3972 T precession term), then the three series. */
  const double T = t / SEC_PER_CENTURY;
  const double L0 = (218.31617 + 481267.88088 * T - 1.3972 * T) * DEG;
  const double l = (134.96292 + 477198.86753 * T) * DEG;
  const double lp = (357.52543 + 35999.04944 * T) * DEG;
  const double F = (93.27283 + 483202.01873 * T) * DEG;
  const double D = (297.85027 + 445267.11135 * T) * DEG;
  const double lam =
      L0 + (A[0] * sin(l) + A[1] * sin(2.0 * l) + A[2] * sin(l - 2.0 * D) + A[3] * sin(2.0 * D) + A[4] * sin(lp) +
            A[5] * sin(2.0 * F) + A[6] * sin(2.0 * l - 2.0 * D) + A[7] * sin(l + lp - 2.0 * D) +
            A[8] * sin(l + 2.0 * D) + A[9] * sin(lp - 2.0 * D) + A[10] * sin(l - lp) + A[11] * sin(D) +
            A[12] * sin(l + lp) + A[13] * sin(2.0 * F - 2.0 * D)) *
               ARCSEC;
  const double beta =
      (B[0] * sin(F + lam - L0 + (412.0 * sin(2.0 * F) + 541.0 * sin(lp)) * ARCSEC) + B[1] * sin(F - 2.0 * D) +
       B[2] * sin(l + F - 2.0 * D) + B[3] * sin(-l + F - 2.0 * D) + B[4] * sin(-2.0 * l + F) +
       B[5] * sin(lp + F - 2.0 * D) + B[6] * sin(-l + F) + B[7] * sin(-lp + F - 2.0 * D)) *
      ARCSEC;
  const double r = R[0] + R[1] * cos(l) + R[2] * cos(2.0 * D - l) + R[3] * cos(2.0 * D) + R[4] * cos(2.0 * l) +
                   R[5] * cos(2.0 * l - 2.0 * D) + R[6] * cos(lp - 2.0 * D) + R[7] * cos(l + 2.0 * D) +
                   R[8] * cos(l + lp - 2.0 * D);
  const double xe = r * cos(lam) * cos(beta);
  const double ye = r * sin(lam) * cos(beta);
  const double ze = r * sin(beta);
  const double ce = cos(OBLIQUITY), se = sin(OBLIQUITY);
  out[0] = xe;
  out[1] = ye * ce - ze * se;
  out[2] = ye * se + ze * ce;
}

void or_moon_position(double t, double out[3]) { or_moon_position_series(t, MOON_LAM, MOON_BETA, MOON_R, out); }

double or_earth_rotation_angle(double t) {
  double th = fmod(THETA_J2000 + OMEGA_EARTH * t, TWO_PI);
  if (th < 0.0) th += TWO_PI;
  return th;
}

/* ------------------------------------------------------------------------- */
/* force model                                                                */
/* ------------------------------------------------------------------------- */
void or_env_at(double t, or_env *e) {
  const double th = or_earth_rotation_angle(t);
  e->ct = cos(th);
  e->st = sin(th);
  or_sun_position(t, e->sun);
  or_moon_position(t, e->moon);
}

void or_acceleration(const double r[3], const double v[3], double t, unsigned mask, const or_params *p,
                     double a[3]) {
  or_env e;
  or_env_at(t, &e);
  const real_params q = {p->aom, p->bc_inv, p->cr, p->activity};
  real_accel_env(r, v, mask, &q, &e, a);
}

void or_yoshida_coefficients(double c[4], double d[3]) {
  const double cb = cbrt(2.0);
  const double w1 = 1.0 / (2.0 - cb);
  const double w0 = -cb / (2.0 - cb);
  c[0] = c[3] = w1 / 2.0;
  c[1] = c[2] = (w0 + w1) / 2.0;
  d[0] = d[2] = w1;
  d[1] = w0;
}

void or_yoshida_steps(size_t n, double *rx, double *ry, double *rz, double *vx, double *vy, double *vz,
                      const double *aom, const double *bc_inv, const double *cr, const int32_t *activity,
                      double t0, double dt, unsigned mask, int nsteps, int nthreads) {
  double c[4], d[3];
  or_yoshida_coefficients(c, d);
  const double tau[3] = {c[0], c[0] + c[1], c[0] + c[1] + c[2]};
  or_env *env = (or_env *)malloc(sizeof(or_env) * 3 * (size_t)(nsteps > 0 ? nsteps : 1));
  for (int s = 0; s < nsteps; ++s) {
    const double t = t0 + (double)s * dt;
    for (int k = 0; k < 3; ++k) or_env_at(t + tau[k] * dt, &env[3 * s + k]);
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (long long i = 0; i < (long long)n; ++i) {
    const real_params p = {aom ? aom[i] : 0.0, bc_inv ? bc_inv[i] : 0.0, cr ? cr[i] : 1.0,
                           activity ? activity[i] : 0};
    double r[3] = {rx[i], ry[i], rz[i]}, v[3] = {vx[i], vy[i], vz[i]};
    real_yoshida_object(r, v, &p, env, c, d, dt, mask, nsteps);
    rx[i] = r[0];
    ry[i] = r[1];
    rz[i] = r[2];
    vx[i] = v[0];
    vy[i] = v[1];
    vz[i] = v[2];
  }
  free(env);
}

/* ------------------------------------------------------------------------- */
/* conjunction predicate                                                      */
/* ------------------------------------------------------------------------- */
/* autopas::utils::ArrayMath::dot: result = 0; result += a[k] * b[k] for k = 0..2 */
static inline double dot3(const double a[3], const double b[3]) {
  double s = 0.0;
  s += a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}

void or_distance_by_linear_interpolation(const double ri[3], const double vi[3], const double rj[3],
                                         const double vj[3], double dt, double dr[3], double p2[3]) {
  /* DistanceApproximation.cpp:25-55, operation for operation */
  double oi[3], oj[3];
  for (int k = 0; k < 3; ++k) {
    oi[k] = ri[k] - vi[k] * dt;
    oj[k] = rj[k] - vj[k] * dt;
  }
  const double vi_ri = dot3(vi, oi);
  const double vi_rj = dot3(vi, oj);
  const double vj_ri = dot3(vj, oi);
  const double vj_rj = dot3(vj, oj);
  const double nominator = vi_rj + vj_ri - vi_ri - vj_rj;
  const double two_vi_vj = 2.0 * dot3(vi, vj);
  const double vi_square = dot3(vi, vi);
  const double vj_square = dot3(vj, vj);
  const double denominator = vi_square + vj_square - two_vi_vj;
  double t = nominator / denominator;
  /* std::clamp(t, 0., dt): NaN passes through */
  t = (t < 0.0) ? 0.0 : ((dt < t) ? dt : t);
  for (int k = 0; k < 3; ++k) {
    const double p1k = oi[k] + vi[k] * t;
    const double p2k = oj[k] + vj[k] * t;
    dr[k] = p1k - p2k;
    p2[k] = p2k;
  }
}

int or_collision_functor(const or_objects *o, size_t i, size_t j, const or_functor_cfg *cfg, or_pair *out) {
  /* CollisionFunctor.cpp:40-131 (constructor scalings :24-28) */
  const double cutoff_sq = cfg->cutoff_km * cfg->cutoff_km;
  const double cdf = cfg->collision_distance_factor / 1000.;
  const double evasion_sq = cfg->evasion_cutoff_km * cfg->evasion_cutoff_km;
  const int ai = o->activity[i], aj = o->activity[j];
  const double Ri = o->radius_m[i], Rj = o->radius_m[j];
  int evaded = 0;
  if ((ai != OR_PASSIVE && aj != OR_PASSIVE) || (ai != OR_PASSIVE && Rj >= cfg->min_detection_radius_m) ||
      (aj != OR_PASSIVE && Ri >= cfg->min_detection_radius_m))
    evaded = 1;
  if (o->parent[i] != UINT64_MAX && o->parent[i] == o->parent[j]) return 0;
  const double ri[3] = {o->rx[i], o->ry[i], o->rz[i]};
  const double rj[3] = {o->rx[j], o->ry[j], o->rz[j]};
  const double d0[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  const double distanceSquare = dot3(d0, d0);
  if (distanceSquare > cutoff_sq) return 0;
  const double vi[3] = {o->vx[i], o->vy[i], o->vz[i]};
  const double vj[3] = {o->vx[j], o->vy[j], o->vz[j]};
  double dr[3], p2[3];
  or_distance_by_linear_interpolation(ri, vi, rj, vj, cfg->dt, dr, p2);
  const double dl = dot3(dr, dr);
  const double scaled = (Ri + Rj) * cdf;
  if (evaded) {
    if (dl > evasion_sq) return 0;
  } else if (cfg->threshold_mode == 1) {
    if (dl > cfg->threshold_km * cfg->threshold_km) return 0;
  } else {
    if (dl > scaled * scaled) return 0;
  }
  for (int k = 0; k < 3; ++k) out->mid[k] = p2[k] + dr[k] * 0.5;
  out->d2 = dl;
  if (o->id[i] < o->id[j]) {
    out->id_a = o->id[i];
    out->id_b = o->id[j];
    out->idx_a = (uint32_t)i;
    out->idx_b = (uint32_t)j;
  } else {
    out->id_a = o->id[j];
    out->id_b = o->id[i];
    out->idx_a = (uint32_t)j;
    out->idx_b = (uint32_t)i;
  }
  return evaded ? 2 : 1;
}

typedef struct {
  or_pair *v;
  size_t n, cap;
} pvec;

static void pvec_push(pvec *p, const or_pair *x) {
  if (p->n == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 64;
    p->v = (or_pair *)realloc(p->v, p->cap * sizeof(or_pair));
  }
  p->v[p->n++] = *x;
}

static void emit(const or_pair *pr, int kind, or_pair *hits, size_t cap, size_t *n, or_pair *evaded, size_t ecap,
                 size_t *ne) {
  if (kind == 1) {
    if (*n < cap) hits[*n] = *pr;
    ++*n;
  } else if (kind == 2) {
    if (*ne < ecap) evaded[*ne] = *pr;
    ++*ne;
  }
}

void or_pairs_bruteforce(const or_objects *o, const or_functor_cfg *cfg, or_pair *hits, size_t cap, size_t *n,
                         or_pair *evaded, size_t ecap, size_t *ne) {
  *n = 0;
  *ne = 0;
  or_pair pr;
  for (size_t i = 0; i < o->n; ++i)
    for (size_t j = i + 1; j < o->n; ++j) {
      const int kind = or_collision_functor(o, i, j, cfg, &pr);
      emit(&pr, kind, hits, cap, n, evaded, ecap, ne);
    }
}

typedef struct {
  uint64_t key;
  uint32_t idx;
} keyidx;

#define GRID_BIAS 1048576LL /* 2^20: cell coordinates stored in 21 bits each */
static uint64_t pack_cell(long long ix, long long iy, long long iz) {
  return ((uint64_t)(ix + GRID_BIAS) << 42) | ((uint64_t)(iy + GRID_BIAS) << 21) | (uint64_t)(iz + GRID_BIAS);
}

static uint64_t splitmix64(uint64_t x);

/* Stable LSD radix sort of (key, idx) records by key, 16 bits per pass.  The input is in
 * index order, so the result is ordered by (key, idx). */
static void radix_sort_keyidx(keyidx *a, size_t n) {
  keyidx *tmp = (keyidx *)malloc(sizeof(keyidx) * (n ? n : 1));
  size_t *cnt = (size_t *)malloc(sizeof(size_t) * 65536);
  for (int pass = 0; pass < 4; ++pass) {
    const int sh = 16 * pass;
    memset(cnt, 0, sizeof(size_t) * 65536);
    for (size_t i = 0; i < n; ++i) ++cnt[(a[i].key >> sh) & 0xFFFFu];
    size_t used = 0;
    for (size_t b = 0; b < 65536; ++b) used += cnt[b] != 0;
    if (used <= 1) continue; /* every key has the same digit: the pass keeps the order */
    size_t run = 0;
    for (size_t b = 0; b < 65536; ++b) {
      const size_t c = cnt[b];
      cnt[b] = run;
      run += c;
    }
    for (size_t i = 0; i < n; ++i) tmp[cnt[(a[i].key >> sh) & 0xFFFFu]++] = a[i];
    memcpy(a, tmp, sizeof(keyidx) * n);
  }
  free(cnt);
  free(tmp);
}

/* What the grid scan does with each reported pair. */
typedef struct {
  pvec hits, ev;        /* list mode */
  uint64_t nh, ne;      /* digest mode: counts */
  uint64_t dh[2], de[2]; /* digest mode: sums of the pair hashes (or_pairset_digest) */
} grid_acc;

static void grid_record(grid_acc *acc, int digest_only, int kind, const or_pair *pr) {
  if (kind == 0) return;
  if (!digest_only) {
    pvec_push(kind == 1 ? &acc->hits : &acc->ev, pr);
    return;
  }
  const uint64_t h = splitmix64(pr->id_a ^ splitmix64(pr->id_b));
  const uint64_t h2 = splitmix64(h ^ 0xD1B54A32D192ED03ull);
  if (kind == 1) {
    acc->nh += 1;
    acc->dh[0] += h;
    acc->dh[1] += h2;
  } else {
    acc->ne += 1;
    acc->de[0] += h;
    acc->de[1] += h2;
  }
}

/* Linked cells of side cutoff*(1+1e-7): every pair within the cutoff lies in neighbouring
 * cells, and AoSFunctor(i, j) is called for it from its lower index i only, so every
 * unordered pair is evaluated exactly once (AutoPas Newton3 traversal).  Cells are found
 * by binary search over the sorted unique cell keys.  Returns one accumulator per thread
 * (*nt of them), or NULL when a coordinate is beyond the packed range (caller falls back
 * to the brute force). */
static grid_acc *grid_scan(const or_objects *o, const or_functor_cfg *cfg, int nthreads, int digest_only, int *nt_out) {
  const size_t N = o->n;
  const double s = cfg->cutoff_km * (1.0 + 1e-7);
  int nt = 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  *nt_out = nt;
  if (!(s > 0.0)) return NULL;
  long long *cell = (long long *)malloc(sizeof(long long) * 3 * N);
  keyidx *ki = (keyidx *)malloc(sizeof(keyidx) * N);
  int fallback = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(| : fallback)
#endif
  for (long long ii = 0; ii < (long long)N; ++ii) {
    const size_t i = (size_t)ii;
    const double u[3] = {o->rx[i] / s, o->ry[i] / s, o->rz[i] / s};
    int bad = 0;
    for (int k = 0; k < 3; ++k) {
      if (!(fabs(u[k]) < (double)(GRID_BIAS - 2))) {
        bad = 1;
        break;
      }
      cell[3 * i + k] = (long long)floor(u[k]);
    }
    fallback |= bad;
    if (!bad) {
      ki[i].key = pack_cell(cell[3 * i], cell[3 * i + 1], cell[3 * i + 2]);
      ki[i].idx = (uint32_t)i;
    }
  }
  if (fallback) {
    free(ki);
    free(cell);
    return NULL;
  }
  radix_sort_keyidx(ki, N);
  /* unique cells and their ranges */
  size_t nu = 0;
  uint64_t *ukey = (uint64_t *)malloc(sizeof(uint64_t) * N);
  size_t *ubeg = (size_t *)malloc(sizeof(size_t) * (N + 1));
  for (size_t i = 0; i < N; ++i) {
    if (i == 0 || ki[i].key != ki[i - 1].key) {
      ukey[nu] = ki[i].key;
      ubeg[nu] = i;
      ++nu;
    }
  }
  ubeg[nu] = N;
  grid_acc *acc = (grid_acc *)calloc((size_t)nt, sizeof(grid_acc));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (long long ii = 0; ii < (long long)N; ++ii) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    const size_t i = (size_t)ii;
    or_pair pr;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          const uint64_t key = pack_cell(cell[3 * i] + dx, cell[3 * i + 1] + dy, cell[3 * i + 2] + dz);
          size_t lo = 0, hi = nu;
          while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (ukey[mid] < key)
              lo = mid + 1;
            else
              hi = mid;
          }
          if (lo == nu || ukey[lo] != key) continue;
          for (size_t q = ubeg[lo]; q < ubeg[lo + 1]; ++q) {
            const size_t j = ki[q].idx;
            if (j <= i) continue;
            grid_record(&acc[tid], digest_only, or_collision_functor(o, i, j, cfg, &pr), &pr);
          }
        }
  }
  free(ubeg);
  free(ukey);
  free(ki);
  free(cell);
  return acc;
}

void or_pairs_grid(const or_objects *o, const or_functor_cfg *cfg, or_pair *hits, size_t cap, size_t *n,
                   or_pair *evaded, size_t ecap, size_t *ne, int nthreads) {
  *n = 0;
  *ne = 0;
  if (o->n < 2) return;
  int nt = 1;
  grid_acc *acc = grid_scan(o, cfg, nthreads, 0, &nt);
  if (!acc) {
    or_pairs_bruteforce(o, cfg, hits, cap, n, evaded, ecap, ne);
    return;
  }
  for (int t = 0; t < nt; ++t) {
    for (size_t k = 0; k < acc[t].hits.n; ++k) emit(&acc[t].hits.v[k], 1, hits, cap, n, evaded, ecap, ne);
    for (size_t k = 0; k < acc[t].ev.n; ++k) emit(&acc[t].ev.v[k], 2, hits, cap, n, evaded, ecap, ne);
    free(acc[t].hits.v);
    free(acc[t].ev.v);
  }
  free(acc);
}

void or_pairs_grid_digest(const or_objects *o, const or_functor_cfg *cfg, uint64_t *n, uint64_t digest[2],
                          uint64_t *ne, uint64_t edigest[2], int nthreads) {
  *n = *ne = 0;
  digest[0] = digest[1] = edigest[0] = edigest[1] = 0;
  if (o->n < 2) return;
  int nt = 1;
  grid_acc *acc = grid_scan(o, cfg, nthreads, 1, &nt);
  if (!acc) {
    /* beyond the packed cell range: the brute force, hashed the same way */
    grid_acc one;
    memset(&one, 0, sizeof(one));
    or_pair pr;
    for (size_t i = 0; i < o->n; ++i)
      for (size_t j = i + 1; j < o->n; ++j) grid_record(&one, 1, or_collision_functor(o, i, j, cfg, &pr), &pr);
    acc = (grid_acc *)calloc(1, sizeof(grid_acc));
    *acc = one;
    nt = 1;
  }
  for (int t = 0; t < nt; ++t) {
    *n += acc[t].nh;
    *ne += acc[t].ne;
    for (int k = 0; k < 2; ++k) {
      digest[k] += acc[t].dh[k];
      edigest[k] += acc[t].de[k];
    }
  }
  free(acc);
}

static int cmp_pair(const void *a, const void *b) {
  const or_pair *x = (const or_pair *)a, *y = (const or_pair *)b;
  if (x->id_a != y->id_a) return x->id_a < y->id_a ? -1 : 1;
  if (x->id_b != y->id_b) return x->id_b < y->id_b ? -1 : 1;
  if (x->idx_a != y->idx_a) return x->idx_a < y->idx_a ? -1 : 1;
  if (x->idx_b != y->idx_b) return x->idx_b < y->idx_b ? -1 : 1;
  return 0;
}

void or_sort_pairs(or_pair *p, size_t n) { qsort(p, n, sizeof(or_pair), cmp_pair); }

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void or_pairset_digest(const or_pair *p, size_t n, uint64_t digest[2]) {
  /* multiset hash: two independent 64-bit sums of a mixed pair hash */
  uint64_t s0 = 0, s1 = 0;
  for (size_t k = 0; k < n; ++k) {
    const uint64_t h = splitmix64(p[k].id_a ^ splitmix64(p[k].id_b));
    s0 += h;
    s1 += splitmix64(h ^ 0xD1B54A32D192ED03ull);
  }
  digest[0] = s0;
  digest[1] = s1;
}

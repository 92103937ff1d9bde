// population.cpp -- seeded synthetic debris populations (host only).
//
// The reference loads catalogue CSVs (src/ladds/io/DatasetReader.cpp:15-50) whose
// data files are not part of its tree, so workloads are generated from Keplerian
// element distributions with a counter-based generator: every draw is
// splitmix64(seed, object, draw), so a population is independent of thread
// count and any object can be regenerated alone.
//
// Element conventions: the altitude draw is the PERIGEE altitude, a = (R_E + h)/(1 - e),
// so no object starts below the 150 km burn-up altitude; elements are converted with
// a Newton solve of Kepler's equation to 1e-15.  Object attributes follow
// src/ladds/particle/Particle.h:62-85 (aom = pi r^2 1e-6 / m [km^2/kg]) and
// Particle::calculateBcInv (Particle.cpp:13-27: bc_inv = c_D pi r^2 / m, c_D = 2.2).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/ladds_amd.h"
#include "constants.h"

namespace ladds {
namespace {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {
  uint64_t seed;
  double uniform(uint64_t obj, uint32_t draw) const {
    const uint64_t h = splitmix64(seed ^ splitmix64(obj * 64u + draw));
    return (double)(h >> 11) * 0x1.0p-53;
  }
  double normal(uint64_t obj, uint32_t draw) const {
    const double u1 = uniform(obj, draw), u2 = uniform(obj, draw + 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * kPi * u2);
  }
  double log_uniform(uint64_t obj, uint32_t draw, double lo, double hi) const {
    return std::exp(std::log(lo) + uniform(obj, draw) * (std::log(hi) - std::log(lo)));
  }
};

void elements_to_state(double a, double e, double inc, double raan, double argp, double M, double r[3],
                       double v[3]) {
  double E = M;
  for (int it = 0; it < 50; ++it) {
    const double f = E - e * std::sin(E) - M;
    const double dE = f / (1.0 - e * std::cos(E));
    E -= dE;
    if (std::fabs(dE) < 1e-15) break;
  }
  const double cE = std::cos(E), sE = std::sin(E);
  const double sq = std::sqrt(1.0 - e * e);
  const double xp = a * (cE - e), yp = a * sq * sE;
  const double rn = a * (1.0 - e * cE);
  const double k = std::sqrt(kMuEarth * a) / rn;
  const double vxp = -k * sE, vyp = k * sq * cE;
  const double cO = std::cos(raan), sO = std::sin(raan);
  const double cw = std::cos(argp), sw = std::sin(argp);
  const double ci = std::cos(inc), si = std::sin(inc);
  // perifocal -> inertial: R3(-raan) R1(-inc) R3(-argp)
  const double P[3] = {cO * cw - sO * sw * ci, sO * cw + cO * sw * ci, sw * si};
  const double Q[3] = {-cO * sw - sO * cw * ci, -sO * sw + cO * cw * ci, cw * si};
  for (int k3 = 0; k3 < 3; ++k3) {
    r[k3] = xp * P[k3] + yp * Q[k3];
    v[k3] = vxp * P[k3] + vyp * Q[k3];
  }
}

struct Attr {
  double radius, mass, aom, bc_inv, cr;
};

Attr attributes(const Rng &g, uint64_t i, int workload) {
  Attr at;
  double radius = g.log_uniform(i, 9, 0.05, 5.0);
  if (workload == 1 && g.uniform(i, 13) < 0.9) radius = g.log_uniform(i, 14, 0.005, 0.1);
  const double rho = 500.0 + 2500.0 * g.uniform(i, 10);
  at.radius = radius;
  at.mass = 4.0 / 3.0 * kPi * radius * radius * radius * rho;
  const double area = kPi * radius * radius;
  at.aom = area * 1e-6 / at.mass;
  at.bc_inv = (workload == 1) ? 0.005 + 0.045 * g.uniform(i, 11) : 2.2 * area / at.mass;
  at.cr = 1.0 + 0.5 * g.uniform(i, 12);
  return at;
}

void background_object(const Rng &g, uint64_t i, double r[3], double v[3]) {
  const double h = 200.0 + 1800.0 * g.uniform(i, 0);
  const double e = 0.02 * g.uniform(i, 1);
  const double sel = g.uniform(i, 2);
  double inc_deg;
  if (sel < 0.5)
    inc_deg = 53.0 + 2.0 * g.normal(i, 3);
  else if (sel < 0.8)
    inc_deg = 98.0 + 1.0 * g.normal(i, 3);
  else
    inc_deg = 180.0 * g.uniform(i, 5);
  inc_deg = std::fmin(std::fmax(inc_deg, 0.0), 180.0);
  const double a = (kREarth + h) / (1.0 - e);
  elements_to_state(a, e, inc_deg * kPi / 180.0, 2.0 * kPi * g.uniform(i, 6), 2.0 * kPi * g.uniform(i, 7),
                    2.0 * kPi * g.uniform(i, 8), r, v);
}

// Walker-like shells: altitude [km], inclination [deg], planes, satellites per plane.
struct Shell {
  double h, inc;
  int planes, sats;
};
constexpr Shell kShells[3] = {{550.0, 53.0, 72, 22}, {1110.0, 87.9, 36, 49}, {1200.0, 70.0, 40, 33}};

void constellation_slot(uint64_t slot, double r[3], double v[3]) {
  uint64_t s = slot;
  for (const Shell &sh : kShells) {
    const uint64_t cnt = (uint64_t)sh.planes * (uint64_t)sh.sats;
    if (s < cnt) {
      const int p = (int)(s / (uint64_t)sh.sats), k = (int)(s % (uint64_t)sh.sats);
      const double raan = 2.0 * kPi * p / sh.planes;
      const double u = 2.0 * kPi * k / sh.sats + 2.0 * kPi * p / ((double)sh.planes * sh.sats);
      elements_to_state(kREarth + sh.h, 0.0, sh.inc * kPi / 180.0, raan, 0.0, u, r, v);
      return;
    }
    s -= cnt;
  }
}

constexpr uint64_t kSlots = 72 * 22 + 36 * 49 + 40 * 33;  // 4668

}  // namespace
}  // namespace ladds

extern "C" int ladds_generate_population(int workload, uint64_t seed, size_t n, size_t planted_pairs, double *rx,
                                         double *ry, double *rz, double *vx, double *vy, double *vz,
                                         double *radius_m, double *mass_kg, double *aom, double *bc_inv,
                                         double *cr) {
  using namespace ladds;
  if (workload < 0 || workload > 2 || !rx || !ry || !rz || !vx || !vy || !vz) return LADDS_E_INVALID;
  if (planted_pairs * 2 > n) return LADDS_E_INVALID;
  const Rng g{seed};
  const size_t nbase = n - planted_pairs;
  const size_t nclustered = (workload == 2) ? (size_t)(0.6 * (double)nbase) : 0;
  auto body = [&](uint64_t i) {
    double r[3], v[3];
    if (i < nclustered) {
      constellation_slot(i % kSlots, r, v);
      for (int k = 0; k < 3; ++k) r[k] += 0.05 * (2.0 * g.uniform(i, 20 + k) - 1.0);
    } else {
      background_object(g, i, r, v);
    }
    rx[i] = r[0];
    ry[i] = r[1];
    rz[i] = r[2];
    vx[i] = v[0];
    vy[i] = v[1];
    vz[i] = v[2];
    const Attr at = attributes(g, i, workload);
    if (radius_m) radius_m[i] = at.radius;
    if (mass_kg) mass_kg[i] = at.mass;
    if (aom) aom[i] = at.aom;
    if (bc_inv) bc_inv[i] = at.bc_inv;
    if (cr) cr[i] = at.cr;
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (uint64_t i = t; i < nbase; i += nt) body(i);
    });
  for (std::thread &th : pool) th.join();
  // planted partners: even k co-orbital (along-track offset), odd k crossing (velocity
  // rotated about the radial direction), both 0.3-0.9 km from their base object.
  const size_t stride = planted_pairs ? nbase / planted_pairs : 1;
  for (size_t k = 0; k < planted_pairs; ++k) {
    const size_t b = k * stride, i = nbase + k;
    const double r[3] = {rx[b], ry[b], rz[b]}, v[3] = {vx[b], vy[b], vz[b]};
    const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double off = 0.3 + 0.6 * g.uniform(i, 30);
    double nr[3], nv[3];
    if (k % 2 == 0) {
      for (int c = 0; c < 3; ++c) {
        nr[c] = r[c] + off * v[c] / vn;
        nv[c] = v[c];
      }
    } else {
      const double ang = (1.0 + 4.0 * g.uniform(i, 31)) * kPi / 180.0;
      const double u[3] = {r[0] / rn, r[1] / rn, r[2] / rn};
      // Rodrigues rotation of v about u
      const double cxv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
      const double udv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
      for (int c = 0; c < 3; ++c) {
        nv[c] = v[c] * std::cos(ang) + cxv[c] * std::sin(ang) + u[c] * udv * (1.0 - std::cos(ang));
        nr[c] = r[c] + off * u[c];
      }
    }
    rx[i] = nr[0];
    ry[i] = nr[1];
    rz[i] = nr[2];
    vx[i] = nv[0];
    vy[i] = nv[1];
    vz[i] = nv[2];
    const Attr at = attributes(g, i, workload);
    if (radius_m) radius_m[i] = at.radius;
    if (mass_kg) mass_kg[i] = at.mass;
    if (aom) aom[i] = at.aom;
    if (bc_inv) bc_inv[i] = at.bc_inv;
    if (cr) cr[i] = at.cr;
  }
  return LADDS_OK;
}

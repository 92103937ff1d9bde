/*
 * ladds_amd.hpp -- C++ adapter with the call surface LADDS's simulation loop uses.
 *
 * Header-only, built on the C ABI in ladds_amd.h.  The classes carry the names and
 * constructor arguments of the objects LADDS builds around its hot path, so the loop in
 * src/ladds/Simulation.cpp keeps its shape when it is pointed at this engine:
 *
 *   LADDS                                                  here
 *   LADDS::Particle (particle/Particle.h:21-195)           ladds_amd::Particle
 *   autopas::AutoPas<Particle> (Simulation.cpp:59-120)     ladds_amd::Container
 *   Acceleration::AccelerationAccumulator (:142-143)       ladds_amd::AccelerationAccumulator
 *   YoshidaIntegrator (:145, integrate() at :297)          ladds_amd::YoshidaIntegrator
 *   CollisionFunctor (CollisionFunctor.h:17-110)           ladds_amd::CollisionFunctor
 *   AutoPas::iteratePairwise (Simulation.cpp:220)          Container::iteratePairwise
 *   DatasetReader::readDataset (io/DatasetReader.cpp)      ladds_amd::DatasetReader
 *   ConjunctionLogger (io/ConjuctionLogger.cpp)            ladds_amd::ConjunctionLogger
 *
 * Ownership follows the reference: the container owns the objects (on the device, with a
 * host mirror that is refreshed on demand); the accumulator, integrator and functor hold
 * non-owning references.  Hit tuples hold Particle* into the container's host mirror and
 * stay valid until the container is next mutated or re-synchronised.  Errors from the
 * engine are rethrown as std::runtime_error.  A container is bound to one host thread.
 */
#ifndef LADDS_AMD_HPP
#define LADDS_AMD_HPP

#include <algorithm>
#include <array>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ladds_amd.h"

namespace ladds_amd {

inline void check(int rc, const ladds_ctx *ctx, const char *what) {
  if (rc == LADDS_OK) return;
  std::string msg = std::string(what) + " failed (status " + std::to_string(rc) + ")";
  if (ctx) {
    const char *detail = ladds_last_error(ctx);
    if (detail && *detail) msg += ": " + std::string(detail);
  }
  throw std::runtime_error(msg);
}

/* The attributes of LADDS::Particle that the hot path reads (Particle.h:37-85). */
class Particle {
 public:
  enum class ActivityState : int { passive = LADDS_PASSIVE, evasive = LADDS_EVASIVE, evasivePreserving = LADDS_EVASIVE_PRESERVING };

  /* Particle.cpp:13-27: bc_inv from c_D and the cross-section when B* is absent or 0,
   * otherwise 2 R_E / 0.1570 * B*.  R_E is the propagator's Earth radius in km, the unit
   * Simulation.cpp:573 adds it to a km altitude in. */
  static double calculateBcInv(double bstar, double mass, double radius, double coefficientOfDrag) {
    if (std::isnan(bstar) || bstar == 0.0) return coefficientOfDrag * (M_PI * radius * radius) / mass;
    const double earth_radius_km = 6378.1363;
    return (2.0 * earth_radius_km / 0.1570) * bstar;
  }

  /* Same argument order as LADDS::Particle's constructor (Particle.h:60-80); cr is the
   * radiation pressure coefficient this engine adds (1.0 matches a plain cannonball). */
  Particle(std::array<double, 3> pos, std::array<double, 3> v, size_t debrisId, std::string identifier,
           ActivityState activityState, double mass, double radius, double bcInv, size_t parentIdentifier,
           double cr = 1.0)
      : r_(pos), v_(v), id_(debrisId), identifier_(std::move(identifier)), activity_(activityState), mass_(mass),
        radius_(radius), bc_inv_(bcInv), aom_(M_PI * radius * radius * 1e-6 / mass), parent_(parentIdentifier),
        cr_(cr) {}

  const std::array<double, 3> &getPosition() const { return r_; }
  const std::array<double, 3> &getVelocity() const { return v_; }
  void setPosition(const std::array<double, 3> &r) { r_ = r; }
  void setVelocity(const std::array<double, 3> &v) { v_ = v; }
  size_t getID() const { return id_; }
  const std::string &getIdentifier() const { return identifier_; }
  ActivityState getActivityState() const { return activity_; }
  double getMass() const { return mass_; }
  double getRadius() const { return radius_; }
  double getBcInv() const { return bc_inv_; }
  double getAom() const { return aom_; }
  size_t getParentIdentifier() const { return parent_; }
  void setParentIdentifier(size_t parent) { parent_ = parent; }
  double getCr() const { return cr_; }
  double getHeight() const { return std::sqrt(r_[0] * r_[0] + r_[1] * r_[1] + r_[2] * r_[2]); }
  double getSpeed() const { return std::sqrt(v_[0] * v_[0] + v_[1] * v_[1] + v_[2] * v_[2]); }

 private:
  std::array<double, 3> r_, v_;
  size_t id_;
  std::string identifier_;
  ActivityState activity_;
  double mass_, radius_, bc_inv_, aom_;
  size_t parent_;
  double cr_;
};

class CollisionFunctor;

/* Parameters fixed when the container is built; Simulation::initAutoPas takes them from
 * the YAML keys named beside each field (Simulation.cpp:59-120). */
struct ContainerConfig {
  double cutoff = 80.0;          /* autopas/cutoff [km]                                  */
  double cellSize = 0.0;         /* 0: chosen by the engine from cutoff and population   */
  int precision = LADDS_FP64;    /* LADDS_FP64 | LADDS_SPLIT | LADDS_SPLIT_ALL32          */
  int scanVariant = 0;           /* 0: automatic, 1: atomic counting sort, 2: plane tiles */
  int device = 0;                /* HIP device ordinal                                   */
  bool tune = true;              /* tune the scan on the first iteratePairwise (AutoPas's
                                    tuning phase) when the container holds enough objects */
};

/* The autopas::AutoPas<Particle> subset the hot path uses: particle storage, iteration,
 * and the pairwise traversal with a CollisionFunctor. */
class Container {
 public:
  using iterator = std::vector<Particle>::iterator;
  using const_iterator = std::vector<Particle>::const_iterator;

  explicit Container(const ContainerConfig &c = ContainerConfig{}) : cc_(c) {
    ladds_config_default(&cfg_);
    cfg_.cutoff_km = c.cutoff;
    cfg_.cell_size_km = c.cellSize;
    cfg_.precision = c.precision;
    cfg_.scan_variant = c.scanVariant;
    cfg_.device = c.device;
    check(ladds_create(&cfg_, &ctx_), nullptr, "ladds_create");
    applied_dt_ = cfg_.dt;
    applied_t0_ = cfg_.t0;
  }
  ~Container() { ladds_destroy(ctx_); }
  Container(const Container &) = delete;
  Container &operator=(const Container &) = delete;

  double getCutoff() const { return cc_.cutoff; }
  void addParticle(const Particle &p) {
    syncHost();
    particles_.push_back(p);
    ++appended_;  // appended on the device (ladds_add_objects) unless a full upload is due
  }
  /* autopas.deleteParticle (BreakupWrapper.cpp:24-25): removed on the device, the mirror
   * follows at its next download; Particle* into the mirror stay valid until then. */
  void deleteParticle(const Particle &p) {
    syncDevice();
    const uint64_t id = p.getID();
    check(ladds_delete_objects(ctx_, &id, 1, nullptr), ctx_, "ladds_delete_objects");
    host_stale_ = true;
  }
  size_t getNumberOfParticles() { syncHost(); return particles_.size(); }
  /* autopas.forEach: f applied to every object of the host mirror, uploaded before the
   * next device call. */
  template <typename F>
  void forEach(F f) {
    syncHost();
    for (Particle &p : particles_) f(p);
    device_dirty_ = true;
  }
  /* Simulation::removeParticleSpawnProtection (Simulation.cpp:470-473) on the device: every
   * parent id reset, without an upload of the population. */
  void removeParticleSpawnProtection() {
    syncDevice();
    check(ladds_remove_spawn_protection(ctx_), ctx_, "ladds_remove_spawn_protection");
    for (Particle &p : particles_) p.setParentIdentifier(std::numeric_limits<size_t>::max());
  }
  /* Host view of the objects; reading it after device work downloads the state first. */
  iterator begin() { syncHost(); device_dirty_ = true; return particles_.begin(); }
  iterator end() { return particles_.end(); }
  const_iterator cbegin() { syncHost(); return particles_.cbegin(); }
  const_iterator cend() const { return particles_.cend(); }

  /* AutoPas::iteratePairwise with a CollisionFunctor (Simulation.cpp:211-222).  Returns
   * whether this call was a tuning iteration (AutoPas's return value): the first call on a
   * container of at least kTuneMinObjects objects times the scan's shapes first
   * (ladds_autotune); its pairs are exact like every other call's. */
  inline bool iteratePairwise(CollisionFunctor *f);
  static constexpr size_t kTuneMinObjects = 16384;

  /* Simulation::deleteBurnUps (Simulation.cpp:561-596): drop objects below the altitude. */
  size_t deleteBurnUps(double minAltitude) {
    syncDevice();
    size_t removed = 0;
    check(ladds_delete_burnups(ctx_, minAltitude, &removed), ctx_, "ladds_delete_burnups");
    host_stale_ = true;
    return removed;
  }

  /* The fused loop body of Simulation::simulationLoop (Simulation.cpp:287-415): nsteps
   * integrations with burn-up removal and a conjunction scan every tpcd iterations,
   * logged on the device.  The logged pairs are returned by fetchLog(). */
  ladds_run_stats simulate(uint64_t firstIteration, int nsteps, int tpcd, double minAltitude) {
    syncDevice();
    ladds_run_stats st{};
    check(ladds_simulate(ctx_, firstIteration, nsteps, tpcd, minAltitude, &st), ctx_, "ladds_simulate");
    host_stale_ = true;
    return st;
  }
  std::vector<ladds_logged_pair> fetchLog() {
    size_t n = 0;
    int rc = ladds_fetch_log(ctx_, nullptr, 0, &n);
    if (rc != LADDS_OK && rc != LADDS_E_OVERFLOW) check(rc, ctx_, "ladds_fetch_log");
    std::vector<ladds_logged_pair> out(n);
    check(ladds_fetch_log(ctx_, out.data(), out.size(), &n), ctx_, "ladds_fetch_log");
    out.resize(n);
    check(ladds_clear_log(ctx_), ctx_, "ladds_clear_log");
    return out;
  }
  /* What HDF5Writer::writeParticles stores (io/hdf5/HDF5Writer.cpp:89-149), packed on the
   * device: float32 xyz positions / velocities (3 per object) and uint32 ids. */
  void particleRecords(std::vector<float> &pos, std::vector<float> &vel, std::vector<uint32_t> &ids) {
    syncDevice();
    const size_t n = ladds_num_objects(ctx_);
    pos.resize(3 * n);
    vel.resize(3 * n);
    ids.resize(n);
    check(ladds_particle_records(ctx_, pos.data(), vel.data(), ids.data()), ctx_, "ladds_particle_records");
  }
  /* What HDF5Writer::writeConjunctions / writeEvasions store for one iteration
   * (HDF5Writer.cpp:151-187), from the device log (call before fetchLog, which clears it). */
  void logRecords(uint64_t iteration, std::vector<ladds_pair_record> &collisions,
                  std::vector<ladds_pair_record> &evasions) {
    size_t n = 0, ne = 0;
    int rc = ladds_log_records(ctx_, iteration, nullptr, 0, &n, nullptr, 0, &ne);
    if (rc != LADDS_OK && rc != LADDS_E_OVERFLOW) check(rc, ctx_, "ladds_log_records");
    collisions.resize(n);
    evasions.resize(ne);
    check(ladds_log_records(ctx_, iteration, collisions.data(), n, &n, evasions.data(), ne, &ne), ctx_,
          "ladds_log_records");
  }

  /* ---- used by the adapter classes ---- */
  ladds_ctx *context() { return ctx_; }
  ladds_config &config() { return cfg_; }
  void applyConfig() {
    // unchanged since the last call (iteratePairwise every iteration with the same functor
    // parameters): nothing to send
    if (configured_ && std::memcmp(&cfg_, &applied_cfg_, sizeof(cfg_)) == 0) return;
    const bool dt_only = cfg_.dt != applied_dt_ && cfg_.t0 == applied_t0_;
    check(ladds_configure(ctx_, &cfg_), ctx_, "ladds_configure");
    // a new step alone keeps the elapsed time: the context's clock restarts at the current time
    if (dt_only) cfg_.t0 = ladds_get_time(ctx_);
    applied_dt_ = cfg_.dt;
    applied_t0_ = cfg_.t0;
    applied_cfg_ = cfg_;
    configured_ = true;
  }
  /* Upload the host mirror if it was changed since the last upload: the whole mirror after
   * host-side edits, only the new tail after addParticle calls (breakup fragments). */
  void syncDevice() {
    const size_t n = particles_.size();
    if (device_dirty_) {
      upload(0, n, ladds_set_objects, "ladds_set_objects");
      host_stale_ = false;  // the device now holds exactly the mirror
    } else if (appended_) {
      upload(n - appended_, n, ladds_add_objects, "ladds_add_objects");
    }
    device_dirty_ = false;
    appended_ = 0;
  }
  /* Download the device state into the host mirror if the device moved ahead of it.
   * Objects removed on the device (burn-ups) are removed from the mirror, order kept. */
  void syncHost() {
    if (!host_stale_) return;
    const size_t n = ladds_num_objects(ctx_);
    std::vector<double> s(6 * n);
    std::vector<uint64_t> id(n);
    check(ladds_get_state(ctx_, s.data(), s.data() + n, s.data() + 2 * n, s.data() + 3 * n, s.data() + 4 * n,
                          s.data() + 5 * n), ctx_, "ladds_get_state");
    check(ladds_get_ids(ctx_, id.data()), ctx_, "ladds_get_ids");
    if (n != particles_.size()) {
      std::unordered_map<uint64_t, size_t> where;
      where.reserve(particles_.size());
      for (size_t i = 0; i < particles_.size(); ++i) where.emplace(particles_[i].getID(), i);
      std::vector<Particle> kept;
      kept.reserve(n);
      for (size_t i = 0; i < n; ++i) kept.push_back(particles_[where.at(id[i])]);
      particles_.swap(kept);
    }
    for (size_t i = 0; i < n; ++i) {
      particles_[i].setPosition({s[i], s[n + i], s[2 * n + i]});
      particles_[i].setVelocity({s[3 * n + i], s[4 * n + i], s[5 * n + i]});
    }
    host_stale_ = false;
  }
  void markDeviceAhead() { host_stale_ = true; }
  /* Bring the mirror entries of the given device indices up to date (the objects a
   * detection returned), or the whole mirror when objects were removed on the device. */
  void refresh(const std::vector<uint32_t> &idx) {
    if (!host_stale_) return;
    if (ladds_num_objects(ctx_) != particles_.size()) {
      syncHost();
      return;
    }
    if (idx.empty()) return;
    std::vector<double> r(3 * idx.size()), v(3 * idx.size());
    check(ladds_gather_state(ctx_, idx.data(), idx.size(), r.data(), v.data()), ctx_, "ladds_gather_state");
    for (size_t k = 0; k < idx.size(); ++k) {
      particles_[idx[k]].setPosition({r[3 * k], r[3 * k + 1], r[3 * k + 2]});
      particles_[idx[k]].setVelocity({v[3 * k], v[3 * k + 1], v[3 * k + 2]});
    }
  }
  /* Particle* of the object at device index idx (host mirror must be current). */
  Particle *particleAt(uint32_t idx) { return &particles_.at(idx); }

 private:
  ContainerConfig cc_;
  ladds_config cfg_{};
  ladds_ctx *ctx_ = nullptr;
  std::vector<Particle> particles_;
  bool device_dirty_ = false;
  bool host_stale_ = false;
  size_t appended_ = 0;
  double applied_dt_ = -1.0, applied_t0_ = 0.0;  // the configuration the context holds
  ladds_config applied_cfg_{};
  bool configured_ = false;
  bool tuned_ = false;
  // pair buffers reused by every iteratePairwise (grown when a detection reports more)
  std::vector<ladds_pair> hit_buf_, ev_buf_;

  void upload(size_t b, size_t e, int (*entry)(ladds_ctx *, const ladds_objects *), const char *what) {
    const size_t n = e - b;
    std::vector<double> col(10 * n);
    std::vector<int32_t> act(n);
    std::vector<uint64_t> id(n), parent(n);
    double *rx = col.data(), *ry = rx + n, *rz = ry + n, *vx = rz + n, *vy = vx + n, *vz = vy + n;
    double *rad = vz + n, *aom = rad + n, *bc = aom + n, *cr = bc + n;
    for (size_t i = 0; i < n; ++i) {
      const Particle &p = particles_[b + i];
      rx[i] = p.getPosition()[0]; ry[i] = p.getPosition()[1]; rz[i] = p.getPosition()[2];
      vx[i] = p.getVelocity()[0]; vy[i] = p.getVelocity()[1]; vz[i] = p.getVelocity()[2];
      rad[i] = p.getRadius(); aom[i] = p.getAom(); bc[i] = p.getBcInv(); cr[i] = p.getCr();
      act[i] = static_cast<int32_t>(p.getActivityState());
      id[i] = p.getID(); parent[i] = p.getParentIdentifier();
    }
    ladds_objects o{n, rx, ry, rz, vx, vy, vz, rad, aom, bc, cr, act.data(), id.data(), parent.data()};
    check(entry(ctx_, &o), ctx_, what);
  }
};

namespace Acceleration {
/* AccelerationAccumulator(components, container, t0, fileOutput) (Simulation.cpp:142-143).
 * Component order KEP, J2, C22, S22, SOL, LUN, SRP, DRAG as the sim/prop/use*Component
 * keys (Simulation.cpp:127-135).  Per-step CSV output of accelerations (the fileOutput
 * argument) is not produced by this engine; pass nullptr. */
class AccelerationAccumulator {
 public:
  AccelerationAccumulator(const std::array<bool, 8> &components, Container &container, double t0,
                          void *fileOutput = nullptr)
      : container_(container) {
    if (fileOutput != nullptr) throw std::runtime_error("AccelerationAccumulator: file output is not supported");
    uint32_t mask = 0;
    for (int k = 0; k < 8; ++k)
      if (components[k]) mask |= 1u << k;
    container_.config().components = mask;
    container_.config().t0 = t0;
    container_.applyConfig();
  }
  Container &container() { return container_; }

 private:
  Container &container_;
};
}  // namespace Acceleration

/* YoshidaIntegrator(container, accumulator, dt) (Simulation.cpp:145). */
class YoshidaIntegrator {
 public:
  YoshidaIntegrator(Container &container, Acceleration::AccelerationAccumulator &accumulator, double deltaT)
      : container_(container) {
    (void)accumulator;
    container_.config().dt = deltaT;
    container_.applyConfig();
  }
  /* One Yoshida-4 step of every object (Simulation.cpp:297).  The flag asks the reference
   * to keep accelerations for file output, which this engine does not produce. */
  void integrate(bool writeAcc = false) {
    (void)writeAcc;
    container_.syncDevice();
    check(ladds_propagate(container_.context(), 1), container_.context(), "ladds_propagate");
    container_.markDeviceAhead();
  }
  double getDeltaT() const { return container_.config().dt; }

 private:
  Container &container_;
};

/* CollisionFunctor(cutoff, dt, collisionDistanceFactor, minDetectionRadius,
 * evasionTrackingCutoffInKM) (CollisionFunctor.h:17-30).  The scan runs on the device when
 * the container iterates with it; getCollisions()/getEvadedCollisions() then hold
 * (Particle*, Particle*, squared distance, collision point) with the smaller id first. */
class CollisionFunctor {
 public:
  using CollisionT = std::tuple<Particle *, Particle *, double, std::array<double, 3>>;
  using CollisionCollectionT = std::vector<CollisionT>;

  CollisionFunctor(double cutoff, double dt, double collisionDistanceFactor, double minDetectionRadius,
                   double evasionTrackingCutoffInKM)
      : cutoff_(cutoff), dt_(dt), cdf_(collisionDistanceFactor), min_det_(minDetectionRadius),
        evasion_(evasionTrackingCutoffInKM) {}

  const CollisionCollectionT &getCollisions() const { return collisions_; }
  const CollisionCollectionT &getEvadedCollisions() const { return evaded_; }

 private:
  friend class Container;
  double cutoff_, dt_, cdf_, min_det_, evasion_;
  CollisionCollectionT collisions_, evaded_;
};

inline bool Container::iteratePairwise(CollisionFunctor *f) {
  if (f->cutoff_ != cc_.cutoff)
    throw std::runtime_error("CollisionFunctor cutoff differs from the container cutoff");
  syncDevice();
  cfg_.collision_distance_factor = f->cdf_;
  cfg_.min_detection_radius_m = f->min_det_;
  cfg_.evasion_cutoff_km = f->evasion_;
  cfg_.threshold_mode = LADDS_THRESHOLD_RADIUS_SCALED;
  applyConfig();
  bool tuning = false;
  if (cc_.tune && cc_.scanVariant != 1 && !tuned_ && ladds_num_objects(ctx_) >= kTuneMinObjects) {
    ladds_tuning t{};
    check(ladds_autotune(ctx_, f->dt_, 2, &t), ctx_, "ladds_autotune");
    tuned_ = tuning = true;
  }
  size_t n = 0, ne = 0;
  if (hit_buf_.size() < 4096) hit_buf_.resize(4096);
  if (ev_buf_.size() < 4096) ev_buf_.resize(4096);
  int rc = ladds_conjunctions(ctx_, f->dt_, hit_buf_.data(), hit_buf_.size(), &n, ev_buf_.data(), ev_buf_.size(), &ne);
  if (rc == LADDS_E_OVERFLOW) {  // rerun with the sizes the first pass reported
    hit_buf_.resize(std::max(n, hit_buf_.size()));
    ev_buf_.resize(std::max(ne, ev_buf_.size()));
    rc = ladds_conjunctions(ctx_, f->dt_, hit_buf_.data(), hit_buf_.size(), &n, ev_buf_.data(), ev_buf_.size(), &ne);
  }
  check(rc, ctx_, "ladds_conjunctions");
  const ladds_pair *hits = hit_buf_.data(), *evaded = ev_buf_.data();
  // only the objects the tuples point at are brought up to date (the rest of the mirror
  // stays behind the device until it is read)
  std::vector<uint32_t> touched;
  touched.reserve(2 * (n + ne));
  for (size_t k = 0; k < n; ++k) {
    touched.push_back(hits[k].idx_a);
    touched.push_back(hits[k].idx_b);
  }
  for (size_t k = 0; k < ne; ++k) {
    touched.push_back(evaded[k].idx_a);
    touched.push_back(evaded[k].idx_b);
  }
  std::sort(touched.begin(), touched.end());
  touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
  refresh(touched);
  auto convert = [this](const ladds_pair *src, size_t count, CollisionFunctor::CollisionCollectionT &dst) {
    dst.clear();
    dst.reserve(count);
    for (size_t k = 0; k < count; ++k) {
      const ladds_pair &p = src[k];
      Particle *a = particleAt(p.idx_a), *b = particleAt(p.idx_b);
      dst.emplace_back(a, b, p.d2, std::array<double, 3>{p.mid[0], p.mid[1], p.mid[2]});
    }
  };
  convert(hits, n, f->collisions_);
  convert(evaded, ne, f->evaded_);
  return tuning;
}

/* DatasetReader::readDataset (src/ladds/io/DatasetReader.cpp:15-50): a CSV file with one
 * header line and the columns id, COSPAR id, name, B* [1/Earth radius], mass [kg],
 * radius [m], activity (passive | evasive | evasivePreserving, Particle.cpp:130-141),
 * r x y z [km], v x y z [km/s].  bc_inv follows Particle::calculateBcInv with the given
 * drag coefficient; parent ids are "none" (max size_t). */
class DatasetReader {
 public:
  static std::vector<Particle> readDataset(const std::string &csvFilepath, double coefficientOfDrag) {
    std::ifstream in(csvFilepath);
    if (!in) throw std::runtime_error("cannot open dataset " + csvFilepath);
    std::string line;
    std::getline(in, line);  // header
    std::vector<Particle> out;
    size_t lineno = 1;
    while (std::getline(in, line)) {
      ++lineno;
      if (line.empty() || line == "\r") continue;
      if (!line.empty() && line.back() == '\r') line.pop_back();
      std::vector<std::string> f;
      std::stringstream ls(line);
      std::string cell;
      while (std::getline(ls, cell, ',')) f.push_back(cell);
      if (f.size() != 13) throw std::runtime_error(csvFilepath + ":" + std::to_string(lineno) + ": expected 13 columns");
      auto num = [&](size_t k) {
        try {
          return std::stod(f[k]);
        } catch (const std::exception &) {
          throw std::runtime_error(csvFilepath + ":" + std::to_string(lineno) + ": bad number '" + f[k] + "'");
        }
      };
      Particle::ActivityState act = Particle::ActivityState::passive;
      if (f[6] == "evasive") act = Particle::ActivityState::evasive;
      if (f[6] == "evasivePreserving") act = Particle::ActivityState::evasivePreserving;
      const double bstar = num(3), mass = num(4), radius = num(5);
      out.emplace_back(std::array<double, 3>{num(7), num(8), num(9)}, std::array<double, 3>{num(10), num(11), num(12)},
                       static_cast<size_t>(std::stoull(f[0])), f[1], act, mass, radius,
                       Particle::calculateBcInv(bstar, mass, radius, coefficientOfDrag),
                       std::numeric_limits<size_t>::max());
    }
    return out;
  }
};

/* ConjunctionLogger (src/ladds/io/ConjuctionLogger.cpp:13-44): CSV with the header
 * "Iteration,P1,P2,SquaredDistance" and one line per pair (smaller id first), doubles in
 * their shortest round-trip form.  Collisions and evasions go to the same file, as in the
 * reference. */
class ConjunctionLogger {
 public:
  explicit ConjunctionLogger(const std::string &path) : out_(path) {
    if (!out_) throw std::runtime_error("cannot open " + path);
    out_ << "Iteration,P1,P2,SquaredDistance\n";
  }
  void log(uint64_t iteration, uint64_t id1, uint64_t id2, double distanceSquared) {
    char buf[32];
    const auto r = std::to_chars(buf, buf + sizeof(buf), distanceSquared);
    out_ << iteration << ',' << id1 << ',' << id2 << ',' << std::string(buf, r.ptr) << '\n';
  }
  void writeConjunctions(uint64_t iteration, const std::vector<std::tuple<Particle *, Particle *, double,
                                                                           std::array<double, 3>>> &collisions) {
    for (const auto &c : collisions) log(iteration, std::get<0>(c)->getID(), std::get<1>(c)->getID(), std::get<2>(c));
  }
  /* Pairs logged on the device by Container::simulate. */
  void writeLog(const std::vector<ladds_logged_pair> &entries) {
    for (const ladds_logged_pair &e : entries) log(e.iteration, e.pair.id_a, e.pair.id_b, e.pair.d2);
  }

 private:
  std::ofstream out_;
};

}  // namespace ladds_amd

#endif

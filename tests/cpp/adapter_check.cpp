// Drives the C++ adapter (include/ladds_amd.hpp) the way LADDS's simulation loop drives
// AutoPas + OrbitPropagator: build a container, an accumulator, an integrator, integrate,
// then iterate a CollisionFunctor over the container (Simulation.cpp:122-148, 211-222, 297).
//
// Reads a scenario, one command per line, from the file named by argv[1] and writes results
// to stdout; tests/test_cpp_adapter.py writes the scenarios and checks the output.
//
//   container <cutoff_km> <precision> <scan_variant>
//   accumulator <KEP J2 C22 S22 SOL LUN SRP DRAG as 0/1> <t0>
//   integrator <dt>
//   particle <rx ry rz vx vy vz id activity mass radius_m bc_inv parent cr>
//   integrate <steps>
//   functor <cutoff dt cdf min_det_radius_m evasion_cutoff_km>   -> "hit|evaded idA idB d2 mx my mz"
//                           and "hitpos id rx ry rz vx vy vz" of both Particle* of every tuple
//   state                                                         -> "state id rx ry rz vx vy vz"
//   burnups <min_altitude_km>                                     -> "removed n"
//   simulate <first_iteration> <steps> <tpcd> <min_altitude_km>   -> "stats ..." and "log ..."
//   conjlog <path>          ConjunctionLogger CSV of the last simulate's log
//   unprotect               Simulation::removeParticleSpawnProtection
//   population <workload> <seed> <n> <planted_pairs>   seeded synthetic objects (ladds_generate_population)
//   timeloop <steps> <cutoff dt cdf min_det_radius_m evasion_cutoff_km>
//                           the per-iteration drop-in sequence of Simulation.cpp:297 and :211-222,
//                           integrate(false) + a fresh CollisionFunctor + iteratePairwise, timed
//                           -> "timing steps seconds collisions evasions"
//   dataset <path> <c_D>    DatasetReader::readDataset (no device needed)
//                           -> "data id cospar activity mass radius bc_inv rx ry rz vx vy vz"
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "ladds_amd.hpp"

using namespace ladds_amd;

static void print_pairs(const char *tag, const CollisionFunctor::CollisionCollectionT &c) {
  for (const auto &[a, b, d2, mid] : c)
    std::printf("%s %zu %zu %.17g %.17g %.17g %.17g\n", tag, a->getID(), b->getID(), d2, mid[0], mid[1], mid[2]);
}

// What the tuples' Particle* hold (the adapter refreshes only these objects per detection).
static void print_hit_particles(const CollisionFunctor::CollisionCollectionT &c) {
  for (const auto &t : c)
    for (const Particle *p : {std::get<0>(t), std::get<1>(t)}) {
      const auto &r = p->getPosition();
      const auto &v = p->getVelocity();
      std::printf("hitpos %zu %.17g %.17g %.17g %.17g %.17g %.17g\n", p->getID(), r[0], r[1], r[2], v[0], v[1], v[2]);
    }
}

// Every field of the line was read and nothing but blanks follows.
static bool parsed(std::istringstream &ls, const std::string &line) {
  std::string rest;
  if (!ls.fail() && !(ls >> rest)) return true;
  std::fprintf(stderr, "malformed line: %s\n", line.c_str());
  return false;
}

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s scenario.txt\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1]);
  if (!in) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  std::unique_ptr<Container> container;
  std::unique_ptr<Acceleration::AccelerationAccumulator> accumulator;
  std::unique_ptr<YoshidaIntegrator> integrator;
  std::vector<ladds_logged_pair> last_log;
  std::string line;
  try {
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string cmd;
      if (!(ls >> cmd) || cmd[0] == '#') continue;
      if (cmd == "dataset") {
        std::string path;
        double cd = 0.0;
        ls >> path >> cd;
        if (!parsed(ls, line)) return 2;
        for (const Particle &p : DatasetReader::readDataset(path, cd)) {
          const auto &r = p.getPosition();
          const auto &v = p.getVelocity();
          std::printf("data %zu %s %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p.getID(),
                      p.getIdentifier().c_str(), static_cast<int>(p.getActivityState()), p.getMass(), p.getRadius(),
                      p.getBcInv(), r[0], r[1], r[2], v[0], v[1], v[2]);
        }
        continue;
      }
      if (cmd != "container" && !container) {
        std::fprintf(stderr, "%s before container\n", cmd.c_str());
        return 2;
      }
      if ((cmd == "integrator" && !accumulator) || (cmd == "integrate" && !integrator)) {
        std::fprintf(stderr, "%s before its accumulator/integrator\n", cmd.c_str());
        return 2;
      }
      if (cmd == "container") {
        ContainerConfig cc;
        ls >> cc.cutoff >> cc.precision >> cc.scanVariant;
        if (!parsed(ls, line)) return 2;
        container = std::make_unique<Container>(cc);
      } else if (cmd == "accumulator") {
        std::array<bool, 8> comp{};
        for (auto &c : comp) {
          int b = 0;
          ls >> b;
          c = b != 0;
        }
        double t0 = 0.0;
        ls >> t0;
        if (!parsed(ls, line)) return 2;
        accumulator = std::make_unique<Acceleration::AccelerationAccumulator>(comp, *container, t0, nullptr);
      } else if (cmd == "integrator") {
        double dt = 0.0;
        ls >> dt;
        if (!parsed(ls, line)) return 2;
        integrator = std::make_unique<YoshidaIntegrator>(*container, *accumulator, dt);
      } else if (cmd == "particle") {
        double rx, ry, rz, vx, vy, vz, mass, radius, bc, cr;
        unsigned long long id, parent;
        int act;
        ls >> rx >> ry >> rz >> vx >> vy >> vz >> id >> act >> mass >> radius >> bc >> parent >> cr;
        if (!parsed(ls, line)) return 2;
        container->addParticle(Particle({rx, ry, rz}, {vx, vy, vz}, id, "", static_cast<Particle::ActivityState>(act),
                                        mass, radius, bc, parent, cr));
      } else if (cmd == "integrate") {
        int steps = 0;
        ls >> steps;
        if (!parsed(ls, line)) return 2;
        for (int s = 0; s < steps; ++s) integrator->integrate(false);
      } else if (cmd == "functor") {
        double cutoff, dt, cdf, min_det, evasion;
        ls >> cutoff >> dt >> cdf >> min_det >> evasion;
        if (!parsed(ls, line)) return 2;
        CollisionFunctor f(cutoff, dt, cdf, min_det, evasion);
        container->iteratePairwise(&f);
        print_pairs("hit", f.getCollisions());
        print_pairs("evaded", f.getEvadedCollisions());
        print_hit_particles(f.getCollisions());
        print_hit_particles(f.getEvadedCollisions());
      } else if (cmd == "unprotect") {
        if (!parsed(ls, line)) return 2;
        container->removeParticleSpawnProtection();
      } else if (cmd == "population") {
        int workload = 0;
        unsigned long long seed = 0, n = 0, planted = 0;
        ls >> workload >> seed >> n >> planted;
        if (!parsed(ls, line)) return 2;
        std::vector<double> a(11 * n);
        double *c[11];
        for (int k = 0; k < 11; ++k) c[k] = a.data() + k * n;
        check(ladds_generate_population(workload, seed, n, planted, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7],
                                        c[8], c[9], c[10]),
              nullptr, "ladds_generate_population");
        for (size_t i = 0; i < n; ++i)
          container->addParticle(Particle({c[0][i], c[1][i], c[2][i]}, {c[3][i], c[4][i], c[5][i]}, i, "",
                                          Particle::ActivityState::passive, c[7][i], c[6][i], c[9][i],
                                          std::numeric_limits<size_t>::max(), c[10][i]));
      } else if (cmd == "timeloop") {
        int steps = 0;
        double cutoff, dt, cdf, min_det, evasion;
        ls >> steps >> cutoff >> dt >> cdf >> min_det >> evasion;
        if (!parsed(ls, line)) return 2;
        container->syncDevice();  // the population upload is not part of the loop
        unsigned long long hits = 0, evasions = 0;
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s) {
          integrator->integrate(false);
          CollisionFunctor f(cutoff, dt, cdf, min_det, evasion);
          container->iteratePairwise(&f);
          hits += f.getCollisions().size();
          evasions += f.getEvadedCollisions().size();
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("timing %d %.9g %llu %llu\n", steps, sec, hits, evasions);
      } else if (cmd == "state") {
        for (auto it = container->cbegin(); it != container->cend(); ++it) {
          const auto &r = it->getPosition();
          const auto &v = it->getVelocity();
          std::printf("state %zu %.17g %.17g %.17g %.17g %.17g %.17g\n", it->getID(), r[0], r[1], r[2], v[0], v[1],
                      v[2]);
        }
      } else if (cmd == "delete") {  // autopas.deleteParticle, as BreakupWrapper.cpp:24-25 does
        unsigned long long id = 0;
        ls >> id;
        if (!parsed(ls, line)) return 2;
        bool found = false;
        for (auto it = container->cbegin(); it != container->cend(); ++it)
          if (it->getID() == id) {
            container->deleteParticle(*it);
            found = true;
            break;
          }
        if (!found) {
          std::fprintf(stderr, "no particle with id %llu\n", id);
          return 2;
        }
      } else if (cmd == "burnups") {
        double h = 0.0;
        ls >> h;
        if (!parsed(ls, line)) return 2;
        std::printf("removed %zu\n", container->deleteBurnUps(h));
      } else if (cmd == "simulate") {
        unsigned long long first = 0;
        int steps = 0, tpcd = 1;
        double h = 0.0;
        ls >> first >> steps >> tpcd >> h;
        if (!parsed(ls, line)) return 2;
        const ladds_run_stats st = container->simulate(first, steps, tpcd, h);
        std::printf("stats %llu %llu %llu\n", (unsigned long long)st.collisions, (unsigned long long)st.evasions,
                    (unsigned long long)st.remaining);
        last_log = container->fetchLog();
        for (const ladds_logged_pair &lp : last_log)
          std::printf("log %llu %s %llu %llu %.17g\n", (unsigned long long)lp.iteration, lp.evaded ? "evaded" : "hit",
                      (unsigned long long)lp.pair.id_a, (unsigned long long)lp.pair.id_b, lp.pair.d2);
      } else if (cmd == "conjlog") {
        std::string path;
        ls >> path;
        if (!parsed(ls, line)) return 2;
        ConjunctionLogger logger(path);
        logger.writeLog(last_log);
      } else {
        std::fprintf(stderr, "unknown command: %s\n", cmd.c_str());
        return 2;
      }
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

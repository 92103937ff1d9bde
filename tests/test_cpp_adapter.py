"""The C++ adapter (include/ladds_amd.hpp) driven like LADDS's simulation loop.

tests/cpp/adapter_check builds a Container, AccelerationAccumulator and
YoshidaIntegrator, integrates, and iterates a CollisionFunctor over the container,
exactly as Simulation.cpp:122-148, 211-222 and 297 do with AutoPas and
OrbitPropagator.  Its output is checked against the reference's known answers
(CollisionFunctorTest.cpp, SimulationTest.cpp:39-57) and against the oracle.
"""
import subprocess

import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol
from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = ROOT / "tests" / "cpp" / "adapter_check"
UMAX = np.iinfo(np.uint64).max


def _run(tmp_path, lines):
    if not EXE.exists():
        raise FileNotFoundError(f"{EXE} is missing: run __graft_entry__.build() first")
    scen = tmp_path / "scenario.txt"
    scen.write_text("\n".join(lines) + "\n")
    res = subprocess.run([str(EXE), str(scen)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = {"hit": [], "evaded": [], "state": {}, "removed": [], "stats": [], "log": [], "hitpos": {}, "timing": []}
    for ln in res.stdout.splitlines():
        tag, *f = ln.split()
        if tag in ("hit", "evaded"):
            out[tag].append((int(f[0]), int(f[1]), float(f[2]), [float(x) for x in f[3:6]]))
        elif tag == "state":
            out["state"][int(f[0])] = np.array([float(x) for x in f[1:7]])
        elif tag == "hitpos":
            out["hitpos"][int(f[0])] = np.array([float(x) for x in f[1:7]])
        elif tag == "timing":
            out["timing"].append((int(f[0]), float(f[1]), int(f[2]), int(f[3])))
        elif tag == "removed":
            out["removed"].append(int(f[0]))
        elif tag == "stats":
            out["stats"].append(tuple(int(x) for x in f))
        elif tag == "log":
            out["log"].append((int(f[0]), f[1], int(f[2]), int(f[3]), float(f[4])))
    return out


def _particle(r, v, pid, act=0, mass=1.0, radius=1.0, bc=0.0, parent=UMAX, cr=1.0):
    vals = [repr(float(x)) for x in (*r, *v)]
    mass, radius, bc, cr = (repr(float(x)) for x in (mass, radius, bc, cr))
    return f"particle {' '.join(vals)} {int(pid)} {int(act)} {mass} {radius} {bc} {int(parent)} {cr}"


def _header(cutoff, dt, comps=(1, 0, 0, 0, 0, 0, 0, 0), precision=la.FP64, variant=0, t0=0.0):
    return [f"container {cutoff!r} {precision} {variant}",
            f"accumulator {' '.join(map(str, comps))} {t0!r}",
            f"integrator {dt!r}"]


def test_three_particles(tmp_path, golden_cases):
    """CollisionFunctorTest.cpp:18-59 through Container::iteratePairwise."""
    case = golden_cases["three_particles"]
    f = case["functor"]
    lines = _header(f["cutoff"], f["dt"])
    lines += [_particle(p["r"], p["v"], p["id"], p["activity"], p["mass"], p["radius"], parent=p["parent"])
              for p in case["particles"]]
    lines.append(f"functor {f['cutoff']!r} {f['dt']!r} {f['cdf']!r} {f['min_det']!r} {f['evasion']!r}")
    out = _run(tmp_path, lines)
    got = {(a, b): (d2, mid) for a, b, d2, mid in out["hit"]}
    exp = {(e["a"], e["b"]): (e["d2"], e["mid"]) for e in case["expected_collisions"]}
    assert got == exp


def test_mix_activity_states(tmp_path, golden_cases):
    """CollisionFunctorTest.cpp:76-209: 14 expected id pairs, all with NaN distance."""
    case = golden_cases["mix_activity_states"]
    f = case["functor"]
    lines = _header(f["cutoff"], f["dt"])
    lines += [_particle(p["r"], p["v"], p["id"], p["activity"], p["mass"], p["radius"], parent=p["parent"])
              for p in case["particles"]]
    lines.append(f"functor {f['cutoff']!r} {f['dt']!r} {f['cdf']!r} {f['min_det']!r} {f['evasion']!r}")
    out = _run(tmp_path, lines)
    assert sorted((a, b) for a, b, _, _ in out["hit"]) == sorted(tuple(p) for p in case["expected_id_pairs"])
    assert all(np.isnan(d2) for _, _, d2, _ in out["hit"])


@pytest.mark.parametrize("cdf", ["1.0", "2.0"])
def test_collision_distance_factor(tmp_path, golden_cases, cdf):
    """CollisionFunctorTest.cpp:215-258: one hit only with the larger distance factor."""
    case = golden_cases["collision_distance_factor"]
    f = case["functor"]
    lines = _header(f["cutoff"], f["dt"])
    lines += [_particle(p["r"], p["v"], p["id"], p["activity"], p["mass"], p["radius"], parent=p["parent"])
              for p in case["particles"]]
    lines.append(f"functor {f['cutoff']!r} {f['dt']!r} {cdf} {f['min_det']!r} {f['evasion']!r}")
    out = _run(tmp_path, lines)
    assert len(out["hit"]) == case["expected_count_by_cdf"][cdf]


def test_kepler_one_second(tmp_path, kepler_1s):
    """SimulationTest.cpp:39-57 through YoshidaIntegrator::integrate."""
    p1, p2 = kepler_1s["p1"], kepler_1s["p2"]
    lines = _header(10.0, kepler_1s["dt"]) + [_particle(p1["r"], p1["v"], 0), "integrate 1", "state"]
    st = _run(tmp_path, lines)["state"][0]
    assert np.max(np.abs(st[:3] - p2["r"])) < 1e-9
    assert np.max(np.abs(st[3:] - p2["v"])) < 1e-11


def test_pos_test_ten_seconds(tmp_path, pos_test_fixture):
    """data/pos_test.csv -> pos_test_10s.csv (Kepler, 10 s) through the adapter."""
    r0, v0, r1, v1 = pos_test_fixture
    lines = _header(10.0, 10.0) + [_particle(r0[i], v0[i], i) for i in range(r0.shape[0])]
    lines += ["integrate 1", "state"]
    st = _run(tmp_path, lines)["state"]
    r = np.array([st[i][:3] for i in range(r0.shape[0])])
    v = np.array([st[i][3:] for i in range(r0.shape[0])])
    assert np.max(np.linalg.norm(r - r1, axis=1) / np.linalg.norm(r1, axis=1)) < 3e-10
    assert np.max(np.abs(v - v1)) < 2e-9


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_population_full_model_against_oracle(tmp_path, precision):
    """2,048 objects + 64 planted pairs, full force model, 4 steps of 10 s, then the
    functor: FP64 positions within 1e-9 relative of the oracle; the pair set of the
    engine's own state is bit-identical to the oracle's brute force."""
    pop = la.generate_population(1, 0x1ADD5 + 7, 2048, planted_pairs=64)
    n = pop.n
    cdf, cutoff, dt, steps = 2.0e5, 10.0, 10.0, 4
    lines = _header(cutoff, dt, comps=(1,) * 8, precision=precision)
    for i in range(n):
        lines.append(_particle((pop.rx[i], pop.ry[i], pop.rz[i]), (pop.vx[i], pop.vy[i], pop.vz[i]), i,
                               mass=pop.mass_kg[i], radius=pop.radius_m[i], bc=pop.bc_inv[i], cr=pop.cr[i]))
    lines += [f"integrate {steps}", "state", f"functor {cutoff!r} {dt!r} {cdf!r} 0.05 0.1"]
    out = _run(tmp_path, lines)
    r = np.array([out["state"][i][:3] for i in range(n)])
    v = np.array([out["state"][i][3:] for i in range(n)])
    aom = np.pi * pop.radius_m * pop.radius_m * 1e-6 / pop.mass_kg
    ro, _ = ol.propagate(pop.r(), pop.v(), 0.0, dt, la.ALL_COMPONENTS, steps, aom=aom, bc_inv=pop.bc_inv,
                         cr=pop.cr)
    err = np.max(np.linalg.norm(r - ro, axis=1) / np.linalg.norm(ro, axis=1))
    assert err < (1e-9 if precision == la.FP64 else 1e-8), err
    objs = ol.ObjectSet(r, v, pop.radius_m, pop.activity, pop.id, pop.parent)
    oh, _ = ol.pairs(objs, ol.functor_cfg(cutoff, dt, cdf=cdf, min_det=0.05, evasion=0.1), method="brute")
    assert oh.shape[0] >= 16
    got = sorted((a, b, d2, tuple(mid)) for a, b, d2, mid in out["hit"])
    exp = sorted((int(p["id_a"]), int(p["id_b"]), float(p["d2"]), tuple(map(float, p["mid"]))) for p in oh)
    assert got == exp


def test_simulate_and_burnups_match_python_binding(tmp_path):
    """Container::simulate and deleteBurnUps give what the Python binding gives on the same input."""
    pop = la.generate_population(0, 0x1ADD5 + 8, 1024, planted_pairs=16)
    scale = (6378.1363 + 120.0) / np.linalg.norm(pop.r()[:8], axis=1)  # 8 objects at 120 km altitude
    pop.rx[:8] *= scale
    pop.ry[:8] *= scale
    pop.rz[:8] *= scale
    lines = _header(10.0, 10.0, comps=(1, 1, 0, 0, 0, 0, 0, 0))
    for i in range(pop.n):
        lines.append(_particle((pop.rx[i], pop.ry[i], pop.rz[i]), (pop.vx[i], pop.vy[i], pop.vz[i]), i,
                               mass=pop.mass_kg[i], radius=pop.radius_m[i], bc=pop.bc_inv[i], cr=pop.cr[i]))
    # the functor call sets the detection parameters simulate() then uses
    lines += ["functor 10.0 10.0 200000.0 0.05 0.1", "burnups 150.0", "simulate 0 6 2 150.0", "state"]
    out = _run(tmp_path, lines)
    cfg = la.EngineConfig(dt=10.0, components=la.KEP | la.J2, cutoff_km=10.0, collision_distance_factor=2.0e5)
    with la.Engine(cfg) as eng:
        p = pop.copy()
        p.aom = np.pi * pop.radius_m * pop.radius_m * 1e-6 / pop.mass_kg
        eng.set_objects(p)
        removed = eng.delete_burnups(150.0)
        st = eng.simulate(6, first_iteration=0, tpcd=2, min_altitude_km=150.0)
        log = eng.fetch_log()
        r, v = eng.state()
        ids = eng.ids()
    assert out["removed"] == [removed] and removed == 8
    assert out["stats"] == [(st["collisions"], st["evasions"], st["remaining"])]
    assert st["collisions"] > 0
    got = sorted(out["log"])
    exp = sorted((int(e["iteration"]), "evaded" if e["evaded"] else "hit", int(e["pair"]["id_a"]),
                  int(e["pair"]["id_b"]), float(e["pair"]["d2"])) for e in log)
    assert got == exp
    assert sorted(out["state"]) == sorted(int(i) for i in ids)
    for k, i in enumerate(ids):
        assert np.array_equal(out["state"][int(i)], np.concatenate([r[k], v[k]]))


def test_breakup_and_conjunction_log_match_python_binding(tmp_path):
    """BreakupWrapper.cpp:13-45 through the adapter (deleteParticle of a colliding pair,
    addParticle of fragments whose parent is the first object), then simulate() and the
    ConjunctionLogger CSV: same state, pairs and CSV rows as the Python binding."""
    rng = np.random.default_rng(3)
    pop = la.generate_population(0, 0x1ADD5 + 9, 1024, planted_pairs=16)
    cfg = la.EngineConfig(dt=10.0, components=la.KEP | la.J2, cutoff_km=10.0, collision_distance_factor=2.0e5)
    with la.Engine(cfg) as eng:
        p = pop.copy()
        p.aom = np.pi * pop.radius_m * pop.radius_m * 1e-6 / pop.mass_kg
        eng.set_objects(p)
        eng.propagate(2)
        hits, _ = eng.conjunctions(10.0)
        h = ol.sort_pairs(hits)[0]
        ida, idb = int(h["id_a"]), int(h["id_b"])
        r, v = eng.state()
        ia = int(h["idx_a"])
        nf = 12
        fr = np.asarray(h["mid"]) + rng.uniform(-0.3, 0.3, (nf, 3))
        fv = v[ia] + rng.normal(0.0, 0.02, (nf, 3))
        frad = rng.uniform(0.05, 0.5, nf)
        fid = np.arange(5000, 5000 + nf, dtype=np.uint64)
        assert eng.delete_objects([ida, idb]) == 2
        eng.add_objects(rx=fr[:, 0], ry=fr[:, 1], rz=fr[:, 2], vx=fv[:, 0], vy=fv[:, 1], vz=fv[:, 2],
                        radius_m=frad, aom=np.zeros(nf), bc_inv=np.zeros(nf), id=fid,
                        parent=np.full(nf, ida, np.uint64))
        after, _ = eng.conjunctions(10.0)
        st = eng.simulate(3, first_iteration=2, tpcd=1, min_altitude_km=150.0)
        log = eng.fetch_log()
        rs, vs = eng.state()
        ids = eng.ids()
        la.write_conjunctions_csv(tmp_path / "py.csv", log)
    lines = _header(10.0, 10.0, comps=(1, 1, 0, 0, 0, 0, 0, 0))
    for i in range(pop.n):
        lines.append(_particle((pop.rx[i], pop.ry[i], pop.rz[i]), (pop.vx[i], pop.vy[i], pop.vz[i]), i,
                               mass=pop.mass_kg[i], radius=pop.radius_m[i], bc=pop.bc_inv[i], cr=pop.cr[i]))
    lines += ["integrate 2", "functor 10.0 10.0 200000.0 0.05 0.1", f"delete {ida}", f"delete {idb}"]
    for k in range(nf):
        lines.append(_particle(fr[k], fv[k], fid[k], radius=frad[k], parent=ida))
    lines += ["functor 10.0 10.0 200000.0 0.05 0.1", "simulate 2 3 1 150.0", f"conjlog {tmp_path / 'cpp.csv'}",
              "state"]
    out = _run(tmp_path, lines)
    first = sorted((int(x["id_a"]), int(x["id_b"])) for x in hits)
    second = sorted((int(x["id_a"]), int(x["id_b"])) for x in after)
    n_first = len(first)
    assert sorted((a, b) for a, b, _, _ in out["hit"][:n_first]) == first
    assert sorted((a, b) for a, b, _, _ in out["hit"][n_first:]) == second
    assert (ida, idb) in first and (ida, idb) not in second
    assert out["stats"] == [(st["collisions"], st["evasions"], st["remaining"])]
    assert sorted(out["state"]) == sorted(int(i) for i in ids)
    for k, i in enumerate(ids):
        assert np.array_equal(out["state"][int(i)], np.concatenate([rs[k], vs[k]]))

    def rows(path):
        text = path.read_text().splitlines()
        assert text[0] == "Iteration,P1,P2,SquaredDistance"
        return sorted((int(a), int(b), int(c), float(d)) for a, b, c, d in (ln.split(",") for ln in text[1:]))

    assert rows(tmp_path / "cpp.csv") == rows(tmp_path / "py.csv") and len(rows(tmp_path / "py.csv")) > 0


def test_iterate_pairwise_refreshes_the_particles_it_returns(tmp_path):
    """After integrate(false), iteratePairwise brings only the objects its tuples point at up
    to date (no download of the population): their Particle positions and velocities equal
    the device state that a full read of the container returns afterwards."""
    pop = la.generate_population(0, 0x1ADD5 + 10, 3000, planted_pairs=40)
    lines = _header(10.0, 10.0, comps=(1, 1, 0, 0, 0, 0, 0, 0))
    for i in range(pop.n):
        lines.append(_particle((pop.rx[i], pop.ry[i], pop.rz[i]), (pop.vx[i], pop.vy[i], pop.vz[i]), i,
                               mass=pop.mass_kg[i], radius=pop.radius_m[i], bc=pop.bc_inv[i], cr=pop.cr[i]))
    lines += ["integrate 3", "functor 10.0 10.0 200000.0 0.05 0.1", "state"]
    out = _run(tmp_path, lines)
    assert len(out["hit"]) >= 20 and len(out["hitpos"]) >= 20
    for pid, st in out["hitpos"].items():
        assert np.array_equal(st, out["state"][pid]), pid
    moved = [pid for pid in out["hitpos"] if not np.array_equal(out["state"][pid][:3], pop.r()[pid])]
    assert len(moved) == len(out["hitpos"])


def test_remove_particle_spawn_protection(tmp_path):
    """Simulation::removeParticleSpawnProtection through the adapter: fragments of one parent
    are skipped by the functor until their parent ids are reset."""
    rng = np.random.default_rng(4)
    lines = _header(10.0, 10.0)
    centre = np.array([7000.0, 100.0, 50.0])
    for k in range(5):
        lines.append(_particle(centre + rng.uniform(-0.2, 0.2, 3), [0.0, 7.5, 0.01 * k], 100 + k, radius=500.0,
                               parent=42))
    lines += ["functor 10.0 10.0 1.0 0.05 0.1", "unprotect", "functor 10.0 10.0 1.0 0.05 0.1"]
    out = _run(tmp_path, lines)
    assert len(out["hit"]) == 10  # none before the reset, all 10 pairs after it

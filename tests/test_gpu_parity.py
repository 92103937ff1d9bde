"""Parity of the gfx950 engine against the CPU oracle (through the C ABI).

Propagation: FP64 positions agree with the oracle within 1e-9 relative per step.
Conjunctions: the sorted pair sets (ids, indices, d^2, midpoint) are bit-identical
to the oracle's on the same state.
"""
import math

import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu

UMAX = np.iinfo(np.uint64).max


VARIANTS = [1, 2]  # pair-scan pipelines: atomic counting sort, plane tiles


def engine(**kw):
    return la.Engine(la.EngineConfig(**kw))


@pytest.fixture(params=VARIANTS, ids=["atomic_sort", "plane_tiles"])
def variant(request):
    return request.param


def _rel_err(r_gpu, r_ref):
    return float(np.max(np.linalg.norm(r_gpu - r_ref, axis=1) / np.linalg.norm(r_ref, axis=1)))


def _oracle_objects(r, v, pop):
    return ol.ObjectSet(r, v, pop.radius_m, pop.activity, pop.id, pop.parent)


def _pair_equal(a, b):
    """Bit equality of pair arrays, treating NaN d^2 / midpoints as equal to NaN."""
    assert a.shape == b.shape, (a.shape, b.shape)
    for f in ("id_a", "id_b", "idx_a", "idx_b"):
        assert np.array_equal(a[f], b[f]), f
    for f in ("d2", "mid"):
        x, y = a[f], b[f]
        nan = np.isnan(x)
        assert np.array_equal(nan, np.isnan(y)), f
        assert np.array_equal(x[~nan], y[~nan]), f


# ---------------------------------------------------------------------------
# propagation
# ---------------------------------------------------------------------------
def test_propagation_kepler_j2_fp64_per_step():
    """16,384 objects, Kepler + J2, dt = 10 s: 1e-9 relative per step for 100 steps."""
    pop = la.generate_population(0, 0x1ADD5, 16384)
    mask = la.KEP | la.J2
    with engine(dt=10.0, components=mask, cutoff_km=10.0) as eng:
        eng.set_objects(pop)
        r, v = pop.r(), pop.v()
        t = 0.0
        for step in range(100):
            eng.propagate(1)
            r, v = ol.propagate(r, v, t, 10.0, mask, 1)
            t += 10.0
            rg, vg = eng.state()
            assert _rel_err(rg, r) < 1e-9, step
            assert _rel_err(vg, v) < 1e-9, step
            # continue the oracle from the GPU state: the bound is per step
            r, v = rg, vg


def test_propagation_full_model_fp64():
    """Full force model, 4,096 objects of the debris-tail workload, 200 steps."""
    pop = la.generate_population(1, 0x1ADD6, 4096)
    with engine(dt=10.0, components=la.ALL_COMPONENTS, cutoff_km=10.0) as eng:
        eng.set_objects(pop)
        r, v = pop.r(), pop.v()
        for chunk in range(4):
            eng.propagate(50)
            r, v = ol.propagate(r, v, chunk * 500.0, 10.0, 0xFF, 50, pop.aom, pop.bc_inv, pop.cr, pop.activity)
            rg, vg = eng.state()
            assert _rel_err(rg, r) < 1e-9 * 50
            r, v = rg, vg


@pytest.mark.parametrize("mask", [la.J2, la.C22, la.S22, la.SOL, la.LUN, la.SRP, la.DRAG])
def test_each_component_fp64(mask):
    """One component at a time (plus Kepler), 1,024 objects, 20 steps from t0 = 1e8 s."""
    pop = la.generate_population(1, 77, 1024)
    pop.activity[::7] = la.EVASIVE_PRESERVING
    m = la.KEP | mask
    with engine(dt=10.0, t0=1.0e8, components=m, cutoff_km=10.0) as eng:
        eng.set_objects(pop)
        eng.propagate(20)
        rg, vg = eng.state()
    r, v = ol.propagate(pop.r(), pop.v(), 1.0e8, 10.0, m, 20, pop.aom, pop.bc_inv, pop.cr, pop.activity)
    assert _rel_err(rg, r) < 1e-11
    assert _rel_err(vg, v) < 1e-11


def test_no_components_straight_line():
    pop = la.generate_population(0, 3, 256)
    with engine(dt=10.0, components=0, cutoff_km=10.0) as eng:
        eng.set_objects(pop)
        eng.propagate(7)
        rg, vg = eng.state()
    assert np.allclose(rg, pop.r() + 70.0 * pop.v(), rtol=1e-14, atol=1e-9)
    assert np.array_equal(vg, pop.v())


@pytest.mark.parametrize("precision,limit_km", [(la.SPLIT, 1e-5), (la.SPLIT_ALL32, 5e-2)])
def test_split_precision_error_after_one_day(precision, limit_km):
    """Split forms vs FP64 after 8,640 steps of 10 s with the full force model."""
    pop = la.generate_population(1, 0x1ADD6, 4096)
    out = {}
    for prec in (la.FP64, precision):
        with engine(dt=10.0, components=la.ALL_COMPONENTS, precision=prec, cutoff_km=10.0) as eng:
            eng.set_objects(pop)
            eng.propagate(8640)
            out[prec] = eng.state()[0]
    err = np.linalg.norm(out[precision] - out[la.FP64], axis=1)
    keep = np.linalg.norm(out[la.FP64], axis=1) > 6378.1363 + 150  # burnt-up objects excluded
    print(f"precision {precision}: max {err[keep].max() * 1e5:.4f} cm, median {np.median(err[keep]) * 1e5:.4f} cm")
    assert err[keep].max() < limit_km


# ---------------------------------------------------------------------------
# conjunctions
# ---------------------------------------------------------------------------
def _cloud(n, seed, box=200.0, centre=(6000.0, 3000.0, 1500.0), vel_sigma=5.0):
    rng = np.random.default_rng(seed)
    r = np.array(centre) + rng.uniform(-box / 2, box / 2, (n, 3))
    v = np.array([0.0, 4.0, 5.0]) + rng.normal(0, vel_sigma, (n, 3))
    pop = la.Population(r[:, 0].copy(), r[:, 1].copy(), r[:, 2].copy(), v[:, 0].copy(), v[:, 1].copy(),
                        v[:, 2].copy(), rng.uniform(0.01, 600.0, n), np.ones(n), np.zeros(n), np.zeros(n),
                        np.ones(n))
    pop.activity = rng.choice([0, 0, 0, 1, 2], n).astype(np.int32)
    pop.id = rng.permutation(n).astype(np.uint64)
    pop.parent = np.where(rng.random(n) < 0.1, rng.integers(0, 50, n), UMAX).astype(np.uint64)
    return pop


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
@pytest.mark.parametrize("mode,thr", [(la.THRESHOLD_RADIUS_SCALED, 1.0), (la.THRESHOLD_ABSOLUTE, 0.6)])
def test_conjunctions_bit_identical(precision, mode, thr, variant):
    pop = _cloud(20000, 11)
    cutoff, dt = 2.0, 10.0
    with engine(scan_variant=variant, dt=dt, cutoff_km=cutoff, precision=precision, threshold_mode=mode, threshold_km=thr,
                collision_distance_factor=1.5, min_detection_radius_m=300.0, evasion_cutoff_km=0.8) as eng:
        eng.set_objects(pop)
        hits, ev = eng.conjunctions(dt)
    objs = _oracle_objects(pop.r(), pop.v(), pop)
    oh, oe = ol.pairs(objs, ol.functor_cfg(cutoff, dt, 1.5, 300.0, 0.8, mode, thr), method="grid")
    assert oh.shape[0] > 20 and oe.shape[0] > 20
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)


def test_conjunctions_after_propagation_match_bruteforce(variant):
    """LEO population with planted pairs: GPU scan of the GPU state == oracle brute force.
    After 3 steps the co-orbital half of the 64 planted pairs is still within 1 km."""
    pop = la.generate_population(0, 0x1ADD5, 4096, planted_pairs=64)
    with engine(scan_variant=variant, dt=10.0, components=la.KEP | la.J2, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=1.0) as eng:
        eng.set_objects(pop)
        eng.propagate(3)
        r, v = eng.state()
        hits, ev = eng.conjunctions()
    objs = _oracle_objects(r, v, pop)
    oh, oe = ol.pairs(objs, ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0), method="brute")
    assert oh.shape[0] >= 30
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)


@pytest.mark.parametrize("name", ["three_particles", "mix_activity_states"])
def test_reference_known_answers_on_gpu(golden_cases, name, variant):
    case = golden_cases[name]
    f = case["functor"]
    parts = case["particles"]
    n = len(parts)
    with engine(scan_variant=variant, dt=f["dt"], cutoff_km=f["cutoff"], collision_distance_factor=f["cdf"],
                min_detection_radius_m=f["min_det"], evasion_cutoff_km=f["evasion"]) as eng:
        eng.set_objects(rx=[p["r"][0] for p in parts], ry=[p["r"][1] for p in parts],
                        rz=[p["r"][2] for p in parts], vx=[p["v"][0] for p in parts],
                        vy=[p["v"][1] for p in parts], vz=[p["v"][2] for p in parts],
                        radius_m=[p["radius"] for p in parts], activity=[p["activity"] for p in parts],
                        id=[p["id"] for p in parts], parent=[p["parent"] for p in parts])
        hits, _ = eng.conjunctions(f["dt"])
    if name == "three_particles":
        got = {(int(h["idx_a"]), int(h["idx_b"])): (float(h["d2"]), list(h["mid"])) for h in hits}
        exp = {(e["a"], e["b"]): (e["d2"], e["mid"]) for e in case["expected_collisions"]}
        assert got == exp
    else:
        got = sorted((int(h["id_a"]), int(h["id_b"])) for h in hits)
        assert got == sorted(tuple(p) for p in case["expected_id_pairs"])
        assert n == 8 and np.all(np.isnan(hits["d2"]))


def test_linear_interpolation_vectors_on_gpu(golden_cases, variant):
    case = golden_cases["linear_interpolation"]
    f = case["functor"]
    for vec in case["vectors"]:
        with engine(scan_variant=variant, dt=vec["dt"], cutoff_km=f["cutoff"], collision_distance_factor=f["cdf"],
                    min_detection_radius_m=f["min_det"], evasion_cutoff_km=f["evasion"]) as eng:
            eng.set_objects(rx=[vec["x1"][0], vec["x2"][0]], ry=[vec["x1"][1], vec["x2"][1]],
                            rz=[vec["x1"][2], vec["x2"][2]], vx=[vec["v1"][0], vec["v2"][0]],
                            vy=[vec["v1"][1], vec["v2"][1]], vz=[vec["v1"][2], vec["v2"][2]],
                            radius_m=[case["radius"]] * 2)
            hits, _ = eng.conjunctions(vec["dt"])
        assert hits.shape[0] == 1
        assert float(hits[0]["d2"]) == vec["d2"] and list(hits[0]["mid"]) == vec["mid"]


def test_digest_matches_oracle(variant):
    pop = _cloud(20000, 12)
    with engine(scan_variant=variant, dt=10.0, cutoff_km=2.0, precision=la.SPLIT, min_detection_radius_m=300.0) as eng:
        eng.set_objects(pop)
        n, d, ne, e = eng.conjunction_digest(10.0)
        hits, ev = eng.conjunctions(10.0)
    objs = _oracle_objects(pop.r(), pop.v(), pop)
    oh, oe = ol.pairs(objs, ol.functor_cfg(2.0, 10.0, 1.0, 300.0, 0.1), method="grid")
    assert n == oh.shape[0] and ne == oe.shape[0]
    assert d == ol.digest(oh) and e == ol.digest(oe)
    assert hits.shape[0] == n


def test_dense_cell_all_pairs(variant):
    """1,000 objects inside 50 m: every pair is scanned once (499,500 candidates)."""
    n = 1000
    rng = np.random.default_rng(3)
    r = np.array([7000.0, 0.0, 0.0]) + rng.uniform(-0.025, 0.025, (n, 3))
    v = np.array([0.0, 7.5, 0.0]) + rng.normal(0, 1e-3, (n, 3))
    with engine(scan_variant=variant, dt=10.0, cutoff_km=5.0, threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=0.5) as eng:
        eng.set_objects(rx=r[:, 0], ry=r[:, 1], rz=r[:, 2], vx=v[:, 0], vy=v[:, 1], vz=v[:, 2],
                        radius_m=np.full(n, 0.1))
        hits, _ = eng.conjunctions()
        cnt, dig, _, _ = eng.conjunction_digest()
    assert hits.shape[0] == n * (n - 1) // 2 == cnt
    objs = ol.ObjectSet(r, v, np.full(n, 0.1))
    oh, _ = ol.pairs(objs, ol.functor_cfg(5.0, 10.0, threshold_mode=1, threshold_km=0.5), method="grid")
    _pair_equal(ol.sort_pairs(hits), oh)
    assert dig == ol.digest(oh)


def test_cell_boundaries_and_tiny_populations(variant):
    cutoff = 1.0
    cell = 2.0 * cutoff * (1 + 1e-9)
    # objects exactly on and around cell faces, within and just outside the cutoff
    base = np.array([cell * 3000, cell * -17, cell * 5])
    offs = np.array([[0, 0, 0], [cutoff, 0, 0], [-cutoff, 0, 0], [0, cutoff, 0], [0, 0, -cutoff],
                     [cutoff * 1.0000001, 0, 0], [0.5 * cell, 0.5 * cell, 0.5 * cell], [-0.5 * cell, 0, 0],
                     [0.7, 0.7, 0.0], [0.0, -0.7, 0.71]])
    r = base + offs
    v = np.zeros_like(r) + np.array([0.0, 7.0, 0.0]) + np.arange(len(r))[:, None] * 1e-3
    n = r.shape[0]
    with engine(scan_variant=variant, dt=10.0, cutoff_km=cutoff, collision_distance_factor=1e6) as eng:
        eng.set_objects(rx=r[:, 0], ry=r[:, 1], rz=r[:, 2], vx=v[:, 0], vy=v[:, 1], vz=v[:, 2],
                        radius_m=np.full(n, 1.0))
        hits, _ = eng.conjunctions()
        # n = 1 and n = 0
        eng.set_objects(rx=r[:1, 0], ry=r[:1, 1], rz=r[:1, 2], vx=v[:1, 0], vy=v[:1, 1], vz=v[:1, 2],
                        radius_m=[1.0])
        h1, e1 = eng.conjunctions()
        eng.set_objects(rx=[], ry=[], rz=[], vx=[], vy=[], vz=[], radius_m=[])
        h0, e0 = eng.conjunctions()
        eng.propagate(3)
    oh, _ = ol.pairs(ol.ObjectSet(r, v, np.full(n, 1.0)), ol.functor_cfg(cutoff, 10.0, 1e6), method="brute")
    assert oh.shape[0] > 5
    _pair_equal(ol.sort_pairs(hits), oh)
    assert h1.shape[0] == e1.shape[0] == h0.shape[0] == e0.shape[0] == 0


# ---------------------------------------------------------------------------
# the simulation loop
# ---------------------------------------------------------------------------
def test_simulate_matches_oracle_step_by_step(variant):
    """integrate + detection every step (tpcd = 1): per-iteration id pair sets equal the oracle's."""
    pop = la.generate_population(0, 0x1ADD5, 8192, planted_pairs=48)
    mask, dt, steps = la.KEP | la.J2, 10.0, 12
    with engine(scan_variant=variant, dt=dt, components=mask, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=1.0) as eng:
        eng.set_objects(pop)
        stats = eng.simulate(steps, first_iteration=0, tpcd=1)
        log = eng.fetch_log()
    assert stats["logged"] == log.shape[0] == stats["collisions"] + stats["evasions"]
    r, v = pop.r(), pop.v()
    total = 0
    for it in range(steps):
        r, v = ol.propagate(r, v, it * dt, dt, mask, 1)
        oh, _ = ol.pairs(_oracle_objects(r, v, pop), ol.functor_cfg(10.0, dt, threshold_mode=1, threshold_km=1.0))
        mine = log[(log["iteration"] == it) & (log["evaded"] == 0)]["pair"]
        got = sorted(zip(mine["id_a"].tolist(), mine["id_b"].tolist()))
        assert got == sorted(zip(oh["id_a"].tolist(), oh["id_b"].tolist())), it
        ms = ol.sort_pairs(mine)
        assert np.allclose(ms["d2"], oh["d2"], rtol=1e-6, atol=1e-12)
        total += oh.shape[0]
    assert total >= 48


def _cluster_population(seed, n=40000, cluster=2500):
    """A LEO population with planted pairs plus `cluster` objects within a 4 km cube moving
    almost together: thousands of conjunctions every step."""
    pop = la.generate_population(1, 0x1ADD5 + seed, n, planted_pairs=200)
    rng = np.random.default_rng(seed)
    c = np.array([pop.rx[0], pop.ry[0], pop.rz[0]])
    for i in range(1, cluster):
        pop.rx[i], pop.ry[i], pop.rz[i] = c + rng.uniform(-4.0, 4.0, 3)
        pop.vx[i], pop.vy[i], pop.vz[i] = pop.vx[0], pop.vy[0], pop.vz[0] + rng.uniform(-1e-3, 1e-3)
    return pop


def _d2_bits(x):
    return int(np.array(x, np.float64).view(np.int64))


@pytest.mark.parametrize("steps,tpcd", [(30, 1), (31, 1), (32, 1), (33, 1), (29, 2), (40, 3)])
@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_simulate_pipeline_equals_step_by_step(steps, tpcd, precision):
    """ladds_simulate's eager loop cycles through three buffers, with the propagation of step
    k+2, the partition of step k+1 and the pair scan of step k in flight together.  Its pair
    log (iteration, ids, d^2 bits, evaded) and its final state equal, bit for bit, those of
    the same steps run one call at a time (ladds_propagate, then ladds_conjunctions at each
    detection iteration).  The step counts leave the final state in each of the buffers
    (three, or four with two pair-scan streams); detection every 1, 2 and 3 iterations."""
    pop = _cluster_population(19)
    cfg = dict(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 22)
    with engine(**cfg) as eng:
        eng.set_objects(pop)
        stats = eng.simulate(steps, first_iteration=0, tpcd=tpcd)
        log = eng.fetch_log()
        r1, v1 = eng.state()
    ref = []
    with engine(**cfg) as eng:
        eng.set_objects(pop)
        for it in range(steps):
            eng.propagate(1)
            if it % tpcd == 0:
                hits, ev = eng.conjunctions(dt_interp=10.0 * tpcd)
                for arr, evaded in ((hits, 0), (ev, 1)):
                    ref += [(it, int(p["id_a"]), int(p["id_b"]), _d2_bits(p["d2"]), evaded) for p in arr]
        r0, v0 = eng.state()
    got = [(int(e["iteration"]), int(e["pair"]["id_a"]), int(e["pair"]["id_b"]), _d2_bits(e["pair"]["d2"]),
            int(e["evaded"])) for e in log]
    assert stats["logged"] == len(got) > 1000 * ((steps + tpcd - 1) // tpcd)
    assert sorted(got) == sorted(ref)
    assert np.array_equal(r1, r0) and np.array_equal(v1, v0)


@pytest.mark.parametrize("streams", [1, 2])
@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_pair_streams_keep_the_pipeline_result(precision, streams):
    """ladds_simulate's pair scans on one stream, or alternating over two so that one starts
    while the previous drains (ladds_tuning.pair_streams): the pair log and the final state
    equal, bit for bit, those of the same steps run one call at a time; detection every 1
    and 2 iterations, step counts ending in each pair stream."""
    pop = _cluster_population(29)
    cfg = dict(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 22)
    for steps, tpcd in ((24, 1), (25, 1), (27, 2)):
        with engine(**cfg) as eng:
            eng.set_objects(pop)
            t = eng.autotune(repetitions=1)
            eng.set_tuning(dict(t, pair_streams=streams))
            eng.simulate(steps, first_iteration=0, tpcd=tpcd)
            log = eng.fetch_log()
            r1, v1 = eng.state()
        ref = []
        with engine(**cfg) as eng:
            eng.set_objects(pop)
            for it in range(steps):
                eng.propagate(1)
                if it % tpcd == 0:
                    hits, ev = eng.conjunctions(dt_interp=10.0 * tpcd)
                    for arr, evaded in ((hits, 0), (ev, 1)):
                        ref += [(it, int(p["id_a"]), int(p["id_b"]), _d2_bits(p["d2"]), evaded) for p in arr]
            r0, v0 = eng.state()
        got = [(int(e["iteration"]), int(e["pair"]["id_a"]), int(e["pair"]["id_b"]), _d2_bits(e["pair"]["d2"]),
                int(e["evaded"])) for e in log]
        same = sorted(got) == sorted(ref)
        assert len(got) > 1000 and same, (steps, tpcd, len(got), len(ref))
        assert np.array_equal(r1, r0) and np.array_equal(v1, v0), (steps, tpcd)


@pytest.mark.parametrize("graph", [0, 1])
def test_split_calls_reuse_the_staged_environments(graph):
    """Sun, Moon and Earth-rotation environments depend only on the clock (t0, dt, step), so a
    call reuses the segment an earlier call staged while it covers the call's steps, at the
    offset of the current step.  Three steps of ladds_propagate, then two ladds_simulate calls
    (4 and 5 steps, each starting inside the staged segment), end in the state one 12-step
    simulate reaches, bit for bit, and log the same pairs from iteration 3 on."""
    pop = _cluster_population(23)
    cfg = dict(dt=10.0, components=la.ALL_COMPONENTS, precision=la.SPLIT, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 22, step_graph=graph)

    def keyed(log):
        return sorted((int(e["iteration"]), int(e["pair"]["id_a"]), int(e["pair"]["id_b"]),
                       _d2_bits(e["pair"]["d2"]), int(e["evaded"])) for e in log if int(e["iteration"]) >= 3)

    with engine(**cfg) as eng:
        eng.set_objects(pop)
        eng.simulate(12, first_iteration=0, tpcd=1)
        whole = keyed(eng.fetch_log())
        r0, v0 = eng.state()
    with engine(**cfg) as eng:
        eng.set_objects(pop)
        eng.propagate(3)
        eng.simulate(4, first_iteration=3, tpcd=1)
        eng.simulate(5, first_iteration=7, tpcd=1)
        parts = keyed(eng.fetch_log())
        r1, v1 = eng.state()
    assert len(whole) > 1000
    same = parts == whole  # (a plain list compare: pytest's diff of 10^4 entries takes minutes)
    assert same, (len(parts), len(whole), sorted(set(parts) ^ set(whole))[:5])
    assert np.array_equal(r1, r0) and np.array_equal(v1, v0)


def test_timesteps_per_collision_detection(variant):
    """SimulationTest.cpp:144-193: head-on pair, tpcd = 2 -> 0 / 0 / 1 conjunctions."""
    re = 6378.1363
    alt = re + 150.0
    for iterations, expected in ((1, 0), (2, 0), (3, 1)):
        with engine(scan_variant=variant, dt=1.0, components=la.KEP, cutoff_km=80.0) as eng:
            eng.set_objects(rx=[alt + 100.0, alt + 100.0], ry=[20.0, -20.0], rz=[0.0, 0.0], vx=[0.0, 0.0],
                            vy=[-10.0, 10.0], vz=[0.0, 0.0], radius_m=[1.0, 1.0], id=[0, 1])
            stats = eng.simulate(iterations, first_iteration=0, tpcd=2, min_altitude_km=150.0)
        assert stats["collisions"] == expected, iterations


def test_burn_up(variant):
    """SimulationTest.cpp:102-136: the object heading to Earth burns up after one iteration."""
    re = 6378.1363
    crit = re + 150.0
    with engine(scan_variant=variant, dt=1.0, components=la.KEP, cutoff_km=80.0) as eng:
        eng.set_objects(rx=[crit + 1.0, 0.0], ry=[0.0, crit + 1.0], rz=[0.0, 0.0], vx=[-10.0, 0.0],
                        vy=[0.0, 10.0], vz=[0.0, 0.0], radius_m=[1.0, 1.0], id=[0, 1])
        stats = eng.simulate(1, first_iteration=0, tpcd=1, min_altitude_km=150.0)
        assert stats["remaining"] == 1 == eng.n
        assert eng.ids().tolist() == [1]


def test_delete_burnups_compacts_in_order():
    pop = la.generate_population(0, 21, 3000)
    r = pop.r()
    low = np.arange(0, 3000, 7)
    for k in low:
        r[k] *= (6378.1363 + 100.0) / np.linalg.norm(r[k])
    pop.rx, pop.ry, pop.rz = r[:, 0].copy(), r[:, 1].copy(), r[:, 2].copy()
    with engine(dt=10.0, cutoff_km=10.0) as eng:
        eng.set_objects(pop)
        removed = eng.delete_burnups(150.0)
        ids = eng.ids()
        rg, _ = eng.state()
    keep = np.setdiff1d(np.arange(3000), low)
    assert removed == low.size
    assert ids.tolist() == keep.tolist()
    assert np.array_equal(rg, r[keep])


def test_log_overflow_is_reported(variant):
    pop = _cloud(5000, 13, box=20.0, vel_sigma=1e-3)
    with engine(scan_variant=variant, dt=10.0, cutoff_km=2.0, collision_distance_factor=1e6, log_capacity=1024) as eng:
        eng.set_objects(pop)
        stats = eng.simulate(1, tpcd=1)
        assert stats["collisions"] + stats["evasions"] > 1024
        with pytest.raises(la.LaddsError):
            eng.fetch_log()


def test_plane_overflow_uses_scratch(variant):
    """6,000 objects inside 2 km: one plane far beyond the LDS capacity of the tile kernel."""
    n = 6000
    rng = np.random.default_rng(5)
    r = np.array([7000.0, 100.0, -50.0]) + rng.uniform(-1.0, 1.0, (n, 3))
    v = np.array([0.0, 7.5, 0.1]) + rng.normal(0, 0.05, (n, 3))
    with engine(scan_variant=variant, dt=10.0, cutoff_km=5.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=0.02) as eng:
        eng.set_objects(rx=r[:, 0], ry=r[:, 1], rz=r[:, 2], vx=v[:, 0], vy=v[:, 1], vz=v[:, 2],
                        radius_m=np.full(n, 0.1))
        hits, _ = eng.conjunctions()
        cnt, dig, _, _ = eng.conjunction_digest()
    objs = ol.ObjectSet(r, v, np.full(n, 0.1))
    oh, _ = ol.pairs(objs, ol.functor_cfg(5.0, 10.0, threshold_mode=1, threshold_km=0.02), method="grid")
    assert oh.shape[0] > 10
    _pair_equal(ol.sort_pairs(hits), oh)
    assert cnt == oh.shape[0] and dig == ol.digest(oh)


def test_variants_agree_on_leo_population():
    """Both pipelines and both forms report the same pairs after a few steps of the full model."""
    pop = la.generate_population(1, 99, 60000, planted_pairs=200)
    got = {}
    for v in VARIANTS:
        for prec in (la.FP64, la.SPLIT):
            with engine(scan_variant=v, precision=prec, dt=10.0, components=la.ALL_COMPONENTS, cutoff_km=10.0,
                        threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0) as eng:
                eng.set_objects(pop)
                st = eng.simulate(5, tpcd=1, min_altitude_km=150.0)
                log = eng.fetch_log()
            keys = sorted(zip(log["iteration"].tolist(), log["pair"]["id_a"].tolist(), log["pair"]["id_b"].tolist()))
            got[(v, prec)] = (keys, st["candidates"] > 0)
    ref = got[(1, la.FP64)][0]
    assert len(ref) >= 100
    for k, (keys, _) in got.items():
        assert keys == ref, k


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_autotune_keeps_the_pair_set(precision):
    """ladds_autotune changes cell size, plane shape and workgroup size; the pair set stays
    bit-identical to the oracle's."""
    pop = la.generate_population(1, 0x1ADD5 + 12, 65536, planted_pairs=256)
    cfg = dict(dt=10.0, components=la.KEP | la.J2, precision=precision, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0)
    with engine(**cfg) as eng:
        eng.set_objects(pop)
        eng.propagate(3)
        t = eng.autotune(repetitions=1)
        assert t["tried"] > 10 and t["scan_us"] <= t["default_us"]
        hits, ev = eng.conjunctions()
        r, v = eng.state()
    oh, oe = ol.pairs(_oracle_objects(r, v, pop), ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0))
    assert oh.shape[0] > 100
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)


@pytest.mark.parametrize("chunks", [2, 3, 16])
@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_chunked_planes_keep_the_pair_set(chunks, precision):
    """Several workgroups per plane (each sorting the plane itself) still report every pair
    exactly once, including a dense cell where every object is close to every other."""
    pop = la.generate_population(1, 0x1ADD5 + 13, 30000, planted_pairs=200)
    c = np.array([pop.rx[0], pop.ry[0], pop.rz[0]])
    rng = np.random.default_rng(13)
    for i in range(1, 1500):  # a dense cluster of 1500 objects within ~0.3 km (three chunks of 512)
        pop.rx[i], pop.ry[i], pop.rz[i] = c + rng.uniform(-0.15, 0.15, 3)
        pop.vx[i], pop.vy[i], pop.vz[i] = pop.vx[0], pop.vy[0], pop.vz[0] + rng.uniform(-1e-3, 1e-3)
    with engine(dt=10.0, precision=precision, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=1.0, scan_chunks=chunks) as eng:
        eng.set_objects(pop)
        hits, ev = eng.conjunctions()
    objs = _oracle_objects(pop.r(), pop.v(), pop)
    oh, oe = ol.pairs(objs, ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0))
    assert oh.shape[0] > 1500 * 1499 // 2
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_lds_and_arena_planes_in_one_scan(precision):
    """A sparse LEO population plus a 12,000-object cluster: the cluster's plane pair exceeds
    the LDS of the pair kernel (sorted in the global arena) while every other plane is
    sorted in LDS, in the same launch; the pair set equals the oracle's."""
    pop = la.generate_population(0, 0x1ADD5 + 31, 20000, planted_pairs=40)
    rng = np.random.default_rng(9)
    m = 12000
    c = np.array([7000.0, 50.0, 20.0]) + rng.uniform(-1.5, 1.5, (m, 3))
    cv = np.array([0.0, 7.5, 0.0]) + rng.normal(0, 0.01, (m, 3))
    n = pop.n + m
    r = np.concatenate([pop.r(), c])
    v = np.concatenate([pop.v(), cv])
    radius = np.concatenate([pop.radius_m, np.full(m, 0.05)])
    with engine(dt=10.0, components=la.KEP, precision=precision, cutoff_km=10.0,
                threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=0.05) as eng:
        eng.set_objects(rx=r[:, 0], ry=r[:, 1], rz=r[:, 2], vx=v[:, 0], vy=v[:, 1], vz=v[:, 2], radius_m=radius)
        hits, ev = eng.conjunctions()
        cnt, dig, _, _ = eng.conjunction_digest()
    objs = ol.ObjectSet(r, v, radius)
    oh, oe = ol.pairs(objs, ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=0.05), method="grid")
    assert oh.shape[0] > 1000 and n == 32000
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)
    assert cnt == oh.shape[0] and dig == ol.digest(oh)


@pytest.mark.parametrize("rows,width", [(2, 1), (2, 2), (1, 2)])
@pytest.mark.parametrize("plane_bits,chunks", [(10, 1), (4, 1), (7, 4), (4, 3)])
@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_two_cell_buckets_keep_the_pair_set(rows, width, plane_bits, chunks, precision):
    """Buckets two cells tall (row_cells = 2: each object scans two rows of the next plane,
    picked by the half of its row it lies in) and/or two cells wide (bucket_cells = 2: two
    buckets of a row, by the half of its bucket), one workgroup per plane and chunked planes
    (index rule), down to 4 buckets per plane (two rows or two buckets a row: the fold puts
    one on both sides of the other).  A LEO population with planted pairs plus a
    2,500-object cluster: pair list and digest equal the oracle's."""
    pop = la.generate_population(1, 0x1ADD5 + 18, 40000, planted_pairs=300)
    rng = np.random.default_rng(18)
    c = np.array([pop.rx[0], pop.ry[0], pop.rz[0]])
    for i in range(1, 2500):
        pop.rx[i], pop.ry[i], pop.rz[i] = c + rng.uniform(-4.0, 4.0, 3)
        pop.vx[i], pop.vy[i], pop.vz[i] = pop.vx[0], pop.vy[0], pop.vz[0] + rng.uniform(-1e-3, 1e-3)
    cell = 10.11 if precision == la.SPLIT else 10.0 * (1 + 1e-8)  # >= the search radius
    shape = {"cell_size_km": cell, "plane_width_km": cell, "plane_bits": plane_bits, "pair_threads": 256,
             "tile_size": 8192, "chunks": chunks, "workgroups_per_cu": 0, "row_cells": rows,
             "bucket_cells": width}
    with engine(dt=10.0, precision=precision, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=1.0) as eng:
        eng.set_objects(pop)
        eng.set_tuning(shape)
        hits, ev = eng.conjunctions()
        cnt, dig, ecnt, edig = eng.conjunction_digest()
    oh, oe = ol.pairs(_oracle_objects(pop.r(), pop.v(), pop),
                      ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0))
    assert oh.shape[0] > 1000
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)
    assert cnt == oh.shape[0] and dig == ol.digest(oh)
    assert ecnt == oe.shape[0] and edig == ol.digest(oe)


@pytest.mark.parametrize("threads,plane_bits,chunks,cells", [(256, 10, 1, 2), (128, 4, 1, 1), (384, 7, 4, 2),
                                                                (1024, 10, 3, 1)])
@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_persistent_pair_kernel_keeps_the_pair_set(threads, plane_bits, chunks, cells, precision):
    """The persistent pair kernel (as many workgroups as the CUs hold, each taking plane pairs
    from its XCD's queue until none is left) over one workgroup size per case, small and
    large bucket counts, chunked planes and the cluster's plane pair sorted in the global
    arena: pair list and digest equal the oracle's, twice in a row on one context (the
    partition clears the queues for each scan)."""
    pop = _cluster_population(20)
    cell = 10.11 if precision == la.SPLIT else 10.0 * (1 + 1e-8)  # >= the search radius
    shape = {"cell_size_km": cell, "plane_width_km": cell, "plane_bits": plane_bits, "pair_threads": threads,
             "tile_size": 8192, "chunks": chunks, "workgroups_per_cu": 0, "plane_order": 1, "row_cells": cells,
             "bucket_cells": cells, "persistent": 1}
    oh, oe = ol.pairs(_oracle_objects(pop.r(), pop.v(), pop),
                      ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0))
    assert oh.shape[0] > 1000
    with engine(dt=10.0, precision=precision, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=1.0) as eng:
        eng.set_objects(pop)
        eng.set_tuning(shape)
        for _ in range(2):
            hits, ev = eng.conjunctions()
            cnt, dig, ecnt, edig = eng.conjunction_digest()
            _pair_equal(ol.sort_pairs(hits), oh)
            _pair_equal(ol.sort_pairs(ev), oe)
            assert cnt == oh.shape[0] and dig == ol.digest(oh)
            assert ecnt == oe.shape[0] and edig == ol.digest(oe)


@pytest.mark.parametrize("threads", [128, 256, 384, 512, 1024])
@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT])
def test_every_workgroup_size_keeps_the_pair_set(threads, precision):
    """The plane-tile pair kernel at each workgroup size (each has its own register budget
    and its own split between the register-resident and the two-pass gather): a LEO
    population with planted pairs plus a 2,500-object cluster, whose plane pair takes the
    two-pass gather below 512 threads.  Pair list and digest equal the oracle's."""
    pop = la.generate_population(1, 0x1ADD5 + 17, 40000, planted_pairs=300)
    rng = np.random.default_rng(17)
    c = np.array([pop.rx[0], pop.ry[0], pop.rz[0]])
    for i in range(1, 2500):
        pop.rx[i], pop.ry[i], pop.rz[i] = c + rng.uniform(-4.0, 4.0, 3)
        pop.vx[i], pop.vy[i], pop.vz[i] = pop.vx[0], pop.vy[0], pop.vz[0] + rng.uniform(-1e-3, 1e-3)
    with engine(dt=10.0, precision=precision, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                threshold_km=1.0, block_size=threads) as eng:
        eng.set_objects(pop)
        hits, ev = eng.conjunctions()
        cnt, dig, ecnt, edig = eng.conjunction_digest()
    oh, oe = ol.pairs(_oracle_objects(pop.r(), pop.v(), pop),
                      ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0))
    assert oh.shape[0] > 1000
    _pair_equal(ol.sort_pairs(hits), oh)
    _pair_equal(ol.sort_pairs(ev), oe)
    assert cnt == oh.shape[0] and dig == ol.digest(oh)
    assert ecnt == oe.shape[0] and edig == ol.digest(oe)

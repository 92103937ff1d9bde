"""Pins the oracle's conjunction predicate against the reference's own known answers.

Mirrors tests/testladds/CollisionFunctorTest.cpp and the container-independence
invariant of tests/testladds/CollisionFunctorIntegrationTest.cpp.
"""
import numpy as np
import pytest

import oracle_lib as ol


def _objset(particles):
    r = [p["r"] for p in particles]
    v = [p["v"] for p in particles]
    return ol.ObjectSet(r, v, [p["radius"] for p in particles], [p["activity"] for p in particles],
                        [p["id"] for p in particles], [p["parent"] for p in particles])


def test_three_particles(golden_cases):
    case = golden_cases["three_particles"]
    f = case["functor"]
    objs = _objset(case["particles"])
    cfg = ol.functor_cfg(f["cutoff"], f["dt"], f["cdf"], f["min_det"], f["evasion"])
    hits, ev = ol.pairs(objs, cfg, method="brute")
    assert ev.shape[0] == 0
    got = {(int(h["idx_a"]), int(h["idx_b"])): (float(h["d2"]), list(h["mid"])) for h in hits}
    exp = {(e["a"], e["b"]): (e["d2"], e["mid"]) for e in case["expected_collisions"]}
    assert got.keys() == exp.keys()
    for k in exp:  # the reference compares with exact equality (gmock UnorderedElementsAreArray)
        assert got[k][0] == exp[k][0]
        assert got[k][1] == exp[k][1]


def test_mix_activity_states(golden_cases):
    case = golden_cases["mix_activity_states"]
    f = case["functor"]
    objs = _objset(case["particles"])
    cfg = ol.functor_cfg(f["cutoff"], f["dt"], f["cdf"], f["min_det"], f["evasion"])
    hits, _ = ol.pairs(objs, cfg, method="brute")
    got = sorted((int(h["id_a"]), int(h["id_b"])) for h in hits)
    assert got == sorted(tuple(p) for p in case["expected_id_pairs"])
    # zero relative velocity: t = 0/0 = NaN passes std::clamp and every threshold test
    assert np.all(np.isnan(hits["d2"]))


@pytest.mark.parametrize("cdf", [1.0, 2.0])
def test_collision_distance_factor(golden_cases, cdf):
    case = golden_cases["collision_distance_factor"]
    f = case["functor"]
    objs = _objset(case["particles"])
    hits, _ = ol.pairs(objs, ol.functor_cfg(f["cutoff"], f["dt"], cdf, f["min_det"], f["evasion"]), method="brute")
    assert hits.shape[0] == case["expected_count_by_cdf"][str(cdf)]


def test_linear_interpolation_vectors(golden_cases):
    case = golden_cases["linear_interpolation"]
    f = case["functor"]
    for vec in case["vectors"]:
        parts = [{"r": vec["x1"], "v": vec["v1"], "id": 0, "activity": 0, "radius": case["radius"],
                  "parent": 2**64 - 1},
                 {"r": vec["x2"], "v": vec["v2"], "id": 1, "activity": 0, "radius": case["radius"],
                  "parent": 2**64 - 1}]
        objs = _objset(parts)
        hits, _ = ol.pairs(objs, ol.functor_cfg(f["cutoff"], vec["dt"], f["cdf"], f["min_det"], f["evasion"]),
                           method="brute")
        assert hits.shape[0] == 1
        assert float(hits[0]["d2"]) == vec["d2"]
        assert list(hits[0]["mid"]) == vec["mid"]


def _closest_approach_closed_form(ri, vi, rj, vj, dt):
    """Textbook closest approach of two straight lines over [0, dt] of the back-extrapolated
    segment, computed in extended precision with numpy longdouble."""
    L = np.longdouble
    ri, vi, rj, vj = (np.array(x, dtype=L) for x in (ri, vi, rj, vj))
    p0 = (ri - vi * L(dt)) - (rj - vj * L(dt))
    w = vi - vj
    ww = np.dot(w, w)
    t = L(0) if ww == 0 else -np.dot(p0, w) / ww
    t = min(max(t, L(0)), L(dt))
    d = p0 + w * t
    return float(np.dot(d, d))


def test_interpolation_matches_closed_form():
    rng = np.random.default_rng(7)
    for _ in range(2000):
        ri = rng.normal(0, 5, 3) + 7000.0
        rj = ri + rng.normal(0, 2, 3)
        vi, vj = rng.normal(0, 7, 3), rng.normal(0, 7, 3)
        dt = float(rng.uniform(0.5, 60))
        dr, p2 = ol.distance_lines(ri, vi, rj, vj, dt)
        ref = _closest_approach_closed_form(ri, vi, rj, vj, dt)
        got = float(np.dot(dr, dr))
        assert abs(got - ref) <= 1e-6 * max(1.0, ref)
        # p2 lies on j's back-extrapolated line
        s = (p2 - (rj - vj * dt))
        assert np.allclose(np.cross(s, vj), 0, atol=1e-6 * np.linalg.norm(vj) * (1 + np.linalg.norm(s)))


def test_planted_mistake_is_caught():
    """A dropped term in the nominator (vj_rj) must fail the closed-form check."""
    ri, vi = np.array([7000.0, 1.0, 2.0]), np.array([0.0, 7.5, 0.2])
    rj, vj = np.array([7000.5, 0.0, 2.5]), np.array([0.1, -7.4, 0.3])
    dt = 10.0
    oi, oj = ri - vi * dt, rj - vj * dt
    t_bad = (vi @ oj + vj @ oi - vi @ oi) / (vi @ vi + vj @ vj - 2 * vi @ vj)
    t_bad = min(max(t_bad, 0.0), dt)
    d_bad = (oi + vi * t_bad) - (oj + vj * t_bad)
    ref = _closest_approach_closed_form(ri, vi, rj, vj, dt)
    assert abs(d_bad @ d_bad - ref) > 1e-3
    dr, _ = ol.distance_lines(ri, vi, rj, vj, dt)
    assert abs(dr @ dr - ref) < 1e-9


def _random_cloud(n, box, seed, vel=0.0):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, box, (n, 3))
    v = rng.normal(0, vel, (n, 3)) if vel else np.zeros((n, 3))
    return r, v, rng


@pytest.mark.parametrize("n,box,cutoff,seed", [(15, 5.0, 1.0, 42), (400, 20.0, 1.5, 1), (3000, 60.0, 2.0, 3)])
def test_grid_equals_bruteforce(n, box, cutoff, seed):
    """CollisionFunctorIntegrationTest: any traversal gives the brute-force pair set."""
    r, v, rng = _random_cloud(n, box, seed, vel=0.5)
    radius = rng.uniform(100, 2000, n)  # large radii so the threshold keeps many pairs
    act = rng.integers(0, 3, n).astype(np.int32)
    parent = np.where(rng.random(n) < 0.2, rng.integers(0, 5, n), 2**64 - 1).astype(np.uint64)
    objs = ol.ObjectSet(r, v, radius, act, np.arange(n), parent)
    cfg = ol.functor_cfg(cutoff, 10.0, 1.0, 500.0, 0.5)
    hb, eb = ol.pairs(objs, cfg, method="brute")
    hg, eg = ol.pairs(objs, cfg, method="grid", nthreads=4)
    assert hb.shape[0] > 0
    assert hb.tobytes() == hg.tobytes()
    assert eb.tobytes() == eg.tobytes()


@pytest.mark.parametrize("n,box,cutoff,seed", [(400, 20.0, 1.5, 1), (3000, 60.0, 2.0, 3)])
def test_grid_digest_equals_digest_of_bruteforce(n, box, cutoff, seed):
    """The digest-only grid scan (used where pair sets are too large to store) reports the
    counts and digests of the brute-force pair lists; coordinates straddle zero so the
    packed cell keys cover negative cells too."""
    r, v, rng = _random_cloud(n, box, seed, vel=0.5)
    r -= box / 2
    radius = rng.uniform(100, 2000, n)
    act = rng.integers(0, 3, n).astype(np.int32)
    parent = np.where(rng.random(n) < 0.2, rng.integers(0, 5, n), 2**64 - 1).astype(np.uint64)
    objs = ol.ObjectSet(r, v, radius, act, rng.permutation(n), parent)
    cfg = ol.functor_cfg(cutoff, 10.0, 1.0, 500.0, 0.5)
    hb, eb = ol.pairs(objs, cfg, method="brute")
    nh, dh, ne, de = ol.grid_digest(objs, cfg, nthreads=4)
    assert hb.shape[0] > 0 and eb.shape[0] > 0
    assert (nh, dh, ne, de) == (hb.shape[0], ol.digest(hb), eb.shape[0], ol.digest(eb))


def test_grid_digest_of_replica_groups():
    """Constellation-like replica groups (every member within 50 m of its slot): all pairs
    of a group are conjunctions, so the count is the closed form sum over groups of
    m (m - 1) / 2, and a planted change of one id changes the digest."""
    rng = np.random.default_rng(11)
    groups, m = 40, 60
    centres = rng.uniform(-7000, 7000, (groups, 3))
    r = np.repeat(centres, m, axis=0) + rng.uniform(-0.05, 0.05, (groups * m, 3))
    v = np.repeat(rng.normal(0, 7.0, (groups, 3)), m, axis=0)
    n = r.shape[0]
    ids = np.arange(n)
    objs = ol.ObjectSet(r, v, np.full(n, 0.1), None, ids)
    cfg = ol.functor_cfg(5.0, 10.0, threshold_mode=1, threshold_km=0.5)
    nh, dh, ne, _ = ol.grid_digest(objs, cfg)
    assert nh == groups * m * (m - 1) // 2 and ne == 0
    hits, _ = ol.pairs(objs, cfg)
    assert dh == ol.digest(hits)
    ids2 = ids.copy()
    ids2[5] = n + 7
    assert ol.grid_digest(ol.ObjectSet(r, v, np.full(n, 0.1), None, ids2), cfg)[1] != dh


def test_cutoff_only_pairs_equal_distance_definition():
    """Integration-test fixture semantics: radius-scaled threshold large, pairs = |r_i - r_j| <= cutoff."""
    r, v, _ = _random_cloud(300, 10.0, 5)
    n = r.shape[0]
    objs = ol.ObjectSet(r, v + 1e-3, np.full(n, 1e7), None, np.arange(n))
    hits, _ = ol.pairs(objs, ol.functor_cfg(1.0, 1.0), method="grid")
    d2 = ((r[:, None, :] - r[None, :, :]) ** 2).sum(-1)
    iu = np.triu_indices(n, 1)
    exp = {(int(i), int(j)) for i, j in zip(*iu) if d2[i, j] <= 1.0}
    assert {(int(h["idx_a"]), int(h["idx_b"])) for h in hits} == exp


def test_absolute_threshold_mode():
    ri = [[7000.0, 0, 0], [7000.6, 0, 0], [7002.0, 0, 0]]
    v = [[0, 7.5, 0], [0, 7.5, 1e-3], [0, 7.5, -1e-3]]
    objs = ol.ObjectSet(ri, v, [0.01, 0.01, 0.01])
    hits, _ = ol.pairs(objs, ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0), method="brute")
    assert [(int(h["idx_a"]), int(h["idx_b"])) for h in hits] == [(0, 1)]
    hits, _ = ol.pairs(objs, ol.functor_cfg(10.0, 10.0, threshold_mode=0), method="brute")
    assert hits.shape[0] == 0


def test_digest_is_order_independent():
    rng = np.random.default_rng(0)
    p = np.zeros(100, ol.PAIR_DTYPE)
    p["id_a"] = rng.integers(0, 1000, 100)
    p["id_b"] = p["id_a"] + rng.integers(1, 1000, 100)
    d1 = ol.digest(p)
    d2 = ol.digest(p[rng.permutation(100)])
    assert d1 == d2
    q = p.copy()
    q["id_b"][3] += 1
    assert ol.digest(q) != d1

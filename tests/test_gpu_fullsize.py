"""Full-size parity against the oracle on the two large single-population workloads.

Clustered (8,388,608-object constellation workload, a 262,144-object window of it that
holds whole replica groups): each group of ~1,078 copies of one constellation slot lies
within +-50 m, so every pair of a group is a conjunction (~7e7 per step in the window) and
the pair set is checked as count + 128-bit digest (the oracle's digest mode stores no
pairs).  Invariant carried over from the reference: any traversal gives the brute-force
pair set (tests/testladds/CollisionFunctorIntegrationTest.cpp:51-67,125); the oracle's
grid scan is pinned to its brute force on the CPU.

Debris tail (1,048,576 objects, full force model): per-iteration pair sets of
ladds_simulate, bit-identical to the oracle's grid scan of the same state, and FP64
positions within 1e-9 relative of the oracle's step from the previous state.
"""
import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu

SLOTS = 72 * 22 + 36 * 49 + 40 * 33  # constellation slots of the clustered workload (4,668)
C2_N = 8388608
C2_SEED = 0x1ADD5 + 2
WINDOW = 262144
MIN_ALT = 150.0


def _rel_err(a, b):
    return float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))


def _pair_equal(a, b):
    """Bit equality of sorted pair arrays; NaN d^2 / midpoints (parallel tracks, the
    reference's NaN path) compare equal to NaN whatever their payload."""
    assert a.shape == b.shape, (a.shape, b.shape)
    for f in ("id_a", "id_b", "idx_a", "idx_b"):
        assert np.array_equal(a[f], b[f]), f
    for f in ("d2", "mid"):
        x, y = a[f], b[f]
        nan = np.isnan(x)
        assert np.array_equal(nan, np.isnan(y)), f
        assert np.array_equal(x[~nan], y[~nan]), f


def _take(pop, idx):
    out = la.Population(*[getattr(pop, k)[idx].copy() for k in ("rx", "ry", "rz", "vx", "vy", "vz", "radius_m",
                                                                "mass_kg", "aom", "bc_inv", "cr")])
    out.activity = pop.activity[idx].copy()
    out.id = pop.id[idx].copy()  # ids of the full population: the window is a true subset
    out.parent = pop.parent[idx].copy()
    return out


@pytest.fixture(scope="module")
def c2_window():
    """Every replica of 120 constellation slots spread over the three shells, plus the first
    background objects, 262,144 objects in all (indices and ids of the full population)."""
    full = la.generate_population(2, C2_SEED, C2_N)
    nclustered = int(0.6 * C2_N)
    slots = np.arange(0, SLOTS, SLOTS // 120)[:120]
    i = np.arange(nclustered)
    reps = i[np.isin(i % SLOTS, slots)]
    back = np.arange(nclustered, nclustered + WINDOW - reps.size)
    idx = np.concatenate([reps, back])
    assert idx.size == WINDOW
    return _take(full, idx), slots.size


def _oracle_digest(r, v, pop, cutoff, dt, threshold):
    objs = ol.ObjectSet(r, v, pop.radius_m, pop.activity, pop.id, pop.parent)
    return ol.grid_digest(objs, ol.functor_cfg(cutoff, dt, threshold_mode=1, threshold_km=threshold))


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT], ids=["fp64", "split"])
def test_clustered_window_digest_matches_oracle(c2_window, precision):
    pop, groups = c2_window
    cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=5.0,
                          threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=0.5, pair_output=la.PAIRS_DIGEST,
                          log_capacity=1 << 16)
    with la.Engine(cfg) as eng:
        eng.set_objects(pop)
        got = eng.conjunction_digest(10.0)
        want = _oracle_digest(pop.r(), pop.v(), pop, 5.0, 10.0, 0.5)
        print(f"step 0: {got[0]} conjunctions (oracle {want[0]})", flush=True)
        # the replica groups alone give sum m (m - 1) / 2 conjunctions
        assert got[0] > groups * 1000 * 999 // 2
        assert got == want
        r_prev, v_prev = pop.r(), pop.v()
        for it in (1, 2):
            st = eng.simulate(1, first_iteration=it, tpcd=1, min_altitude_km=MIN_ALT)
            assert st["remaining"] == pop.n
            r, v = eng.state()
            if precision == la.FP64:
                ro, vo = ol.propagate(r_prev, v_prev, (it - 1) * 10.0, 10.0, la.ALL_COMPONENTS, 1, pop.aom,
                                      pop.bc_inv, pop.cr, pop.activity)
                assert _rel_err(r, ro) < 1e-9 and _rel_err(v, vo) < 1e-9, it
            want = _oracle_digest(r, v, pop, 5.0, 10.0, 0.5)
            got = (st["collisions"], st["digest"], st["evasions"], st["edigest"])
            print(f"step {it}: {got[0]} conjunctions (oracle {want[0]})", flush=True)
            assert got == want, it
            r_prev, v_prev = r, v


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT], ids=["fp64", "split"])
def test_debris_tail_1m_pair_sets_match_oracle(precision):
    """1,048,576 objects, full force model, 10 km cutoff, 1 km threshold, three iterations of
    ladds_simulate: each iteration's sorted pair list (ids, indices, d^2, midpoint) is
    bit-identical to the oracle's grid scan of the engine's state."""
    n = 1 << 20
    pop = la.generate_population(1, 0x1ADD5 + 1, n, planted_pairs=512)
    cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
                          threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, log_capacity=1 << 20)
    fcfg = ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0)
    total = 0
    with la.Engine(cfg) as eng:
        eng.set_objects(pop)
        r_prev, v_prev = pop.r(), pop.v()
        for it in range(3):
            st = eng.simulate(1, first_iteration=it, tpcd=1, min_altitude_km=MIN_ALT)
            log = eng.fetch_log()
            eng.clear_log()
            assert st["remaining"] == n and np.all(log["iteration"] == it)
            r, v = eng.state()
            if precision == la.FP64:
                ro, vo = ol.propagate(r_prev, v_prev, it * 10.0, 10.0, la.ALL_COMPONENTS, 1, pop.aom, pop.bc_inv,
                                      pop.cr, pop.activity)
                assert _rel_err(r, ro) < 1e-9 and _rel_err(v, vo) < 1e-9, it
            oh, oe = ol.pairs(ol.ObjectSet(r, v, pop.radius_m, pop.activity, pop.id, pop.parent), fcfg)
            hits = ol.sort_pairs(log[log["evaded"] == 0]["pair"])
            ev = ol.sort_pairs(log[log["evaded"] == 1]["pair"])
            print(f"iteration {it}: {hits.shape[0]} conjunctions (oracle {oh.shape[0]})", flush=True)
            _pair_equal(hits, oh)
            _pair_equal(ev, oe)
            total += oh.shape[0]
            r_prev, v_prev = r, v
    assert total >= 3 * 256

"""Context bookkeeping through the C ABI: the clock across configure calls, the shared
pair log, the gather of object states, and the burn-up compaction of ladds_simulate."""
import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _engine(**kw):
    cfg = dict(dt=10.0, components=la.KEP | la.J2, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
               threshold_km=1.0)
    cfg.update(kw)
    return la.Engine(la.EngineConfig(**cfg))


def test_clock_survives_configure_calls():
    pop = la.generate_population(1, 17, 512)
    with _engine(components=la.ALL_COMPONENTS) as eng:
        eng.set_objects(pop)
        eng.propagate(3)
        assert eng.time == 30.0
        eng.configure(components=la.KEP)  # components only: the clock keeps running
        assert eng.time == 30.0
        eng.configure(dt=5.0)  # a new step keeps the elapsed time
        assert eng.time == 30.0
        eng.propagate(2)
        assert eng.time == 40.0
        eng.set_time(1.0e6)
        eng.configure(components=la.ALL_COMPONENTS)  # set_time then configure: the new clock stays
        assert eng.time == 1.0e6
        eng.propagate(1)
        assert eng.time == 1.0e6 + 5.0
        r, _ = eng.state()
    # the Sun/Moon/Earth-rotation environment of the last step was evaluated at t = 1e6 s
    r0, v0 = pop.r(), pop.v()
    ro, vo = ol.propagate(r0, v0, 0.0, 10.0, 0xFF, 3, pop.aom, pop.bc_inv, pop.cr, pop.activity)
    ro, vo = ol.propagate(ro, vo, 30.0, 5.0, la.KEP, 2)
    ro, vo = ol.propagate(ro, vo, 1.0e6, 5.0, 0xFF, 1, pop.aom, pop.bc_inv, pop.cr, pop.activity)
    assert np.max(np.linalg.norm(r - ro, axis=1) / np.linalg.norm(ro, axis=1)) < 1e-9


def test_autotune_refuses_an_unfetched_log():
    pop = la.generate_population(0, 0x1ADD5, 8192, planted_pairs=64)
    with _engine() as eng:
        eng.set_objects(pop)
        st = eng.simulate(2, tpcd=1)
        assert st["collisions"] > 0
        with pytest.raises(la.LaddsError):
            eng.autotune(repetitions=1)
        with pytest.raises(la.LaddsError):
            eng.conjunctions()
        log = eng.fetch_log()
        assert log.shape[0] == st["collisions"] + st["evasions"]
        eng.clear_log()
        eng.autotune(repetitions=1)
        hits, _ = eng.conjunctions()
        assert hits.shape[0] > 0


def test_gather_state_equals_full_state():
    pop = la.generate_population(1, 23, 20000)
    with _engine(components=la.ALL_COMPONENTS) as eng:
        eng.set_objects(pop)
        eng.propagate(4)
        idx = np.array([0, 19999, 7, 7, 12345, 500], np.uint32)
        r, v = eng.gather_state(idx)
        rf, vf = eng.state()
        assert np.array_equal(r, rf[idx]) and np.array_equal(v, vf[idx])
        with pytest.raises(la.LaddsError):
            eng.gather_state([20000])


@pytest.mark.parametrize("graph", [0, 1])
def test_simulate_compacts_only_after_burnups(graph):
    """Objects pushed below 150 km burn up inside ladds_simulate and are compacted at its
    end, order kept; a call without burn-ups leaves the population as it is."""
    pop = la.generate_population(0, 31, 5000)
    r = pop.r()
    low = np.arange(3, 5000, 97)
    for k in low:
        r[k] *= (6378.1363 + 140.0) / np.linalg.norm(r[k])
    pop.rx, pop.ry, pop.rz = r[:, 0].copy(), r[:, 1].copy(), r[:, 2].copy()
    with _engine(step_graph=graph) as eng:
        eng.set_objects(pop)
        st = eng.simulate(3, tpcd=1, min_altitude_km=150.0)
        assert st["remaining"] == 5000 - low.size == eng.n
        assert eng.ids().tolist() == np.setdiff1d(np.arange(5000), low).tolist()
        st = eng.simulate(3, first_iteration=3, tpcd=1, min_altitude_km=150.0)
        assert st["remaining"] == 5000 - low.size == eng.n


def test_simulate_tuning_restores_the_context():
    """ladds_config.tune_simulate: the tuner times every shape over a burst of simulate steps
    run from the current state.  Afterwards the state, the time (later steps see the same
    Sun/Moon positions), the pair log (empty) and the spawn-protection count are those from
    before, so the run continues bit for bit as on a context tuned on detections alone."""
    pop = la.generate_population(1, 0x1ADD5 + 15, 65536, planted_pairs=256)
    pop.parent[:64] = 7  # protected objects: the bursts run the spawn-protection bookkeeping
    runs = []
    for tune_simulate in (0, 1):
        cfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=la.SPLIT, cutoff_km=10.0,
                              threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, tune_simulate=tune_simulate)
        with la.Engine(cfg) as eng:
            eng.set_objects(pop)
            eng.propagate(2)
            r0, v0 = eng.state()
            t = eng.autotune(repetitions=1)
            r1, v1 = eng.state()
            assert np.array_equal(r0, r1) and np.array_equal(v0, v1)
            assert t["tried"] > 10
            assert eng.fetch_log().shape[0] == 0
            hits, _ = eng.conjunctions()
            eng.simulate(5, first_iteration=2, tpcd=1)
            log = eng.fetch_log()
            r2, v2 = eng.state()
        runs.append((sorted(zip(hits["id_a"].tolist(), hits["id_b"].tolist())),
                     sorted((int(e["iteration"]), int(e["pair"]["id_a"]), int(e["pair"]["id_b"])) for e in log), r2, v2))
    assert len(runs[0][0]) > 0 and len(runs[0][1]) > 0
    assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][2], runs[1][2]) and np.array_equal(runs[0][3], runs[1][3])


def test_set_tuning_applies_a_saved_shape():
    """ladds_set_tuning: the shape one context's tuner chose, changed in workgroup size, tile,
    dispatch order, bucket shape and pair-kernel mode and applied to a second context without
    timing, gives the same pair set (and the oracle's); values outside the tuner's range are
    refused."""
    pop = la.generate_population(1, 0x1ADD5 + 14, 65536, planted_pairs=256)
    with _engine(precision=la.SPLIT) as eng:
        eng.set_objects(pop)
        t = eng.autotune(repetitions=1)
        hits_tuned, _ = eng.conjunctions()
    t2 = dict(t, pair_threads=512 if t["pair_threads"] != 512 else 384, tile_size=4096,
              plane_order=1 - t["plane_order"], row_cells=1 if t["row_cells"] == 2 else 2,
              bucket_cells=1 if t["bucket_cells"] == 2 else 2, persistent=1 - t["persistent"])
    with _engine(precision=la.SPLIT) as eng:
        eng.set_objects(pop)
        eng.set_tuning(t2)
        hits, _ = eng.conjunctions()
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, pair_threads=300))
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, cell_size_km=1.0))
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, plane_order=2))
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, row_cells=3))
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, bucket_cells=-1))
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, persistent=2))
        with pytest.raises(la.LaddsError):
            eng.set_tuning(dict(t2, pair_streams=3))
    oh, _ = ol.pairs(ol.ObjectSet(pop.r(), pop.v(), pop.radius_m, pop.activity, pop.id, pop.parent),
                     ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0))
    key = lambda p: sorted(zip(p["id_a"].tolist(), p["id_b"].tolist()))
    assert oh.shape[0] > 100 and key(hits) == key(hits_tuned) == key(oh)

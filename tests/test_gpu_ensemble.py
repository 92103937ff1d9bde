"""Batched ensembles: independent scenarios in one context (ladds_set_scenarios).

Every scenario's pair set must equal the oracle's for that scenario alone, no pair may
join two scenarios (scenario 1 is an exact copy of scenario 0, so any leak between
scenarios would produce distance-zero pairs), and compaction must keep the scenarios.
"""
import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu

SIZES = [700, 700, 2300, 1024, 5000, 64, 3100, 9000]


def _ensemble(planted=24):
    pop, off = la.generate_ensemble(SIZES, 0x1ADD5 + 100, workload=1, planted_pairs=planted)
    # scenario 1 := exact copy of scenario 0 (same positions, its own ids)
    a0, a1, b1 = int(off[0]), int(off[1]), int(off[2])
    for f in ("rx", "ry", "rz", "vx", "vy", "vz", "radius_m", "mass_kg", "aom", "bc_inv", "cr"):
        getattr(pop, f)[a1:b1] = getattr(pop, f)[a0:a1]
    return pop, off


def _scenario_oracle(pop, off, s, r, v, cfg):
    lo, hi = int(off[s]), int(off[s + 1])
    objs = ol.ObjectSet(r[lo:hi], v[lo:hi], pop.radius_m[lo:hi], pop.activity[lo:hi], pop.id[lo:hi],
                        pop.parent[lo:hi])
    return ol.pairs(objs, cfg, method="brute")


def _key(p, shift=0):
    return sorted((int(x["id_a"]), int(x["id_b"]), int(x["idx_a"]) + shift, int(x["idx_b"]) + shift,
                   float(x["d2"]).hex(), tuple(float(m).hex() for m in x["mid"])) for x in p)


@pytest.mark.parametrize("precision,chunks", [(la.FP64, 0), (la.SPLIT, 0), (la.SPLIT, 3)])
def test_pairs_per_scenario(precision, chunks):
    pop, off = _ensemble()
    cfg = dict(dt=10.0, components=la.KEP | la.J2, precision=precision, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, scan_chunks=chunks)
    with la.Engine(la.EngineConfig(**cfg)) as eng:
        eng.set_objects(pop)
        eng.set_scenarios(off)
        eng.propagate(2)
        hits, ev = eng.conjunctions()
        r, v = eng.state()
    assert all((int(h["id_a"]) >> 32) == (int(h["id_b"]) >> 32) for h in hits)
    ocfg = ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0)
    total = 0
    for s in range(len(SIZES)):
        oh, oe = _scenario_oracle(pop, off, s, r, v, ocfg)
        mine = [h for h in hits if (int(h["id_a"]) >> 32) == s]
        mine_e = [h for h in ev if (int(h["id_a"]) >> 32) == s]
        assert _key(mine) == _key(oh, int(off[s])), s
        assert _key(mine_e) == _key(oe, int(off[s])), s
        total += len(mine)
    assert total >= 8 * 24 // 2


def test_simulate_logs_scenarios_and_compaction_keeps_them():
    pop, off = _ensemble()
    # put the first 5 objects of scenario 3 at 120 km altitude: they burn up
    lo = int(off[3])
    for i in range(lo, lo + 5):
        sc = (6378.1363 + 120.0) / np.linalg.norm([pop.rx[i], pop.ry[i], pop.rz[i]])
        pop.rx[i] *= sc
        pop.ry[i] *= sc
        pop.rz[i] *= sc
    cfg = dict(dt=10.0, components=la.KEP | la.J2, precision=la.SPLIT, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0)
    with la.Engine(la.EngineConfig(**cfg)) as eng:
        eng.set_objects(pop)
        eng.set_scenarios(off)
        st = eng.simulate(12, tpcd=1, min_altitude_km=150.0)
        log = eng.fetch_log()
        eng.clear_log()
        new_off = eng.scenarios()
        ids = eng.ids()
        hits, _ = eng.conjunctions()
    assert st["collisions"] > 0
    assert np.all(log["scenario"] == (log["pair"]["id_a"] >> np.uint64(32)))
    assert np.all(log["scenario"] == (log["pair"]["id_b"] >> np.uint64(32)))
    expect = np.diff(off.astype(np.int64))
    expect[3] -= 5
    assert np.array_equal(np.diff(new_off.astype(np.int64)), expect)
    for s in range(len(SIZES)):
        assert np.all((ids[int(new_off[s]):int(new_off[s + 1])] >> np.uint64(32)) == s)
    assert all((int(h["id_a"]) >> 32) == (int(h["id_b"]) >> 32) for h in hits)


@pytest.mark.parametrize("tune_simulate", [0, 1])
def test_autotune_with_scenarios_keeps_pairs(tune_simulate):
    """Autotuning an ensemble, on single detections or on simulate bursts (which run over
    every scenario and are undone): the pair set, the state and the scenario offsets after it
    are those before."""
    pop, off = _ensemble()
    cfg = dict(dt=10.0, components=la.KEP, precision=la.SPLIT, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, tune_simulate=tune_simulate)
    with la.Engine(la.EngineConfig(**cfg)) as eng:
        eng.set_objects(pop)
        eng.set_scenarios(off)
        before = eng.conjunctions()[0]
        r0, v0 = eng.state()
        eng.autotune(repetitions=1)
        after = eng.conjunctions()[0]
        r1, v1 = eng.state()
        assert np.array_equal(eng.scenarios(), off)
    assert _key(ol.sort_pairs(before)) == _key(ol.sort_pairs(after))
    assert np.array_equal(r0, r1) and np.array_equal(v0, v1)


def test_bad_offsets_are_rejected():
    pop, off = _ensemble(planted=0)
    with la.Engine(la.EngineConfig(dt=10.0, cutoff_km=10.0)) as eng:
        eng.set_objects(pop)
        bad = off.copy()
        bad[-1] -= 1
        with pytest.raises(la.LaddsError):
            eng.set_scenarios(bad)
        bad = off.copy()
        bad[2], bad[3] = bad[3], bad[2]
        with pytest.raises(la.LaddsError):
            eng.set_scenarios(bad)


# a plane-tile shape with the persistent pair kernel (workgroups taking plane pairs from
# per-XCD queues), valid for the split form's 10 km cutoff
PERSISTENT_SHAPE = {"cell_size_km": 12.625125012625125, "plane_width_km": 25.25025002525025, "plane_bits": 9,
                    "pair_threads": 128, "tile_size": 4096, "chunks": 1, "workgroups_per_cu": 0, "plane_order": 1,
                    "ensemble_plane_objects": 128, "row_cells": 2, "bucket_cells": 2, "persistent": 1}


def _run(pop, off, graph, steps, rounds, shape=None):
    cfg = dict(dt=10.0, components=la.ALL_COMPONENTS, precision=la.SPLIT, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, step_graph=graph)
    out = []
    with la.Engine(la.EngineConfig(**cfg)) as eng:
        eng.set_objects(pop)
        if off is not None:
            eng.set_scenarios(off)
        if shape is not None:
            eng.set_tuning(shape)
        for k in range(rounds):
            st = eng.simulate(steps, first_iteration=k * steps, tpcd=1, min_altitude_km=150.0)
            log = eng.fetch_log()
            eng.clear_log()
            # eager work between the replays
            hits, _ = eng.conjunctions()
            r, v = eng.state()
            keys = np.sort(np.array([(int(e["iteration"]), int(e["pair"]["id_a"]), int(e["pair"]["id_b"]),
                                      int(e["scenario"])) for e in log], dtype=np.uint64).reshape(-1, 4), axis=0)
            out.append((st["collisions"], st["evasions"], st["remaining"], keys, r, v, len(hits)))
    return out


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("ensemble", [False, True])
def test_graph_replay_matches_eager_launches(ensemble, persistent):
    """Each simulate() call captures one step graph and replays it per step; three calls in a
    row, with eager conjunction scans and synchronisations between them, must reproduce the
    eagerly launched run exactly (pair logs with iterations, states, counters), with one pair
    workgroup per plane pair and with the persistent pair kernel (its queues are cleared by
    the partition inside the graph, never by a memset)."""
    if ensemble:
        pop, off = _ensemble()
    else:
        pop, off = la.generate_population(1, 0x1ADD5 + 21, 20000, planted_pairs=64), None
    shape = PERSISTENT_SHAPE if persistent else None
    eager = _run(pop, off, 0, 30, 3, shape)
    graph = _run(pop, off, 1, 30, 3, shape)
    for a, b in zip(eager, graph):
        assert a[:3] == b[:3]
        assert np.array_equal(a[3], b[3])
        assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])
        assert a[6] == b[6]
    assert sum(x[0] for x in eager) > 0


@pytest.mark.parametrize("plane_objects", [32, 256, 512])
def test_ensemble_plane_sizes_keep_pairs(plane_objects):
    """Ensemble planes sized for more (or fewer) objects fold more (or fewer) distant x-slabs
    into one plane of a scenario, in plane order and in the heaviest-first dispatch order
    over all scenarios' planes, with buckets of one or two cells, one workgroup per plane pair
    or persistent workgroups; each scenario's pair set still equals the oracle's."""
    pop, off = _ensemble()
    cfg = dict(dt=10.0, components=la.KEP | la.J2, precision=la.SPLIT, cutoff_km=10.0,
               threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0)
    with la.Engine(la.EngineConfig(**cfg)) as eng:
        eng.set_objects(pop)
        eng.set_scenarios(off)
        t = eng.autotune(repetitions=1)
        eng.set_tuning(dict(t, ensemble_plane_objects=plane_objects, plane_order=int(plane_objects >= 256),
                            row_cells=2 if plane_objects != 256 else 1, bucket_cells=2 if plane_objects == 32 else 1,
                            persistent=int(plane_objects != 256)))
        eng.propagate(2)
        hits, ev = eng.conjunctions()
        r, v = eng.state()
    ocfg = ol.functor_cfg(10.0, 10.0, threshold_mode=1, threshold_km=1.0)
    total = 0
    for s in range(len(SIZES)):
        oh, oe = _scenario_oracle(pop, off, s, r, v, ocfg)
        mine = [h for h in hits if (int(h["id_a"]) >> 32) == s]
        mine_e = [h for h in ev if (int(h["id_a"]) >> 32) == s]
        assert _key(mine) == _key(oh, int(off[s])), s
        assert _key(mine_e) == _key(oe, int(off[s])), s
        total += len(mine)
    assert total >= 8 * 24 // 2

"""Breakup hook on the device: remove a colliding pair, append its fragments, keep going.

Mirrors particle/BreakupWrapper.cpp:13-45 (deleteParticle of both objects, addParticle of
each fragment with the first object's id as parent).  After the edit the state and the
pair sets of the following steps agree with the oracle run on the same objects; pairs of
fragments with one parent are skipped (CollisionFunctor.cpp:65-69).
"""
import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu

UMAX = np.iinfo(np.uint64).max
MASK, DT, CUTOFF = la.KEP | la.J2, 10.0, 10.0


def _engine(variant):
    return la.Engine(la.EngineConfig(scan_variant=variant, dt=DT, components=MASK, cutoff_km=CUTOFF,
                                     threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0))


def _group(centre, v, parent, first_id, count, rng, spread=0.4):
    r = centre + rng.uniform(-spread, spread, (count, 3))
    vv = v + rng.normal(0.0, 0.02, (count, 3))
    return dict(rx=r[:, 0], ry=r[:, 1], rz=r[:, 2], vx=vv[:, 0], vy=vv[:, 1], vz=vv[:, 2],
                radius_m=rng.uniform(0.05, 1.0, count), activity=np.zeros(count, np.int32),
                id=np.arange(first_id, first_id + count, dtype=np.uint64),
                parent=np.full(count, parent, np.uint64))


class Mirror:
    """Host copy of the per-object constants, edited like the device population."""

    def __init__(self, pop):
        self.radius, self.act, self.parent = pop.radius_m.copy(), pop.activity.copy(), pop.parent.copy()

    def delete(self, keep):
        self.radius, self.act, self.parent = self.radius[keep], self.act[keep], self.parent[keep]

    def add(self, g):
        self.radius = np.concatenate([self.radius, g["radius_m"]])
        self.act = np.concatenate([self.act, g["activity"]])
        self.parent = np.concatenate([self.parent, g["parent"]])

    def pairs(self, r, v, ids):
        objs = ol.ObjectSet(r, v, self.radius, self.act, ids, self.parent)
        return ol.pairs(objs, ol.functor_cfg(CUTOFF, DT, threshold_mode=1, threshold_km=1.0), method="grid")


@pytest.mark.parametrize("variant", [1, 2], ids=["atomic_sort", "plane_tiles"])
def test_breakup_delete_add_then_continue(variant):
    rng = np.random.default_rng(7)
    pop = la.generate_population(0, 0x1ADD5, 4096, planted_pairs=32)
    host = Mirror(pop)
    with _engine(variant) as eng:
        eng.set_objects(pop)
        eng.propagate(2)
        hits, _ = eng.conjunctions()
        assert hits.shape[0] >= 2
        r, v = eng.state()
        ids = eng.ids()
        # two collisions -> two breakups: remove both objects, add 40 fragments each
        groups, doomed = [], []
        next_id = int(ids.max()) + 1
        for h in ol.sort_pairs(hits)[:2]:
            a, b = int(h["idx_a"]), int(h["idx_b"])
            doomed += [int(ids[a]), int(ids[b])]
            groups.append(_group(np.asarray(h["mid"]), v[a], int(ids[a]), next_id, 40, rng))
            next_id += 40
        # unrelated objects passing through the first cloud: pairs with its fragments count
        first = groups[0]
        centre = np.array([first["rx"][0], first["ry"][0], first["rz"][0]])
        groups.append(_group(centre, np.array([first["vx"][0], first["vy"][0], first["vz"][0]]), UMAX, next_id, 6,
                             rng))
        next_id += 6
        assert eng.delete_objects(doomed) == 4
        keep = ~np.isin(ids, doomed)
        host.delete(keep)
        for g in groups:
            eng.add_objects(**g)          # the first append grows the device arrays
            host.add(g)
        assert eng.n == 4096 - 4 + 86
        r2, v2 = eng.state()
        cat = {k: np.concatenate([g[k] for g in groups]) for k in groups[0]}
        exp_r = np.concatenate([r[keep], np.stack([cat["rx"], cat["ry"], cat["rz"]], 1)])
        exp_v = np.concatenate([v[keep], np.stack([cat["vx"], cat["vy"], cat["vz"]], 1)])
        assert np.array_equal(r2, exp_r) and np.array_equal(v2, exp_v)
        all_ids = eng.ids()
        assert np.array_equal(all_ids, np.concatenate([ids[keep], cat["id"]]))

        hits2, ev2 = eng.conjunctions()
        oh, oe = host.pairs(r2, v2, all_ids)
        got = ol.sort_pairs(hits2)
        for f in ("id_a", "id_b", "idx_a", "idx_b", "d2"):
            assert np.array_equal(got[f], oh[f]), f
        assert np.array_equal(ol.sort_pairs(ev2)["id_a"], oe["id_a"])
        frag = host.parent != UMAX
        pa, pb = host.parent[got["idx_a"]], host.parent[got["idx_b"]]
        assert not np.any((pa == pb) & (pa != UMAX))          # same-breakup pairs skipped
        assert np.any(frag[got["idx_a"]] ^ frag[got["idx_b"]])  # fragment x passer-by found
        assert np.count_nonzero(frag) == 80

        # an append that fits the grown capacity, then the fused loop on the edited set
        far = _group(exp_r[0] + 30.0, exp_v[0], UMAX, next_id, 5, rng)
        eng.add_objects(**far)
        host.add(far)
        all_ids = eng.ids()
        rr, vv = eng.state()
        t0 = eng.time
        eng.clear_log()
        steps = 4
        eng.simulate(steps, first_iteration=0, tpcd=1)
        log = eng.fetch_log()
        for it in range(steps):
            rr, vv = ol.propagate(rr, vv, t0 + it * DT, DT, MASK, 1)
            oh, _ = host.pairs(rr, vv, all_ids)
            mine = log[(log["iteration"] == it) & (log["evaded"] == 0)]["pair"]
            assert sorted(zip(mine["id_a"].tolist(), mine["id_b"].tolist())) == \
                sorted(zip(oh["id_a"].tolist(), oh["id_b"].tolist())), it
        rg, _ = eng.state()
        rel = np.linalg.norm(rg - rr, axis=1) / np.linalg.norm(rr, axis=1)
        assert rel.max() < steps * 1e-9


def test_add_objects_to_an_ensemble_is_refused():
    pop, offsets = la.generate_ensemble([300, 300], 3)
    with _engine(2) as eng:
        eng.set_objects(pop)
        eng.set_scenarios(offsets)
        with pytest.raises(la.LaddsError):
            eng.add_objects(rx=[7000.0], ry=[0.0], rz=[0.0], vx=[0.0], vy=[7.5], vz=[0.0], radius_m=[1.0])
        assert eng.n == 600


def test_delete_unknown_ids_is_a_no_op():
    pop = la.generate_population(0, 5, 1000)
    with _engine(2) as eng:
        eng.set_objects(pop)
        assert eng.delete_objects([10**12, 10**12 + 1]) == 0
        assert eng.n == 1000

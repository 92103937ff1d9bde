"""Spawn protection of breakup fragments inside the fused loop (ladds_simulate).

Simulation::simulationLoop counts iterations since the last one whose detection found
collisions; when the count is 100 at the start of an iteration it resets every parent id
(Simulation.cpp:287-293, 399-402, 438, 470-473), after which fragments of one breakup
conjunct with each other (CollisionFunctor.cpp:65-69 no longer skips them).

The population: a background, a crossing pair that meets at t = 300 s (its collisions
around iteration 29 reset the count) and a cloud of six co-orbital fragments with one parent
id, all within 1 km of each other.  The oracle loop below restates the reference's loop
(count, reset, integrate, detect) and runs on the engine's state, re-read after every call
of ten iterations, so the count also has to carry over between calls.
"""
import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu

UMAX = np.iinfo(np.uint64).max
MASK, DT, CUTOFF = la.KEP | la.J2, 10.0, 10.0
RE = 6378.1363


def _circular(alt_km, inc_deg, raan_deg, u_deg):
    a = RE + alt_km
    vc = np.sqrt(398600.4418 / a)
    i, O, u = np.radians([inc_deg, raan_deg, u_deg])
    r_orb = a * np.array([np.cos(u), np.sin(u), 0.0])
    v_orb = vc * np.array([-np.sin(u), np.cos(u), 0.0])
    rot_i = np.array([[1, 0, 0], [0, np.cos(i), -np.sin(i)], [0, np.sin(i), np.cos(i)]])
    rot_O = np.array([[np.cos(O), -np.sin(O), 0], [np.sin(O), np.cos(O), 0], [0, 0, 1]])
    m = rot_O @ rot_i
    return m @ r_orb, m @ v_orb


def _population():
    bg = la.generate_population(0, 0x5A11, 3000)
    r, v = [bg.r()], [bg.v()]
    # crossing pair: D reaches C's t = 300 s position (0.3 km radially above) on a velocity
    # turned by 20 degrees about the radial direction; D's start is D's state stepped back
    rc, vc = _circular(900.0, 40.0, 10.0, 0.0)
    r300, v300 = ol.propagate(rc[None], vc[None], 0.0, DT, MASK, 30)
    u = r300[0] / np.linalg.norm(r300[0])
    ang = np.radians(20.0)
    vd = v300[0] * np.cos(ang) + np.cross(u, v300[0]) * np.sin(ang) + u * (u @ v300[0]) * (1 - np.cos(ang))
    rd0, vd0 = ol.propagate((r300[0] + 0.3 * u)[None], (-vd)[None], 0.0, DT, MASK, 30)
    r += [rc[None], rd0]
    v += [vc[None], -vd0]
    # fragment cloud: six co-orbital objects 0.15 km apart along-track, one parent id
    rf, vf = _circular(1500.0, 75.0, 200.0, 33.0)
    t_hat = vf / np.linalg.norm(vf)
    r.append(rf + np.outer(0.15 * np.arange(6), t_hat))
    v.append(np.repeat(vf[None], 6, 0))
    r, v = np.concatenate(r), np.concatenate(v)
    n = r.shape[0]
    parent = np.full(n, UMAX, np.uint64)
    parent[-6:] = 777
    return dict(rx=r[:, 0].copy(), ry=r[:, 1].copy(), rz=r[:, 2].copy(), vx=v[:, 0].copy(), vy=v[:, 1].copy(),
                vz=v[:, 2].copy(), radius_m=np.full(n, 1.0), id=np.arange(n, dtype=np.uint64), parent=parent)


def _oracle_loop(r, v, parent, radius, ids, first, count, tpcd, quiet):
    """The reference loop's spawn protection around oracle steps and scans."""
    cfg = ol.functor_cfg(CUTOFF, DT * tpcd, threshold_mode=1, threshold_km=1.0)
    out = {}
    for it in range(first, first + count):
        if quiet == 100:
            parent = np.full_like(parent, UMAX)
        r, v = ol.propagate(r, v, it * DT, DT, MASK, 1)
        if it % tpcd == 0:
            hits, _ = ol.pairs(ol.ObjectSet(r, v, radius, None, ids, parent), cfg)
            out[it] = sorted(zip(hits["id_a"].tolist(), hits["id_b"].tolist()))
            if hits.shape[0]:
                quiet = 0
        quiet += 1
    return out, parent, quiet


@pytest.mark.parametrize("tpcd,graph", [(1, 0), (3, 0), (1, 1)], ids=["every", "tpcd3", "graph"])
def test_spawn_protection_lifts_after_100_quiet_iterations(tpcd, graph):
    objs = _population()
    n = objs["rx"].shape[0]
    frag = set(range(n - 6, n))
    parent = objs["parent"].copy()
    quiet = 0
    lifted_at = None
    last_hit = None  # the latest iteration whose detection found collisions before the lift
    with la.Engine(la.EngineConfig(dt=DT, components=MASK, cutoff_km=CUTOFF, threshold_mode=la.THRESHOLD_ABSOLUTE,
                                   threshold_km=1.0, step_graph=graph)) as eng:
        eng.set_objects(**objs)
        for call in range(14):
            first = 10 * call
            r, v = eng.state()
            eng.simulate(10, first_iteration=first, tpcd=tpcd)
            log = eng.fetch_log()
            want, parent, quiet = _oracle_loop(r, v, parent, objs["radius_m"], objs["id"], first, 10, tpcd, quiet)
            for it, pairs in want.items():
                mine = log[(log["iteration"] == it) & (log["evaded"] == 0)]["pair"]
                got = sorted(zip(mine["id_a"].tolist(), mine["id_b"].tolist()))
                assert got == pairs, it
                if lifted_at is None and any(a in frag and b in frag for a, b in pairs):
                    lifted_at = it
                if lifted_at is None and pairs:
                    last_hit = it
    # The reset comes 100 iterations after the last collision (or after iteration 0 when there
    # is none); the fragments' 15 pairs appear at the first detection from there on, not before.
    # Detecting every iteration sees the crossing pair around iteration 29; 30-s windows
    # interpolate its arcs as chords and miss it.
    reset = 100 if last_hit is None else last_hit + 100
    assert lifted_at is not None and reset <= lifted_at < reset + tpcd
    if tpcd == 1:
        assert last_hit is not None and 25 <= last_hit <= 32


def test_remove_spawn_protection_entry_point():
    """ladds_remove_spawn_protection for loops driven through ladds_conjunctions."""
    objs = _population()
    with la.Engine(la.EngineConfig(dt=DT, components=MASK, cutoff_km=CUTOFF, threshold_mode=la.THRESHOLD_ABSOLUTE,
                                   threshold_km=1.0)) as eng:
        eng.set_objects(**objs)
        before, _ = eng.conjunctions()
        eng.remove_spawn_protection()
        after, _ = eng.conjunctions()
    r = np.stack([objs["rx"], objs["ry"], objs["rz"]], 1)
    v = np.stack([objs["vx"], objs["vy"], objs["vz"]], 1)
    cfg = ol.functor_cfg(CUTOFF, DT, threshold_mode=1, threshold_km=1.0)
    oh, _ = ol.pairs(ol.ObjectSet(r, v, objs["radius_m"], None, objs["id"], np.full(r.shape[0], UMAX, np.uint64)),
                     cfg)
    assert after.shape[0] == before.shape[0] + 15
    assert sorted(zip(ol.sort_pairs(after)["id_a"].tolist(), ol.sort_pairs(after)["id_b"].tolist())) == \
        sorted(zip(oh["id_a"].tolist(), oh["id_b"].tolist()))

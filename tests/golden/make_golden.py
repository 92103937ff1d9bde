"""Writes the golden fixtures of tests/golden/ from the LADDS reference tree.

Run in the development container (where the reference tree is mounted):
    python tests/golden/make_golden.py /root/reference

Fixtures (data only: inputs and expected outputs):
  collision_functor_cases.json  known-answer vectors of tests/testladds/CollisionFunctorTest.cpp
                                (ThreeParticles :18-59, MixActivityStates :76-209,
                                CollisionDistanceFactorTest :215-258, LinearInterpolation :322-354)
  kepler_1s.json                the 1-s Kepler pair of tests/testladds/SimulationTest.cpp:39-57
  pos_test.csv, v_test.csv,     10 LEO objects and their state 10 s later (data/ of the reference;
  pos_test_10s.csv, v_test_10s.csv  no reference test reads them; Kepler-only motion reproduces them)
"""
from __future__ import annotations

import json
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
UMAX = 2**64 - 1


def particle(r, v, pid, activity, mass, radius, parent=UMAX):
    return {"r": r, "v": v, "id": pid, "activity": activity, "mass": mass, "radius": radius, "parent": parent}


def cases():
    out = {}
    # ThreeParticles: radius 0.6 m, 1 m apart on x, v = (0, i, 0); functor(cutoff 1, dt 10, cdf 1, 0.01, 0.1)
    out["three_particles"] = {
        "source": "tests/testladds/CollisionFunctorTest.cpp:18-59",
        "functor": {"cutoff": 1.0, "dt": 10.0, "cdf": 1.0, "min_det": 0.01, "evasion": 0.1},
        "particles": [particle([i / 1000.0, 0.0, 0.0], [0.0, float(i), 0.0], i, 0, 1.0, 0.6) for i in range(3)],
        "expected_collisions": [
            {"a": 0, "b": 1, "d2": 1.0 / (1000 * 1000), "mid": [0.0005, 0.0, 0.0]},
            {"a": 1, "b": 2, "d2": 1.0 / (1000 * 1000), "mid": [0.0015, 0.0, 0.0]},
        ],
    }
    # MixActivityStates: two particles per activity state, 0.1 km apart, zero velocity (NaN path)
    P, E, EP = 0, 1, 2
    mix = [
        particle([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 0, P, 1.0, 0.001),
        particle([0.0, 0.0, 0.1], [0.0, 0.0, 0.0], 0, P, 1.0, 0.001),
        particle([0.1, 0.0, 0.0], [0.0, 0.0, 0.0], 1, P, 1.0, 1.0),
        particle([0.1, 0.0, 0.1], [0.0, 0.0, 0.0], 1, P, 1.0, 1.0),
        particle([0.0, 0.1, 0.0], [0.0, 0.0, 0.0], 2, E, 1.0, 1.0),
        particle([0.0, 0.1, 0.1], [0.0, 0.0, 0.0], 3, E, 1.0, 1.0),
        particle([0.1, 0.1, 0.0], [0.0, 0.0, 0.0], 4, EP, 1.0, 1.0),
        particle([0.1, 0.1, 0.1], [0.0, 0.0, 0.0], 5, EP, 1.0, 1.0),
    ]
    expected_idx = [(0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (0, 6), (0, 7),
                    (1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (1, 7), (2, 3)]
    out["mix_activity_states"] = {
        "source": "tests/testladds/CollisionFunctorTest.cpp:76-209",
        "functor": {"cutoff": 1.0, "dt": 10.0, "cdf": 1.0, "min_det": 0.01, "evasion": 0.1},
        "particles": mix,
        "expected_id_pairs": [[mix[a]["id"], mix[b]["id"]] for a, b in expected_idx],
    }
    # CollisionDistanceFactorTest: radius 1 m, 3 m apart, head-on at 1 km/s each, dt 1
    out["collision_distance_factor"] = {
        "source": "tests/testladds/CollisionFunctorTest.cpp:215-258",
        "functor": {"cutoff": 80.0, "dt": 1.0, "min_det": 0.1, "evasion": 0.1},
        "particles": [particle([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], 1, P, 1.0, 1.0),
                      particle([3.0 / 1000, 0.0, 0.0], [-1.0, 0.0, 0.0], 2, P, 1.0, 1.0)],
        "expected_count_by_cdf": {"1.0": 0, "2.0": 1},
    }
    # LinearInterpolationTest: radius 1000 m, cdf 10, cutoff 80
    vecs = [
        ([0., 1., 2.], [2., 4., 6.], [0., 0., 0.], [1., 2., 3.], 1., 3., [0.5, 1.5, 2.5]),
        ([5., 0., 0.], [1., 1., 1.], [-1., -1., -1.], [1., 1., 1.], 1., 18., [3., 0.5, 0.5]),
        ([1.5, 0., 1.], [3., 0., 2.], [-3., 0., -2.], [3., 0., 2.], 2., 0., [2.25, 0., 1.5]),
        ([0., 0., 0.], [0., 0.5, 0.], [1., 0., 0.], [0., 0., 1.], 42., .25, [0., 0.25, 0.]),
    ]
    out["linear_interpolation"] = {
        "source": "tests/testladds/CollisionFunctorTest.cpp:260-354",
        "functor": {"cutoff": 80.0, "cdf": 10.0, "min_det": 0.1, "evasion": 0.1},
        "radius": 1000.0,
        "vectors": [{"x1": a, "x2": b, "v1": c, "v2": d, "dt": e, "d2": f, "mid": g} for a, b, c, d, e, f, g in vecs],
    }
    return out


def main(ref: Path):
    (HERE / "collision_functor_cases.json").write_text(json.dumps(cases(), indent=1))
    kepler = {
        "source": "tests/testladds/SimulationTest.cpp:39-57 (p2 is p1 after 1 s of Kepler motion)",
        "dt": 1.0,
        "p1": {"r": [6871., 0., 0.], "v": [0, 4.8958309146899, 5.83462408131549]},
        "p2": {"r": [6870.99577848984, 4.89582991243564, 5.83462288687537],
               "v": [-0.00844301944979238, 4.89582790671436, 5.83462049654983]},
    }
    (HERE / "kepler_1s.json").write_text(json.dumps(kepler, indent=1))
    for name in ("pos_test.csv", "v_test.csv", "pos_test_10s.csv", "v_test_10s.csv"):
        shutil.copyfile(ref / "data" / name, HERE / name)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))

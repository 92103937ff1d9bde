"""Output records packed on the device (the payloads of LADDS's HDF5 writer) and the
checkpoint reload built on them.

particle_records: io/hdf5/HDF5Writer.cpp:89-149 (float32 Vec3 state, uint32 ids);
log_records: HDF5Writer.cpp:151-187 ({uint32, uint32, float32} per pair and iteration);
population_from_records: io/hdf5/HDF5Reader.cpp:25-65.  Expected values are numpy's casts
of the FP64 state and of the device log (round to nearest, low 32 bits of ids).
"""
import numpy as np
import pytest

import ladds_amd as la

pytestmark = pytest.mark.gpu


def _cast_log(log, it, evaded):
    sel = log[(log["iteration"] == it) & ((log["evaded"] != 0) == evaded)]["pair"]
    out = np.zeros(sel.shape[0], la.RECORD_DTYPE)
    out["id_a"] = (sel["id_a"] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    out["id_b"] = (sel["id_b"] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    out["d2"] = sel["d2"].astype(np.float32)
    return np.sort(out, order=["id_a", "id_b", "d2"])


def test_particle_and_log_records_match_casts():
    pop = la.generate_population(0, 0x1ADD5, 6000, planted_pairs=40)
    pop.id = pop.id + np.uint64(3 << 32)          # ids past 2^32: records keep the low 32 bits
    pop.activity = np.where(np.arange(pop.n) % 5 == 0, la.EVASIVE, la.PASSIVE).astype(np.int32)
    cfg = la.EngineConfig(dt=10.0, components=la.KEP | la.J2, cutoff_km=10.0,
                          threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, min_detection_radius_m=0.01,
                          evasion_cutoff_km=1.0)
    with la.Engine(cfg) as eng:
        eng.set_objects(pop)
        eng.simulate(5, first_iteration=10, tpcd=1)
        log = eng.fetch_log()
        r, v = eng.state()
        ids = eng.ids()
        pos, vel, rid = eng.particle_records()
        assert np.array_equal(pos, r.astype(np.float32)) and np.array_equal(vel, v.astype(np.float32))
        assert np.array_equal(rid, (ids & np.uint64(0xFFFFFFFF)).astype(np.uint32))
        seen_ev = 0
        for it in range(10, 15):
            h, e = eng.log_records(it)
            assert np.array_equal(np.sort(h, order=["id_a", "id_b", "d2"]), _cast_log(log, it, False)), it
            assert np.array_equal(np.sort(e, order=["id_a", "id_b", "d2"]), _cast_log(log, it, True)), it
            seen_ev += e.shape[0]
        assert log.shape[0] > 100 and seen_ev > 0
        h, e = eng.log_records(9)
        assert h.shape[0] == 0 and e.shape[0] == 0


def test_log_records_of_a_large_log():
    """1.1M pairs in one iteration: more log entries than one launch has threads."""
    rng = np.random.default_rng(5)
    n = 1500
    r = np.array([7000.0, 0.0, 0.0]) + rng.uniform(-0.5, 0.5, (n, 3))
    v = np.array([0.0, 7.5, 0.0]) + rng.normal(0, 1e-3, (n, 3))
    cfg = la.EngineConfig(dt=1.0, components=la.KEP, cutoff_km=2.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                          threshold_km=2.0, log_capacity=1 << 21)
    with la.Engine(cfg) as eng:
        eng.set_objects(rx=r[:, 0], ry=r[:, 1], rz=r[:, 2], vx=v[:, 0], vy=v[:, 1], vz=v[:, 2],
                        radius_m=np.ones(n))
        eng.simulate(1, first_iteration=0, tpcd=1)
        log = eng.fetch_log()
        assert log.shape[0] == n * (n - 1) // 2
        h, e = eng.log_records(0)
    assert np.array_equal(np.sort(h, order=["id_a", "id_b", "d2"]), _cast_log(log, 0, False))
    assert e.shape[0] == 0


def test_checkpoint_round_trip_through_records():
    """Write records, rebuild the population as HDF5Reader::readParticles does, continue."""
    pop = la.generate_population(0, 77, 3000)
    pop.activity = (np.arange(pop.n) % 3).astype(np.int32)
    table = {"id": pop.id.astype(np.uint32), "mass": pop.mass_kg.astype(np.float32),
             "radius": pop.radius_m.astype(np.float32), "bc_inv": pop.bc_inv.astype(np.float32),
             "activity": pop.activity}
    cfg = la.EngineConfig(dt=10.0, components=la.KEP | la.J2, cutoff_km=10.0)
    with la.Engine(cfg) as eng:
        eng.set_objects(pop)
        eng.propagate(4)
        pos, vel, ids = eng.particle_records()
    back = la.population_from_records(pos[::-1], vel[::-1], ids[::-1], table)
    assert np.array_equal(back.r(), pos[::-1].astype(np.float64))
    assert np.array_equal(back.activity, pop.activity[::-1])
    assert np.array_equal(back.radius_m, pop.radius_m.astype(np.float32).astype(np.float64)[::-1])
    with la.Engine(cfg) as eng:
        eng.set_objects(back)
        assert np.array_equal(eng.ids(), ids[::-1].astype(np.uint64))
        p2, v2, _ = eng.particle_records()
        assert np.array_equal(p2, pos[::-1]) and np.array_equal(v2, vel[::-1])


def test_log_records_overflow_reports_counts():
    pop = la.generate_population(0, 0x1ADD5, 4000, planted_pairs=40)
    cfg = la.EngineConfig(dt=10.0, components=la.KEP, cutoff_km=10.0, threshold_mode=la.THRESHOLD_ABSOLUTE,
                          threshold_km=1.0)
    with la.Engine(cfg) as eng:
        eng.set_objects(pop)
        eng.simulate(1, first_iteration=0, tpcd=1)
        full, _ = eng.log_records(0)
        assert full.shape[0] > 4
        import ctypes as C
        small = np.zeros(2, la.RECORD_DTYPE)
        n, ne = C.c_size_t(0), C.c_size_t(0)
        rc = la.lib.ladds_log_records(eng._h, 0, small.ctypes.data, 2, C.byref(n), None, 0, C.byref(ne))
        assert rc == la._abi.E_OVERFLOW and n.value == full.shape[0]

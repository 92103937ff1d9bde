"""ladds_amd -- MI355X (gfx950) propagate + conjunction engine for LADDS.

Python mirror of the reference's hot-path call surface, over the C ABI of
``include/ladds_amd.h`` (``libladds_amd.so``):

=========================================  ================================================
reference (LADDS / OrbitPropagator)         here
=========================================  ================================================
``Simulation::initAutoPas`` + container     :class:`Engine` (owns device state)
``AutoPas::addParticle`` (bulk)             :meth:`Engine.set_objects`
``YoshidaIntegrator::integrate(false)``     :meth:`Engine.propagate`
``Simulation::collisionDetection``          :meth:`Engine.conjunctions`
``Simulation::simulationLoop`` hot path     :meth:`Engine.simulate`
``Simulation::deleteBurnUps``               :meth:`Engine.delete_burnups`
=========================================  ================================================

Errors are raised as :class:`LaddsError` (the reference throws ``std::runtime_error``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import (ALL_COMPONENTS, C22, DRAG, EVASIVE, EVASIVE_PRESERVING, FP64, J2, KEP, LUN, PAIRS_DIGEST,
                   PAIRS_LIST, PASSIVE, S22, SOL, SPLIT, SPLIT_ALL32, SRP, THRESHOLD_ABSOLUTE,
                   THRESHOLD_RADIUS_SCALED, LOGGED_DTYPE, PAIR_DTYPE, RECORD_DTYPE)

__all__ = [
    "Engine", "EngineConfig", "LaddsError", "generate_population", "generate_ensemble", "read_dataset",
    "write_conjunctions_csv", "bc_inv_from", "device_count",
    "components_mask", "KEP", "J2", "C22", "S22", "SOL", "LUN", "SRP", "DRAG", "ALL_COMPONENTS", "FP64", "SPLIT",
    "SPLIT_ALL32",
    "THRESHOLD_RADIUS_SCALED", "THRESHOLD_ABSOLUTE", "PASSIVE", "EVASIVE", "EVASIVE_PRESERVING",
    "PAIRS_LIST", "PAIRS_DIGEST", "PAIR_DTYPE", "LOGGED_DTYPE", "RECORD_DTYPE", "population_from_records",
]

lib = _abi.lib


class LaddsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ladds error {code}: {message}")
        self.code = code


def components_mask(kep=False, j2=False, c22=False, s22=False, sol=False, lun=False, srp=False, drag=False) -> int:
    """Bit mask from the ``sim/prop/use*Component`` flags (Simulation.cpp:127-135)."""
    flags = (kep, j2, c22, s22, sol, lun, srp, drag)
    return sum(1 << k for k, f in enumerate(flags) if f)


def device_count() -> int:
    n = C.c_int(0)
    lib.ladds_device_count(C.byref(n))
    return n.value


@dataclass
class EngineConfig:
    """Mirror of ``ladds_config``; field names follow the reference's YAML keys."""
    dt: float = 10.0                      # sim/deltaT
    t0: float = 0.0                       # accumulator start time (Simulation.cpp:143)
    components: int = KEP                 # sim/prop/use*Component
    precision: int = FP64
    cutoff_km: float = 80.0               # autopas/cutoff
    collision_distance_factor: float = 1.0  # sim/collisionDistanceFactor
    min_detection_radius_m: float = 0.05  # sim/minDetectionRadius
    evasion_cutoff_km: float = 0.1        # sim/evasionTrackingCutoffInKM
    threshold_mode: int = THRESHOLD_RADIUS_SCALED
    threshold_km: float = 1.0
    cell_size_km: float = 0.0
    scan_variant: int = 0
    block_size: int = 0
    device: int = 0
    pair_output: int = PAIRS_LIST
    log_capacity: int = 1 << 20
    step_graph: int = 0                   # replay each simulate() step from a captured HIP graph
    scan_chunks: int = 0                  # workgroups per plane of the plane-tile scan (0: auto)
    tune_simulate: int = 0                # autotune() times each shape over simulate() steps

    def to_c(self) -> _abi.Config:
        c = _abi.Config()
        lib.ladds_config_default(C.byref(c))
        for name, _ in _abi.Config._fields_:
            setattr(c, name, getattr(self, name))
        return c


def _ptr(a, ctype=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def _f64(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n,):
        raise ValueError(f"expected shape ({n},), got {a.shape}")
    return a


@dataclass
class Population:
    rx: np.ndarray
    ry: np.ndarray
    rz: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    radius_m: np.ndarray
    mass_kg: np.ndarray
    aom: np.ndarray
    bc_inv: np.ndarray
    cr: np.ndarray
    activity: np.ndarray = field(default=None)
    id: np.ndarray = field(default=None)
    parent: np.ndarray = field(default=None)

    @property
    def n(self) -> int:
        return self.rx.shape[0]

    def r(self) -> np.ndarray:
        return np.stack([self.rx, self.ry, self.rz], 1)

    def v(self) -> np.ndarray:
        return np.stack([self.vx, self.vy, self.vz], 1)

    def copy(self) -> "Population":
        return Population(**{k: (None if v is None else v.copy()) for k, v in self.__dict__.items()})


def generate_population(workload: int, seed: int, n: int, planted_pairs: int = 0) -> Population:
    """Seeded synthetic population (``ladds_generate_population``); host only."""
    arrs = [np.empty(n, np.float64) for _ in range(11)]
    rc = lib.ladds_generate_population(workload, seed, n, planted_pairs, *[_ptr(a) for a in arrs])
    if rc != 0:
        raise LaddsError(rc, "ladds_generate_population")
    pop = Population(*arrs)
    pop.activity = np.zeros(n, np.int32)
    pop.id = np.arange(n, dtype=np.uint64)
    pop.parent = np.full(n, np.iinfo(np.uint64).max, np.uint64)
    return pop


_ACTIVITY = {"passive": PASSIVE, "evasive": EVASIVE, "evasivePreserving": EVASIVE_PRESERVING}
_R_EARTH_KM = 6378.1363


def bc_inv_from(bstar, mass, radius, coefficient_of_drag):
    """``Particle::calculateBcInv`` (src/ladds/particle/Particle.cpp:13-27): c_D A / m when B* is
    NaN or 0, else 2 R_E / 0.1570 B* with R_E in km (the unit src/ladds/Simulation.cpp:573 adds it
    to a km altitude in)."""
    bstar, mass, radius = (np.asarray(x, np.float64) for x in (bstar, mass, radius))
    by_cd = coefficient_of_drag * (np.pi * radius * radius) / mass
    return np.where(np.isnan(bstar) | (bstar == 0.0), by_cd, (2.0 * _R_EARTH_KM / 0.1570) * bstar)


def read_dataset(path, coefficient_of_drag: float = 2.2):
    """``DatasetReader::readDataset`` (src/ladds/io/DatasetReader.cpp:15-50): CSV with a header
    and the columns id, COSPAR id, name, B*, mass [kg], radius [m], activity, r [km], v [km/s].
    Returns the Population and the COSPAR identifiers."""
    import csv

    rows = []
    with open(path, newline="") as f:
        rd = csv.reader(f)
        next(rd)
        for lineno, row in enumerate(rd, start=2):
            if not row:
                continue
            if len(row) != 13:
                raise ValueError(f"{path}:{lineno}: expected 13 columns, got {len(row)}")
            rows.append(row)
    n = len(rows)
    col = lambda k: np.array([float(r[k]) for r in rows], np.float64)
    mass, radius, bstar = col(4), col(5), col(3)
    pop = Population(rx=col(7), ry=col(8), rz=col(9), vx=col(10), vy=col(11), vz=col(12), radius_m=radius,
                     mass_kg=mass, aom=np.pi * radius * radius * 1e-6 / mass,
                     bc_inv=bc_inv_from(bstar, mass, radius, coefficient_of_drag), cr=np.ones(n))
    pop.activity = np.array([_ACTIVITY.get(r[6], PASSIVE) for r in rows], np.int32)
    pop.id = np.array([int(r[0]) for r in rows], np.uint64)
    pop.parent = np.full(n, np.iinfo(np.uint64).max, np.uint64)
    return pop, [r[1] for r in rows]


def population_from_records(pos, vel, ids, constant_properties) -> Population:
    """Checkpoint reload (``HDF5Reader::readParticles``, io/hdf5/HDF5Reader.cpp:25-65): float32
    state and uint32 ids of one output iteration plus the ConstantProperties table (a mapping
    or structured array with id, mass, radius, bc_inv, activity), widened to float64.
    Parents are reset to "none", as the reference does."""
    table = {k: np.asarray(constant_properties[k]) for k in ("id", "mass", "radius", "bc_inv", "activity")}
    where = {int(i): k for k, i in enumerate(table["id"])}
    rows = np.array([where[int(i)] for i in np.asarray(ids)], np.int64)
    pos = np.asarray(pos, np.float32).astype(np.float64)
    vel = np.asarray(vel, np.float32).astype(np.float64)
    mass = table["mass"][rows].astype(np.float32).astype(np.float64)
    radius = table["radius"][rows].astype(np.float32).astype(np.float64)
    n = pos.shape[0]
    pop = Population(rx=pos[:, 0].copy(), ry=pos[:, 1].copy(), rz=pos[:, 2].copy(), vx=vel[:, 0].copy(),
                     vy=vel[:, 1].copy(), vz=vel[:, 2].copy(), radius_m=radius, mass_kg=mass,
                     aom=np.pi * radius * radius * 1e-6 / mass,
                     bc_inv=table["bc_inv"][rows].astype(np.float32).astype(np.float64), cr=np.ones(n))
    pop.activity = table["activity"][rows].astype(np.int32)
    pop.id = np.asarray(ids).astype(np.uint64)
    pop.parent = np.full(n, np.iinfo(np.uint64).max, np.uint64)
    return pop


def write_conjunctions_csv(path, log):
    """``ConjunctionLogger`` format (src/ladds/io/ConjuctionLogger.cpp:13-44): header
    ``Iteration,P1,P2,SquaredDistance``, one line per logged pair (smaller id first), the
    squared distance in its shortest round-trip form."""
    with open(path, "w") as f:
        f.write("Iteration,P1,P2,SquaredDistance\n")
        for e in log:
            f.write(f"{int(e['iteration'])},{int(e['pair']['id_a'])},{int(e['pair']['id_b'])},"
                    f"{float(e['pair']['d2'])!r}\n")


def generate_ensemble(sizes, seed: int, workload: int = 1, planted_pairs: int = 0):
    """Concatenated seeded scenarios (scenario s: ``generate_population(workload, seed + s, sizes[s],
    planted_pairs)``: sizes[s] objects, planted partners included) and their offsets.  Ids are
    ``(s << 32) | local index``, so pairs name their scenario."""
    parts = [generate_population(workload, seed + s, int(k), planted_pairs) for s, k in enumerate(sizes)]
    sizes = [p.n for p in parts]
    cat = {f: np.concatenate([getattr(p, f) for p in parts]) for f in Population.__dataclass_fields__}
    cat["id"] = np.concatenate([(np.uint64(s) << np.uint64(32)) | np.arange(k, dtype=np.uint64)
                                for s, k in enumerate(sizes)])
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return Population(**cat), offsets


class Engine:
    """One device context: the AutoPas container + Yoshida integrator + CollisionFunctor."""

    def __init__(self, config: EngineConfig | None = None, **kw):
        self.config = config or EngineConfig(**kw)
        self._c = self.config.to_c()
        h = C.c_void_p()
        rc = lib.ladds_create(C.byref(self._c), C.byref(h))
        if rc != 0:
            raise LaddsError(rc, "ladds_create failed (no usable gfx950 device?)")
        self._h = h

    # -- lifetime -------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib.ladds_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = lib.ladds_last_error(self._h)
            raise LaddsError(rc, f"{what}: {msg.decode() if msg else ''}")

    def configure(self, **changes):
        """Update dt, t0, components and the CollisionFunctor parameters (ladds_configure)."""
        for k, v in changes.items():
            if not hasattr(self.config, k):
                raise AttributeError(k)
            setattr(self.config, k, v)
        self._c = self.config.to_c()
        self._check(lib.ladds_configure(self._h, C.byref(self._c)), "configure")
        if "dt" in changes and "t0" not in changes:
            # a new step alone keeps the elapsed time: the context's clock restarts at the current time
            self.config.t0 = self.time

    # -- state ----------------------------------------------------------
    @staticmethod
    def _pack(pop, arrays):
        src = dict(pop.__dict__) if pop is not None else {}
        src.update(arrays)
        n = len(src["rx"])
        keep = {}
        o = _abi.Objects()
        o.n = n
        for name in ("rx", "ry", "rz", "vx", "vy", "vz", "radius_m"):
            keep[name] = _f64(src[name], n)
            setattr(o, name, _ptr(keep[name]))
        for name in ("aom", "bc_inv", "cr"):
            if src.get(name) is not None:
                keep[name] = _f64(src[name], n)
                setattr(o, name, _ptr(keep[name]))
        if src.get("activity") is not None:
            keep["activity"] = np.ascontiguousarray(src["activity"], dtype=np.int32)
            o.activity = _ptr(keep["activity"], C.c_int32)
        for name in ("id", "parent"):
            if src.get(name) is not None:
                keep[name] = np.ascontiguousarray(src[name], dtype=np.uint64)
                setattr(o, name, _ptr(keep[name], C.c_uint64))
        return o, keep

    def set_objects(self, pop: Population | None = None, **arrays):
        """``AutoPas::addParticle`` for a whole population (replaces all objects)."""
        o, keep = self._pack(pop, arrays)
        self._check(lib.ladds_set_objects(self._h, C.byref(o)), "set_objects")

    def add_objects(self, pop: Population | None = None, **arrays):
        """Append objects and keep the current state: the breakup hook
        (particle/BreakupWrapper.cpp:38-43 adds each fragment to the container).
        Fragments of one breakup share ``parent``; pairs with equal parent ids
        are skipped (CollisionFunctor.cpp:65-69)."""
        o, keep = self._pack(pop, arrays)
        self._check(lib.ladds_add_objects(self._h, C.byref(o)), "add_objects")

    @property
    def n(self) -> int:
        return int(lib.ladds_num_objects(self._h))

    @property
    def time(self) -> float:
        return float(lib.ladds_get_time(self._h))

    def set_time(self, t: float):
        self._check(lib.ladds_set_time(self._h, t), "set_time")
        self.config.t0 = float(t)  # later configure() calls keep this clock

    def state(self):
        """(r, v) as (n, 3) arrays."""
        n = self.n
        a = [np.empty(n, np.float64) for _ in range(6)]
        self._check(lib.ladds_get_state(self._h, *[_ptr(x) for x in a]), "get_state")
        return np.stack(a[:3], 1), np.stack(a[3:], 1)

    def ids(self) -> np.ndarray:
        out = np.empty(self.n, np.uint64)
        self._check(lib.ladds_get_ids(self._h, _ptr(out, C.c_uint64)), "get_ids")
        return out

    # -- hot path -------------------------------------------------------
    def propagate(self, nsteps: int = 1):
        """``YoshidaIntegrator::integrate(false)`` nsteps times."""
        self._check(lib.ladds_propagate(self._h, int(nsteps)), "propagate")

    def conjunctions(self, dt_interp: float | None = None):
        """``Simulation::collisionDetection``: (collisions, evasions) as PAIR_DTYPE arrays."""
        dt = self.config.dt if dt_interp is None else dt_interp
        cap = 1024
        while True:
            hits = np.zeros(cap, PAIR_DTYPE)
            ev = np.zeros(cap, PAIR_DTYPE)
            n, ne = C.c_size_t(0), C.c_size_t(0)
            rc = lib.ladds_conjunctions(self._h, dt, hits.ctypes.data, cap, C.byref(n), ev.ctypes.data, cap,
                                        C.byref(ne))
            if rc == _abi.E_OVERFLOW:
                cap = max(n.value, ne.value) + 1
                continue
            self._check(rc, "conjunctions")
            return hits[: n.value], ev[: ne.value]

    def conjunction_digest(self, dt_interp: float | None = None):
        dt = self.config.dt if dt_interp is None else dt_interp
        n, ne = C.c_uint64(0), C.c_uint64(0)
        d = (C.c_uint64 * 2)()
        e = (C.c_uint64 * 2)()
        self._check(lib.ladds_conjunction_digest(self._h, dt, C.byref(n), d, C.byref(ne), e), "conjunction_digest")
        return n.value, (d[0], d[1]), ne.value, (e[0], e[1])

    def simulate(self, nsteps: int, first_iteration: int = 0, tpcd: int = 1, min_altitude_km: float = 0.0):
        """``Simulation::simulationLoop`` hot path; returns the run statistics as a dict."""
        st = _abi.RunStats()
        self._check(lib.ladds_simulate(self._h, int(first_iteration), int(nsteps), int(tpcd),
                                       float(min_altitude_km), C.byref(st)), "simulate")
        return {"collisions": st.collisions, "evasions": st.evasions, "logged": st.logged,
                "candidates": st.candidates, "digest": (st.digest[0], st.digest[1]),
                "edigest": (st.edigest[0], st.edigest[1]), "remaining": st.remaining}

    def gather_state(self, idx):
        """(r, v) of the objects at the given device indices, as (k, 3) arrays."""
        idx = np.ascontiguousarray(idx, np.uint32)
        r = np.empty((idx.size, 3))
        v = np.empty((idx.size, 3))
        self._check(lib.ladds_gather_state(self._h, _ptr(idx, C.c_uint32), idx.size, _ptr(r), _ptr(v)),
                    "gather_state")
        return r, v

    def remove_spawn_protection(self):
        """``Simulation::removeParticleSpawnProtection``: every parent id reset to none."""
        self._check(lib.ladds_remove_spawn_protection(self._h), "remove_spawn_protection")

    def fetch_log(self) -> np.ndarray:
        n = C.c_size_t(0)
        rc = lib.ladds_fetch_log(self._h, None, 0, C.byref(n))
        if rc not in (0, _abi.E_OVERFLOW):
            self._check(rc, "fetch_log")
        out = np.zeros(max(n.value, 1), LOGGED_DTYPE)
        self._check(lib.ladds_fetch_log(self._h, out.ctypes.data, out.shape[0], C.byref(n)), "fetch_log")
        return out[: n.value]

    def clear_log(self):
        self._check(lib.ladds_clear_log(self._h), "clear_log")

    def particle_records(self):
        """Payload of ``HDF5Writer::writeParticles`` (io/hdf5/HDF5Writer.cpp:89-149), packed on
        the device: float32 (n,3) positions and velocities and uint32 ids."""
        n = self.n
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        ids = np.zeros(n, np.uint32)
        self._check(lib.ladds_particle_records(self._h, pos.ctypes.data, vel.ctypes.data, ids.ctypes.data),
                    "particle_records")
        return pos, vel, ids

    def log_records(self, iteration: int):
        """Payload of ``HDF5Writer::writeConjunctions`` / ``writeEvasions`` (HDF5Writer.cpp:151-187)
        for one iteration of the device log: two RECORD_DTYPE arrays, unordered."""
        n, ne = C.c_size_t(0), C.c_size_t(0)
        rc = lib.ladds_log_records(self._h, int(iteration), None, 0, C.byref(n), None, 0, C.byref(ne))
        if rc not in (0, _abi.E_OVERFLOW):
            self._check(rc, "log_records")
        hits = np.zeros(max(n.value, 1), RECORD_DTYPE)
        ev = np.zeros(max(ne.value, 1), RECORD_DTYPE)
        self._check(lib.ladds_log_records(self._h, int(iteration), hits.ctypes.data, hits.shape[0], C.byref(n),
                                          ev.ctypes.data, ev.shape[0], C.byref(ne)), "log_records")
        return hits[: n.value], ev[: ne.value]

    def delete_burnups(self, min_altitude_km: float) -> int:
        removed = C.c_size_t(0)
        self._check(lib.ladds_delete_burnups(self._h, float(min_altitude_km), C.byref(removed)), "delete_burnups")
        return removed.value

    def delete_objects(self, ids) -> int:
        """Remove the objects with these ids, order of the rest kept (``autopas.deleteParticle``
        of the colliding pair, BreakupWrapper.cpp:23-25)."""
        a = np.ascontiguousarray(np.atleast_1d(ids), dtype=np.uint64)
        removed = C.c_size_t(0)
        self._check(lib.ladds_delete_objects(self._h, _ptr(a, C.c_uint64), a.size, C.byref(removed)),
                    "delete_objects")
        return removed.value

    def set_scenarios(self, offsets):
        """Declare the current objects an ensemble of independent scenarios (``ladds_set_scenarios``);
        ``offsets`` has one entry per scenario plus the total."""
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self._check(lib.ladds_set_scenarios(self._h, off.shape[0] - 1, _ptr(off, C.c_uint64)), "set_scenarios")

    def scenarios(self) -> np.ndarray:
        """Current scenario offsets (after burn-up compaction)."""
        off = np.zeros(int(lib.ladds_num_scenarios(self._h)) + 1, np.uint64)
        self._check(lib.ladds_get_scenarios(self._h, _ptr(off, C.c_uint64)), "get_scenarios")
        return off

    def autotune(self, dt_interp: float | None = None, repetitions: int = 3) -> dict:
        """Pick the plane-tile cell size, plane shape and pair workgroup size by timing the
        conjunction scan on the current state (``ladds_autotune``)."""
        t = _abi.Tuning()
        dt = self.config.dt if dt_interp is None else dt_interp
        self._check(lib.ladds_autotune(self._h, float(dt), int(repetitions), C.byref(t)), "autotune")
        return {"cell_size_km": t.cell_size_km, "plane_width_km": t.plane_width_km, "plane_bits": t.plane_bits,
                "pair_threads": t.pair_threads, "tile_size": t.tile_size, "chunks": t.chunks,
                "workgroups_per_cu": t.workgroups_per_cu, "plane_order": t.plane_order,
                "ensemble_plane_objects": t.ensemble_plane_objects, "row_cells": t.row_cells,
                "bucket_cells": t.bucket_cells, "persistent": t.persistent, "pair_streams": t.pair_streams,
                "scan_us": t.scan_us,
                "default_us": t.default_us, "tried": t.tried}

    def set_tuning(self, tuning: dict):
        """Apply a shape returned by ``autotune`` (``ladds_set_tuning``) without timing anything."""
        t = _abi.Tuning()
        for k in ("cell_size_km", "plane_width_km", "plane_bits", "pair_threads", "tile_size", "chunks",
                  "workgroups_per_cu"):
            setattr(t, k, tuning[k])
        t.plane_order = int(tuning.get("plane_order", 0))
        t.ensemble_plane_objects = int(tuning.get("ensemble_plane_objects", 0))
        t.row_cells = int(tuning.get("row_cells", 0))
        t.bucket_cells = int(tuning.get("bucket_cells", 0))
        t.persistent = int(tuning.get("persistent", 0))
        t.pair_streams = int(tuning.get("pair_streams", 0))
        self._check(lib.ladds_set_tuning(self._h, C.byref(t)), "set_tuning")

    # -- measurement ----------------------------------------------------
    def set_profiling(self, enable: bool):
        self._check(lib.ladds_set_profiling(self._h, int(bool(enable))), "set_profiling")

    def kernel_stats(self, name: str):
        launches, ms = C.c_uint64(0), C.c_double(0.0)
        self._check(lib.ladds_kernel_stats(self._h, name.encode(), C.byref(launches), C.byref(ms)), "kernel_stats")
        return launches.value, ms.value

    def synchronize(self):
        self._check(lib.ladds_synchronize(self._h), "synchronize")

    @property
    def stream(self) -> int:
        return int(lib.ladds_stream(self._h) or 0)

"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker the HIP engine is compared against; only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

KEP, J2, C22, S22, SOL, LUN, SRP, DRAG = (1 << k for k in range(8))

PAIR_DTYPE = np.dtype([("id_a", "<u8"), ("id_b", "<u8"), ("idx_a", "<u4"), ("idx_b", "<u4"),
                       ("d2", "<f8"), ("mid", "<f8", (3,))])


class Params(C.Structure):
    _fields_ = [("aom", C.c_double), ("bc_inv", C.c_double), ("cr", C.c_double), ("activity", C.c_int32)]


class FunctorCfg(C.Structure):
    _fields_ = [("cutoff_km", C.c_double), ("dt", C.c_double), ("collision_distance_factor", C.c_double),
                ("min_detection_radius_m", C.c_double), ("evasion_cutoff_km", C.c_double),
                ("threshold_mode", C.c_int), ("threshold_km", C.c_double)]


_dp = C.POINTER(C.c_double)


class Objects(C.Structure):
    _fields_ = [("n", C.c_size_t), ("rx", _dp), ("ry", _dp), ("rz", _dp), ("vx", _dp), ("vy", _dp), ("vz", _dp),
                ("radius_m", _dp), ("activity", C.POINTER(C.c_int32)), ("id", C.POINTER(C.c_uint64)),
                ("parent", C.POINTER(C.c_uint64))]


def _load():
    if not ORACLE_SO.exists():
        raise ImportError(f"{ORACLE_SO} missing: build it with `make -C oracle`")
    lib = C.CDLL(str(ORACLE_SO))
    lib.or_constant.restype = C.c_double
    lib.or_constant.argtypes = [C.c_char_p]
    for f in ("or_sun_position", "or_moon_position"):
        getattr(lib, f).argtypes = [C.c_double, _dp]
    lib.or_moon_series_default.argtypes = [_dp, _dp, _dp]
    lib.or_moon_position_series.argtypes = [C.c_double, _dp, _dp, _dp, _dp]
    lib.or_earth_rotation_angle.restype = C.c_double
    lib.or_earth_rotation_angle.argtypes = [C.c_double]
    lib.or_density.restype = C.c_double
    lib.or_density.argtypes = [C.c_double]
    lib.or_acceleration.argtypes = [_dp, _dp, C.c_double, C.c_uint, C.POINTER(Params), _dp]
    lib.or_yoshida_coefficients.argtypes = [_dp, _dp]
    lib.or_yoshida_steps.argtypes = [C.c_size_t] + [_dp] * 6 + [_dp, _dp, _dp, C.POINTER(C.c_int32), C.c_double,
                                                               C.c_double, C.c_uint, C.c_int, C.c_int]
    lib.or_yoshida_vjp.argtypes = [C.c_size_t] + [_dp] * 9 + [C.POINTER(C.c_int32), C.c_double, C.c_double,
                                                              C.c_uint, C.c_int] + [_dp] * 13 + [C.c_int]
    lib.or_distance_by_linear_interpolation.argtypes = [_dp, _dp, _dp, _dp, C.c_double, _dp, _dp]
    lib.or_collision_functor.restype = C.c_int
    lib.or_collision_functor.argtypes = [C.POINTER(Objects), C.c_size_t, C.c_size_t, C.POINTER(FunctorCfg),
                                         C.c_void_p]
    for f in ("or_pairs_bruteforce",):
        getattr(lib, f).argtypes = [C.POINTER(Objects), C.POINTER(FunctorCfg), C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_size_t), C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    lib.or_pairs_grid.argtypes = [C.POINTER(Objects), C.POINTER(FunctorCfg), C.c_void_p, C.c_size_t,
                                  C.POINTER(C.c_size_t), C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.c_int]
    lib.or_pairs_grid_digest.argtypes = [C.POINTER(Objects), C.POINTER(FunctorCfg), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.c_int]
    lib.or_sort_pairs.argtypes = [C.c_void_p, C.c_size_t]
    lib.or_pairset_digest.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint64)]
    return lib


lib = _load()


def constant(name: str) -> float:
    return lib.or_constant(name.encode())


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def sun(t: float) -> np.ndarray:
    out = np.zeros(3)
    lib.or_sun_position(t, _p(out))
    return out


def moon(t: float) -> np.ndarray:
    out = np.zeros(3)
    lib.or_moon_position(t, _p(out))
    return out


def moon_series_default():
    """Published amplitudes of the lunar series: (longitude [arcsec], latitude [arcsec], distance [km])."""
    lam, beta, r = np.zeros(14), np.zeros(8), np.zeros(9)
    lib.or_moon_series_default(_p(lam), _p(beta), _p(r))
    return lam, beta, r


def moon_with(t: float, lam, beta, r) -> np.ndarray:
    """The oracle's lunar series evaluated with the given amplitudes (planted-mistake tests)."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (lam, beta, r)]
    out = np.zeros(3)
    lib.or_moon_position_series(t, *[_p(a) for a in arrs], _p(out))
    return out


def acceleration(r, v, t, mask, aom=0.0, bc_inv=0.0, cr=1.0, activity=0) -> np.ndarray:
    r = np.ascontiguousarray(r, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    a = np.zeros(3)
    p = Params(aom, bc_inv, cr, activity)
    lib.or_acceleration(_p(r), _p(v), t, mask, C.byref(p), _p(a))
    return a


def yoshida_coefficients():
    c, d = np.zeros(4), np.zeros(3)
    lib.or_yoshida_coefficients(_p(c), _p(d))
    return c, d


def propagate(r, v, t0, dt, mask, nsteps, aom=None, bc_inv=None, cr=None, activity=None, nthreads=0):
    """Yoshida steps of (n,3) r, v; returns new (r, v)."""
    r = np.asarray(r, np.float64)
    v = np.asarray(v, np.float64)
    n = r.shape[0]
    cols = [np.array(r[:, k], dtype=np.float64, copy=True) for k in range(3)] + \
        [np.array(v[:, k], dtype=np.float64, copy=True) for k in range(3)]
    args = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (aom, bc_inv, cr)]
    act = None if activity is None else np.ascontiguousarray(activity, np.int32)
    lib.or_yoshida_steps(n, *[_p(c) for c in cols], *[None if a is None else _p(a) for a in args],
                         None if act is None else _p(act, C.c_int32), t0, dt, mask, nsteps, nthreads)
    return np.stack(cols[:3], 1), np.stack(cols[3:], 1)


def vjp(r, v, t0, dt, mask, nsteps, g_r, g_v, aom=None, bc_inv=None, cr=None, activity=None, nthreads=0):
    """Complex-step vector-Jacobian product of ``propagate``: returns (g_r0, g_v0, g_bc)
    for upstream gradients g_r, g_v of the final (r, v)."""
    r = np.asarray(r, np.float64)
    v = np.asarray(v, np.float64)
    n = r.shape[0]
    cols = [np.ascontiguousarray(x[:, k], np.float64) for x in (r, v) for k in range(3)]
    gcols = [np.ascontiguousarray(x[:, k], np.float64) for x in (np.asarray(g_r), np.asarray(g_v)) for k in range(3)]
    args = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (aom, bc_inv, cr)]
    act = None if activity is None else np.ascontiguousarray(activity, np.int32)
    outs = [np.zeros(n) for _ in range(7)]
    lib.or_yoshida_vjp(n, *[_p(c) for c in cols], *[None if a is None else _p(a) for a in args],
                       None if act is None else _p(act, C.c_int32), t0, dt, mask, nsteps,
                       *[_p(g) for g in gcols], *[_p(o) for o in outs], nthreads)
    return np.stack(outs[:3], 1), np.stack(outs[3:6], 1), outs[6]


def distance_lines(ri, vi, rj, vj, dt):
    arrs = [np.ascontiguousarray(x, np.float64) for x in (ri, vi, rj, vj)]
    dr, p2 = np.zeros(3), np.zeros(3)
    lib.or_distance_by_linear_interpolation(*[_p(a) for a in arrs], dt, _p(dr), _p(p2))
    return dr, p2


class ObjectSet:
    """Keeps numpy arrays alive behind an or_objects view."""

    def __init__(self, r, v, radius, activity=None, ids=None, parent=None):
        r = np.asarray(r, np.float64)
        v = np.asarray(v, np.float64)
        n = r.shape[0]
        self.cols = [np.ascontiguousarray(r[:, k]) for k in range(3)] + [np.ascontiguousarray(v[:, k])
                                                                         for k in range(3)]
        self.radius = np.ascontiguousarray(radius, np.float64)
        self.activity = np.ascontiguousarray(np.zeros(n) if activity is None else activity, np.int32)
        self.ids = np.ascontiguousarray(np.arange(n) if ids is None else ids, np.uint64)
        self.parent = np.ascontiguousarray(np.full(n, np.iinfo(np.uint64).max) if parent is None else parent,
                                           np.uint64)
        self.c = Objects(n, *[_p(c) for c in self.cols], _p(self.radius), _p(self.activity, C.c_int32),
                         _p(self.ids, C.c_uint64), _p(self.parent, C.c_uint64))


def functor_cfg(cutoff, dt, cdf=1.0, min_det=0.05, evasion=0.1, threshold_mode=0, threshold_km=1.0):
    return FunctorCfg(cutoff, dt, cdf, min_det, evasion, threshold_mode, threshold_km)


def pairs(objs: ObjectSet, cfg: FunctorCfg, method="grid", nthreads=0):
    """(collisions, evasions) as PAIR_DTYPE arrays, sorted by (id_a, id_b, idx_a, idx_b)."""
    cap = 1024
    while True:
        hits = np.zeros(cap, PAIR_DTYPE)
        ev = np.zeros(cap, PAIR_DTYPE)
        n, ne = C.c_size_t(0), C.c_size_t(0)
        if method == "brute":
            lib.or_pairs_bruteforce(C.byref(objs.c), C.byref(cfg), hits.ctypes.data, cap, C.byref(n),
                                    ev.ctypes.data, cap, C.byref(ne))
        else:
            lib.or_pairs_grid(C.byref(objs.c), C.byref(cfg), hits.ctypes.data, cap, C.byref(n), ev.ctypes.data,
                              cap, C.byref(ne), nthreads)
        if n.value > cap or ne.value > cap:
            cap = max(n.value, ne.value)
            continue
        hits, ev = hits[: n.value], ev[: ne.value]
        lib.or_sort_pairs(hits.ctypes.data, hits.shape[0])
        lib.or_sort_pairs(ev.ctypes.data, ev.shape[0])
        return hits, ev


def grid_digest(objs: ObjectSet, cfg: FunctorCfg, nthreads=0):
    """(collisions, digest, evasions, evasion digest) of the grid scan without storing pairs;
    digests are (u64, u64) tuples with the definition of ``digest``."""
    n, ne = C.c_uint64(0), C.c_uint64(0)
    d, ed = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
    lib.or_pairs_grid_digest(C.byref(objs.c), C.byref(cfg), C.byref(n), d, C.byref(ne), ed, nthreads)
    return n.value, (d[0], d[1]), ne.value, (ed[0], ed[1])


def sort_pairs(p: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p.copy())
    lib.or_sort_pairs(p.ctypes.data, p.shape[0])
    return p


def digest(p: np.ndarray):
    p = np.ascontiguousarray(p)
    out = (C.c_uint64 * 2)()
    lib.or_pairset_digest(p.ctypes.data, p.shape[0], out)
    return out[0], out[1]

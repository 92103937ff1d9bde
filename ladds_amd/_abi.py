"""ctypes declarations of the C ABI in ``include/ladds_amd.h``.

The library is loaded from the package directory (``ladds_amd/libladds_amd.so``,
built by ``ladds_amd.build``).  There is no fallback: if the library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libladds_amd.so"

# status codes
OK = 0
E_INVALID = -1
E_HIP = -2
E_OVERFLOW = -3
E_NOMEM = -4
E_STATE = -5
E_NODEVICE = -6

KEP, J2, C22, S22, SOL, LUN, SRP, DRAG = (1 << k for k in range(8))
ALL_COMPONENTS = 0xFF

FP64, SPLIT, SPLIT_ALL32 = 0, 1, 2
THRESHOLD_RADIUS_SCALED, THRESHOLD_ABSOLUTE = 0, 1
PASSIVE, EVASIVE, EVASIVE_PRESERVING = 0, 1, 2
PAIRS_LIST, PAIRS_DIGEST = 0, 1


class Config(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("t0", C.c_double),
        ("components", C.c_uint32),
        ("precision", C.c_int32),
        ("cutoff_km", C.c_double),
        ("collision_distance_factor", C.c_double),
        ("min_detection_radius_m", C.c_double),
        ("evasion_cutoff_km", C.c_double),
        ("threshold_mode", C.c_int32),
        ("threshold_km", C.c_double),
        ("cell_size_km", C.c_double),
        ("scan_variant", C.c_int32),
        ("block_size", C.c_int32),
        ("device", C.c_int32),
        ("pair_output", C.c_int32),
        ("log_capacity", C.c_uint64),
        ("step_graph", C.c_int32),
        ("scan_chunks", C.c_int32),
        ("tune_simulate", C.c_int32),
    ]


_dp = C.POINTER(C.c_double)


class Objects(C.Structure):
    _fields_ = [
        ("n", C.c_size_t),
        ("rx", _dp), ("ry", _dp), ("rz", _dp),
        ("vx", _dp), ("vy", _dp), ("vz", _dp),
        ("radius_m", _dp), ("aom", _dp), ("bc_inv", _dp), ("cr", _dp),
        ("activity", C.POINTER(C.c_int32)),
        ("id", C.POINTER(C.c_uint64)),
        ("parent", C.POINTER(C.c_uint64)),
    ]


class Pair(C.Structure):
    _fields_ = [
        ("id_a", C.c_uint64), ("id_b", C.c_uint64),
        ("idx_a", C.c_uint32), ("idx_b", C.c_uint32),
        ("d2", C.c_double),
        ("mid", C.c_double * 3),
    ]


class LoggedPair(C.Structure):
    _fields_ = [
        ("pair", Pair),
        ("iteration", C.c_uint64),
        ("evaded", C.c_uint32),
        ("scenario", C.c_uint32),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("collisions", C.c_uint64),
        ("evasions", C.c_uint64),
        ("logged", C.c_uint64),
        ("candidates", C.c_uint64),
        ("digest", C.c_uint64 * 2),
        ("edigest", C.c_uint64 * 2),
        ("remaining", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("cell_size_km", C.c_double),
        ("plane_bits", C.c_int32),
        ("pair_threads", C.c_int32),
        ("scan_us", C.c_double),
        ("default_us", C.c_double),
        ("tried", C.c_int32),
        ("tile_size", C.c_int32),
        ("chunks", C.c_int32),
        ("workgroups_per_cu", C.c_int32),
        ("plane_width_km", C.c_double),
        ("plane_order", C.c_int32),
        ("ensemble_plane_objects", C.c_int32),
        ("row_cells", C.c_int32),
        ("bucket_cells", C.c_int32),
        ("persistent", C.c_int32),
        ("pair_streams", C.c_int32),
    ]


class AdjointConfig(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("nsteps", C.c_int32),
        ("checkpoint_every", C.c_int32),
        ("dt", C.c_double),
        ("t0", C.c_double),
        ("components", C.c_uint32),
        ("device", C.c_int32),
    ]


assert C.sizeof(Pair) == 56 and C.sizeof(LoggedPair) == 72

# numpy structured dtypes with the same layout (for zero-copy downloads)
try:
    import numpy as np

    PAIR_DTYPE = np.dtype([("id_a", "<u8"), ("id_b", "<u8"), ("idx_a", "<u4"), ("idx_b", "<u4"),
                           ("d2", "<f8"), ("mid", "<f8", (3,))])
    LOGGED_DTYPE = np.dtype([("pair", PAIR_DTYPE), ("iteration", "<u8"), ("evaded", "<u4"), ("scenario", "<u4")])
    RECORD_DTYPE = np.dtype([("id_a", "<u4"), ("id_b", "<u4"), ("d2", "<f4")])
    assert PAIR_DTYPE.itemsize == 56 and LOGGED_DTYPE.itemsize == 72 and RECORD_DTYPE.itemsize == 12
except ImportError:  # pragma: no cover
    np = None

class DeviceObjects(C.Structure):
    """ladds_device_objects: device pointers of caller-owned object arrays."""
    _fields_ = [("n", C.c_size_t), ("rx", C.c_void_p), ("ry", C.c_void_p), ("rz", C.c_void_p),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p), ("radius_m", C.c_void_p),
                ("activity", C.c_void_p), ("id", C.c_void_p), ("parent", C.c_void_p)]


NOALLOC_MAX_OBJECTS = 65536  # LADDS_NOALLOC_MAX_OBJECTS

_ctx = C.c_void_p
_SIGNATURES = {
    "ladds_version": (C.c_char_p, []),
    "ladds_abi_version": (C.c_int, []),
    "ladds_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ladds_config_default": (None, [C.POINTER(Config)]),
    "ladds_create": (C.c_int, [C.POINTER(Config), C.POINTER(_ctx)]),
    "ladds_destroy": (C.c_int, [_ctx]),
    "ladds_last_error": (C.c_char_p, [_ctx]),
    "ladds_configure": (C.c_int, [_ctx, C.POINTER(Config)]),
    "ladds_set_objects": (C.c_int, [_ctx, C.POINTER(Objects)]),
    "ladds_num_objects": (C.c_size_t, [_ctx]),
    "ladds_add_objects": (C.c_int, [_ctx, C.POINTER(Objects)]),
    "ladds_particle_records": (C.c_int, [_ctx, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ladds_log_records": (C.c_int, [_ctx, C.c_uint64, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.c_void_p,
                                    C.c_size_t, C.POINTER(C.c_size_t)]),
    "ladds_delete_objects": (C.c_int, [_ctx, C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(C.c_size_t)]),
    "ladds_set_scenarios": (C.c_int, [_ctx, C.c_size_t, C.POINTER(C.c_uint64)]),
    "ladds_num_scenarios": (C.c_size_t, [_ctx]),
    "ladds_get_scenarios": (C.c_int, [_ctx, C.POINTER(C.c_uint64)]),
    "ladds_get_state": (C.c_int, [_ctx, _dp, _dp, _dp, _dp, _dp, _dp]),
    "ladds_get_ids": (C.c_int, [_ctx, C.POINTER(C.c_uint64)]),
    "ladds_get_time": (C.c_double, [_ctx]),
    "ladds_set_time": (C.c_int, [_ctx, C.c_double]),
    "ladds_propagate": (C.c_int, [_ctx, C.c_int]),
    "ladds_conjunctions": (C.c_int, [_ctx, C.c_double, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t),
                                     C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ladds_conjunction_digest": (C.c_int, [_ctx, C.c_double, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ladds_simulate": (C.c_int, [_ctx, C.c_uint64, C.c_int, C.c_int, C.c_double, C.POINTER(RunStats)]),
    "ladds_remove_spawn_protection": (C.c_int, [_ctx]),
    "ladds_gather_state": (C.c_int, [_ctx, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "ladds_fetch_log": (C.c_int, [_ctx, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ladds_clear_log": (C.c_int, [_ctx]),
    "ladds_delete_burnups": (C.c_int, [_ctx, C.c_double, C.POINTER(C.c_size_t)]),
    "ladds_autotune": (C.c_int, [_ctx, C.c_double, C.c_int, C.POINTER(Tuning)]),
    "ladds_set_tuning": (C.c_int, [_ctx, C.POINTER(Tuning)]),
    "ladds_set_profiling": (C.c_int, [_ctx, C.c_int]),
    "ladds_kernel_stats": (C.c_int, [_ctx, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "ladds_synchronize": (C.c_int, [_ctx]),
    "ladds_stream": (C.c_void_p, [_ctx]),
    "ladds_adjoint_create": (C.c_int, [C.POINTER(AdjointConfig), C.POINTER(_ctx)]),
    "ladds_adjoint_destroy": (C.c_int, [_ctx]),
    "ladds_adjoint_last_error": (C.c_char_p, [_ctx]),
    "ladds_adjoint_checkpoint_doubles": (C.c_uint64, [_ctx]),
    "ladds_adjoint_forward": (C.c_int, [_ctx] + [C.c_void_p] * 10),
    "ladds_adjoint_backward": (C.c_int, [_ctx] + [C.c_void_p] * 11),
    "ladds_generate_population": (C.c_int, [C.c_int, C.c_uint64, C.c_size_t, C.c_size_t] + [_dp] * 11),
    "ladds_conjunction_digest_noalloc": (C.c_int, [C.POINTER(DeviceObjects), C.POINTER(Config), C.c_double,
                                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_void_p]),
}


def load(path: Path | str = LIB_PATH) -> C.CDLL:
    path = Path(path)
    if not path.exists():
        raise ImportError(
            f"{path} is missing: build the gfx950 engine first (python ladds_amd/build.py)")
    lib = C.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()

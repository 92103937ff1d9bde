"""ladds_conjunction_digest_noalloc: the conjunction digest on caller-owned device arrays
with its scratch in a static device array, against ladds_conjunction_digest (the engine's
context-owned path) on the same objects: two calls in a row, the second on fewer objects,
in both precision forms."""
import ctypes as C

import numpy as np
import pytest
import torch

import ladds_amd as la
from ladds_amd import _abi

pytestmark = pytest.mark.gpu

lib = la.lib


def _device_objects(pop, lo, hi):
    dev = torch.device("cuda:0")
    keep = {}

    def put(a, dtype):
        t = torch.from_numpy(np.ascontiguousarray(a[lo:hi])).to(dev, dtype)
        keep[len(keep)] = t
        return t.data_ptr()

    d = _abi.DeviceObjects()
    d.n = hi - lo
    d.rx, d.ry, d.rz = (put(a, torch.float64) for a in (pop.rx, pop.ry, pop.rz))
    d.vx, d.vy, d.vz = (put(a, torch.float64) for a in (pop.vx, pop.vy, pop.vz))
    d.radius_m = put(pop.radius_m, torch.float64)
    d.activity = put(pop.activity, torch.int32)
    d.id = put(pop.id.view(np.int64), torch.int64)
    d.parent = put(pop.parent.view(np.int64), torch.int64)
    return d, keep


def _noalloc(d, cfg, dt):
    n, ne = C.c_uint64(0), C.c_uint64(0)
    dg, edg = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.ladds_conjunction_digest_noalloc(C.byref(d), C.byref(cfg), dt, C.byref(n), dg, C.byref(ne), edg,
                                              C.c_void_p(stream))
    if rc != 0:
        raise la.LaddsError(rc, lib.ladds_last_error(None).decode())
    return n.value, (dg[0], dg[1]), ne.value, (edg[0], edg[1])


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT], ids=["fp64", "split"])
def test_noalloc_digest_equals_context_digest(precision):
    pop = la.generate_population(1, 0x1ADD5 + 40, 60000, planted_pairs=600)
    pop.activity[::11] = la.EVASIVE
    pop.parent[100:140] = 7  # spawn-protected fragments
    ecfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
                           threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0, evasion_cutoff_km=0.9)
    # all 60,000 objects, then the last 20,000 (they hold a third of the planted pairs): the
    # second, smaller call reuses the scratch array the first one filled
    for lo, hi in ((0, 60000), (40000, 60000)):
        sub = la.Population(*[getattr(pop, k)[lo:hi].copy() for k in ("rx", "ry", "rz", "vx", "vy", "vz", "radius_m",
                                                                      "mass_kg", "aom", "bc_inv", "cr")])
        sub.activity, sub.id, sub.parent = pop.activity[lo:hi].copy(), pop.id[lo:hi].copy(), pop.parent[lo:hi].copy()
        with la.Engine(ecfg) as eng:
            eng.set_objects(sub)
            want = eng.conjunction_digest(10.0)
        d, keep = _device_objects(pop, lo, hi)
        got = _noalloc(d, ecfg.to_c(), 10.0)
        torch.cuda.synchronize()
        print(f"objects [{lo}, {hi}): {got[0]} conjunctions, {got[2]} evasions (context path {want[0]}, {want[2]})")
        assert want[0] > 100 and want[2] > 0
        assert got == want, (lo, hi)


def test_noalloc_rejects_more_than_its_maximum():
    n = _abi.NOALLOC_MAX_OBJECTS + 1
    d = _abi.DeviceObjects()
    d.n = n
    cfg = la.EngineConfig(cutoff_km=10.0).to_c()
    with pytest.raises(la.LaddsError, match="LADDS_NOALLOC_MAX_OBJECTS"):
        _noalloc(d, cfg, 10.0)


@pytest.mark.parametrize("precision", [la.FP64, la.SPLIT], ids=["fp64", "split"])
def test_noalloc_dense_plane_sorted_in_the_arena(precision):
    """A 5,000-object cluster inside 4 km: its plane pair exceeds the pair kernel's LDS and
    is sorted in the global arena.  The context path (whose set-up sees the dense plane and
    stores the predicate's fields beside the arena records) and the no-allocation path (the
    predicate gathers from the object arrays) give the same digest."""
    pop = la.generate_population(1, 0x1ADD5 + 41, 30000, planted_pairs=100)
    rng = np.random.default_rng(41)
    c = np.array([pop.rx[0], pop.ry[0], pop.rz[0]])
    for i in range(1, 5000):
        pop.rx[i], pop.ry[i], pop.rz[i] = c + rng.uniform(-2.0, 2.0, 3)
        pop.vx[i], pop.vy[i], pop.vz[i] = pop.vx[0], pop.vy[0], pop.vz[0] + rng.uniform(-1e-3, 1e-3)
    ecfg = la.EngineConfig(dt=10.0, components=la.ALL_COMPONENTS, precision=precision, cutoff_km=10.0,
                           threshold_mode=la.THRESHOLD_ABSOLUTE, threshold_km=1.0)
    with la.Engine(ecfg) as eng:
        eng.set_objects(pop)
        want = eng.conjunction_digest(10.0)
    d, keep = _device_objects(pop, 0, pop.n)
    got = _noalloc(d, ecfg.to_c(), 10.0)
    torch.cuda.synchronize()
    print(f"dense plane: {got[0]} conjunctions (context path {want[0]})")
    assert want[0] > 100000
    assert got == want

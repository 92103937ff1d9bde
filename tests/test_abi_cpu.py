"""CPU-only checks of the engine library: it loads, exports the C ABI, and its
host-side population generator behaves as documented.  No kernels run here."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import ladds_amd as la
from ladds_amd import _abi

HEADER = Path(__file__).resolve().parents[1] / "include" / "ladds_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^LADDS_API\s+[^;(]*?\b(ladds_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = C.CDLL(str(_abi.LIB_PATH))
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi._SIGNATURES, f"{name} missing from the ctypes table"


def test_version_and_defaults():
    assert la.lib.ladds_abi_version() == 1
    assert b"gfx950" in la.lib.ladds_version()
    c = _abi.Config()
    la.lib.ladds_config_default(C.byref(c))
    # defaults of the reference's config keys (Simulation.cpp:233,245-246; default_cfg.yaml)
    assert c.dt == 10.0 and c.cutoff_km == 80.0 and c.collision_distance_factor == 1.0
    assert c.min_detection_radius_m == 0.05 and c.evasion_cutoff_km == 0.1


def test_struct_layouts():
    assert C.sizeof(_abi.Pair) == 56
    assert C.sizeof(_abi.LoggedPair) == 72
    assert la.PAIR_DTYPE.itemsize == 56


# the ctypes mirrors of the header's structures
MIRRORS = {"ladds_config": _abi.Config, "ladds_objects": _abi.Objects, "ladds_pair": _abi.Pair,
           "ladds_logged_pair": _abi.LoggedPair, "ladds_run_stats": _abi.RunStats, "ladds_tuning": _abi.Tuning,
           "ladds_adjoint_config": _abi.AdjointConfig, "ladds_device_objects": _abi.DeviceObjects}


def test_ctypes_mirrors_match_the_header(tmp_path):
    """Every ctypes mirror has the C structure's size and every field its offset, as gcc lays
    the header out (fields appended to the ABI structures must be appended to both)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ladds_amd.h"', "int main(void) {"]
    for cname, py in MIRRORS.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-I", str(HEADER.parent), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for cname, py in MIRRORS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


def test_create_without_gpu_fails_loudly():
    if la.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(la.LaddsError) as e:
        la.Engine(la.EngineConfig())
    assert e.value.code == _abi.E_NODEVICE


def test_components_mask_order():
    # bit order of the flag array at Simulation.cpp:127-135
    assert la.components_mask(kep=True) == 1
    assert la.components_mask(j2=True) == 2
    assert la.components_mask(drag=True) == 128
    assert la.components_mask(*([True] * 8)) == 0xFF


def _elements(pop, mu=398600.4418):
    r, v = pop.r(), pop.v()
    rn = np.linalg.norm(r, axis=1)
    h = np.cross(r, v)
    evec = np.cross(v, h) / mu - r / rn[:, None]
    e = np.linalg.norm(evec, axis=1)
    a = 1 / (2 / rn - (v * v).sum(1) / mu)
    inc = np.degrees(np.arccos(h[:, 2] / np.linalg.norm(h, axis=1)))
    return a, e, inc


def test_population_workload0_distributions():
    pop = la.generate_population(0, 0x1ADD5, 16384)
    a, e, inc = _elements(pop)
    hp = a * (1 - e) - 6378.1363
    assert hp.min() > 199.9 and hp.max() < 2000.1
    assert e.max() <= 0.02 + 1e-9
    frac53 = np.mean(np.abs(inc - 53) < 8)
    frac98 = np.mean(np.abs(inc - 98) < 4)
    assert 0.50 < frac53 < 0.58  # 50% normal + part of the 20% uniform
    assert 0.30 < frac98 < 0.34
    assert pop.radius_m.min() >= 0.05 and pop.radius_m.max() <= 5.0
    rho = pop.mass_kg / (4 / 3 * math.pi * pop.radius_m ** 3)
    assert rho.min() >= 500 and rho.max() <= 3000
    # Particle.h:72 and Particle.cpp:13-27
    assert np.allclose(pop.aom, math.pi * pop.radius_m ** 2 * 1e-6 / pop.mass_kg, rtol=1e-14)
    assert np.allclose(pop.bc_inv, 2.2 * math.pi * pop.radius_m ** 2 / pop.mass_kg, rtol=1e-14)


def test_population_workload1_debris_tail():
    pop = la.generate_population(1, 0x1ADD6, 20000)
    small = pop.radius_m <= 0.1
    assert 0.89 < small.mean() < 0.92
    assert pop.bc_inv.min() >= 0.005 and pop.bc_inv.max() <= 0.05
    assert pop.cr.min() >= 1.0 and pop.cr.max() <= 1.5


def test_population_workload2_clusters():
    n = 20000
    pop = la.generate_population(2, 0x1ADD7, n)
    slots = 72 * 22 + 36 * 49 + 40 * 33
    r = pop.r()
    nclustered = int(0.6 * n)
    d = np.linalg.norm(r[:nclustered - slots] - r[slots:nclustered], axis=1)
    assert d.max() <= 0.1 * math.sqrt(3) + 1e-9


def test_population_is_deterministic_and_prefix_stable():
    a = la.generate_population(0, 5, 1000)
    b = la.generate_population(0, 5, 1000)
    c = la.generate_population(0, 5, 3000)
    assert np.array_equal(a.rx, b.rx) and np.array_equal(a.vz, b.vz)
    assert np.array_equal(a.rx, c.rx[:1000])
    d = la.generate_population(0, 6, 1000)
    assert not np.array_equal(a.rx, d.rx)


def test_planted_pairs():
    n, p = 2000, 16
    pop = la.generate_population(0, 9, n, planted_pairs=p)
    nbase = n - p
    r = pop.r()
    for k in range(p):
        b = k * (nbase // p)
        d = np.linalg.norm(r[nbase + k] - r[b])
        assert 0.29 < d < 0.91


def test_cpp_adapter_header_compiles(tmp_path):
    """include/ladds_amd.hpp (the C++ adapter) compiles stand-alone with -Wall -Wextra."""
    import subprocess
    from conftest import ROOT
    src = tmp_path / "use.cpp"
    src.write_text('#include "ladds_amd.hpp"\nint main() { return 0; }\n')
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr

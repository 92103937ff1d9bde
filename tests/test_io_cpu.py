"""Input and output formats around the hot path (no device needed).

* ``read_dataset`` / ``ladds_amd::DatasetReader`` follow src/ladds/io/DatasetReader.cpp:15-50
  on the reference's own fixture data/breakup_test.csv (copied as data to tests/golden).
* ``write_conjunctions_csv`` follows the ConjunctionLogger format
  (src/ladds/io/ConjuctionLogger.cpp:13-44).
"""
import subprocess

import numpy as np

import ladds_amd as la
from conftest import GOLDEN, ROOT

EXE = ROOT / "tests" / "cpp" / "adapter_check"


def test_read_breakup_fixture():
    pop, cospar = la.read_dataset(GOLDEN / "breakup_test.csv", 2.2)
    assert cospar == ["BREAKUP_1", "BREAKUP_2"]
    assert list(pop.id) == [0, 1] and list(pop.activity) == [la.PASSIVE] * 2
    np.testing.assert_array_equal(pop.r(), [[-6751.0, 0.0001, 0.0], [-6751.0, -0.0001, 0.0]])
    np.testing.assert_array_equal(pop.v(), [[0.0, -7.68, 0.0], [0.0, 7.68, 0.0]])
    np.testing.assert_array_equal(pop.mass_kg, [10.0, 10.0])
    np.testing.assert_array_equal(pop.radius_m, [0.5, 0.5])
    # B* = 0.01 per Earth radius: bc_inv = 2 R_E[km] / 0.1570 * B*
    np.testing.assert_allclose(pop.bc_inv, 2 * 6378.1363 / 0.1570 * 0.01, rtol=1e-15)


def test_bc_inv_from_drag_coefficient_when_bstar_missing():
    bc = la.bc_inv_from([0.0, np.nan], [4.0, 4.0], [1.0, 1.0], 2.2)
    np.testing.assert_allclose(bc, [2.2 * np.pi / 4.0] * 2, rtol=1e-15)


def test_cpp_reader_matches_python(tmp_path):
    ds = tmp_path / "ds.csv"
    ds.write_text(",COSPAR_ID,NAME,BSTAR,M,R,TYPE,rx,ry,rz,vx,vy,vz\n"
                  "7,A 1,first,0.0,3.5,0.25,evasive,7000.5,-1.25,3.0,0.1,7.5,-0.2\n"
                  "9,B 2,second,0.002,120.0,1.5,evasivePreserving,-6900.0,100.0,-2.5,1.0,-7.0,0.5\n"
                  + (GOLDEN / "breakup_test.csv").read_text().splitlines()[1] + "\n")
    scen = tmp_path / "scen.txt"
    scen.write_text(f"dataset {ds} 2.2\n")
    res = subprocess.run([str(EXE), str(scen)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    rows = [ln.split() for ln in res.stdout.splitlines() if ln.startswith("data ")]
    pop, cospar = la.read_dataset(ds, 2.2)
    assert len(rows) == pop.n == 3
    for k, r in enumerate(rows):
        assert int(r[1]) == int(pop.id[k]) and int(r[-10]) == int(pop.activity[k])
        vals = np.array([float(x) for x in r[-9:]])
        np.testing.assert_array_equal(vals[:3], [pop.mass_kg[k], pop.radius_m[k], pop.bc_inv[k]])
        np.testing.assert_array_equal(vals[3:6], pop.r()[k])
        np.testing.assert_array_equal(vals[6:], pop.v()[k])
    assert [int(a) for a in pop.activity] == [la.EVASIVE, la.EVASIVE_PRESERVING, la.PASSIVE]


def test_conjunction_csv_format(tmp_path):
    log = np.zeros(3, la.LOGGED_DTYPE)
    log["iteration"] = [0, 4, 4]
    log["pair"]["id_a"] = [1, 2, 5]
    log["pair"]["id_b"] = [3, 9, 6]
    log["pair"]["d2"] = [0.1 + 0.2, 1e-5, 0.25]
    out = tmp_path / "c.csv"
    la.write_conjunctions_csv(out, log)
    assert out.read_text().splitlines() == ["Iteration,P1,P2,SquaredDistance", "0,1,3,0.30000000000000004",
                                            "4,2,9,1e-05", "4,5,6,0.25"]

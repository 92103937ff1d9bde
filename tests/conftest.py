import json
import sys
from pathlib import Path

# One HIP runtime per process: torch ships its own libamdhip64 (same soname as ROCm's).
# Importing torch first makes the engine library bind to that runtime, so the torch-based
# adjoint tests and the engine tests share one runtime in this process.
import torch  # noqa: F401

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden_cases():
    return json.loads((GOLDEN / "collision_functor_cases.json").read_text())


@pytest.fixture(scope="session")
def kepler_1s():
    return json.loads((GOLDEN / "kepler_1s.json").read_text())


@pytest.fixture(scope="session")
def pos_test_fixture():
    load = lambda n: np.loadtxt(GOLDEN / n, delimiter=",")
    return load("pos_test.csv"), load("v_test.csv"), load("pos_test_10s.csv"), load("v_test_10s.csv")

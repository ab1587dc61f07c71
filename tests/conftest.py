import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DECKS = os.path.join(GOLDEN, "decks")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmocgpu.so)")


@pytest.fixture(scope="session")
def deck_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("decks"))


@pytest.fixture(scope="session")
def c1_path(deck_dir):
    from mocc_amd import lattice_gen as lg
    return lg.write_config("c1", deck_dir)


@pytest.fixture(scope="session")
def c1(c1_path):
    from mocc_amd.native import HostProblem
    return HostProblem.load(c1_path)

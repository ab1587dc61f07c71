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
def pins3x3_path(deck_dir):
    from mocc_amd import lattice_gen as lg
    return lg.write_config("pins3x3", deck_dir)


@pytest.fixture(scope="session")
def pins3x3(pins3x3_path):
    from mocc_amd.native import HostProblem
    return HostProblem.load(pins3x3_path)


@pytest.fixture(scope="session")
def cmfd_decks(deck_dir):
    """Decks for the CMFD and eigenvalue parity tests: the 3x3-pin deck
    itself, one 17x17 assembly (289 cells) and the 3x3-assembly core geometry
    (51x51 cells, vacuum east/south) with coarser rays so the FP64 oracle
    finishes in seconds, and square lattices with vacuum east/south faces.
    A coarse solve of up to 1024 cells runs on up to 512 lanes, 1 or 2 cells
    per lane (pins3x3, asm17, lat24: 9, 289, 576 cells). Larger meshes run
    the wide solve: on 512 lanes with the rows in registers up to 3072 cells
    (lat34, lat40, lat48, core51, lat55: 1156, 1600, 2304, 2601, 3025 cells at
    3, 4, 5, 6, 6 cells per lane), on 1024 lanes re-reading the rows from
    cache beyond (lat64, lat70, lat78: 4096, 4900, 6084 cells at 4, 5, 6). A
    batch runs the same two solves, up to 3072 cells."""
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    paths = {
        "pins3x3": lg.write_config("pins3x3", deck_dir),
        "asm17": lg.write_config("assembly17", deck_dir, name="asm17_cmfd", n_azimuthal=4, spacing=0.1, groups=2),
        "core51": lg.write_config("core51", deck_dir, name="core51_cmfd", n_azimuthal=4, spacing=0.1),
        "lat48": lg.write_deck(lg.pin_lattice(48, 2, 4, 0.1, "lat48_cmfd", cmfd=True,
                                              bc={"east": "vacuum", "south": "vacuum"}), deck_dir, 48),
    }
    for n in (24, 34, 40, 55, 64, 70, 78):
        paths[f"lat{n}"] = lg.write_deck(lg.pin_lattice(n, 2, 4, 0.1, f"lat{n}_cmfd", cmfd=True,
                                                        bc={"east": "vacuum", "south": "vacuum"}), deck_dir, n)
    return {k: HostProblem.load(v) for k, v in paths.items()}

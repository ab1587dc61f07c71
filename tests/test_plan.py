"""The band plan (the sweep's HBM layout) against the problem it encodes,
built on the host: every segment exactly once per direction (backward
reversed), padding steps in the lane's own dummy slot, band-local FSR slots
and their per-FSR lists, longest-first and Gauss-Seidel band orders, current
events. CPU only."""
import ctypes as C

import pytest

from mocc_amd.native import gpu


def _check(hp, width):
    buf = C.create_string_buffer(512)
    rc = gpu().mocgpu_plan_check(hp.view, width, buf, 512)
    return rc, buf.value.decode()


@pytest.mark.parametrize("width", [64, 48, 17, 1])
def test_plan_pins3x3(pins3x3, width):
    rc, msg = _check(pins3x3, width)
    assert rc == 0, msg
    assert msg.startswith("ok")


def test_plan_vacuum_and_cmfd_geometry(cmfd_decks):
    for name in ("asm17", "lat48", "core51"):
        rc, msg = _check(cmfd_decks[name], 64)
        assert rc == 0, (name, msg)


def test_plan_assembly17(deck_dir):
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    hp = HostProblem.load(lg.write_config("assembly17", deck_dir))
    rc, msg = _check(hp, 64)
    assert rc == 0, msg


@pytest.mark.parametrize("width", [64, 32])
def test_plan_independent_of_thread_count(deck_dir, monkeypatch, width):
    """A problem of over a million segments builds its azimuths' bands on
    several threads (MOCC_HOST_THREADS) and joins them in azimuth order: the
    plan's digest is the same at any thread count."""
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    hp = HostProblem.load(lg.write_deck(lg.pin_lattice(24, 2, 8, 0.05, "plan24v", cmfd=True,
                                                       bc={"east": "vacuum", "south": "vacuum"}), deck_dir, 24))
    assert hp.p.n_seg > 1_000_000
    digests = set()
    for n in ("1", "3", "8"):
        monkeypatch.setenv("MOCC_HOST_THREADS", n)
        rc, msg = _check(hp, width)
        assert rc == 0, msg
        digests.add(msg.rsplit("digest ", 1)[1])
    assert len(digests) == 1

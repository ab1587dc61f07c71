"""Device CMFD (FP64) against the oracle's CMFD::solve restatement
(src/core/cmfd.cpp:171-638, eigen_solver.cpp:263-291) on identical inputs:
the fine flux and the coarse-surface currents of three oracle outer
iterations are loaded into the device handle, then one do_cmfd runs on each.

With tight CMFD tolerances both converge to the same coarse solution, so the
matrix, k, the coarse flux and the rescaled fine flux agree to ~1e-12
relative (bar 1e-8); with the default tolerances k agrees to the tolerance
(bar 1e-6).
"""
import numpy as np
import pytest

from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

TIGHT = dict(cmfd_k_tol=1e-12, cmfd_psi_tol=1e-12, cmfd_resid_reduction=1e-10, cmfd_max_iter=500)


def _oracle_state(hp, outer=3, **cmfd):
    ora = Oracle(hp, n_inner=hp.params().n_inner, use_cmfd=True, **cmfd)
    ora.initialize()
    ora.calc_fission_source(1.0)
    for _ in range(outer):
        ora.fss_step(with_current=True)
    return ora


def _device_from(hp, ora, **cmfd):
    from mocc_amd.sweeper import GpuMoCSweeper
    dev = GpuMoCSweeper(hp, use_cmfd=True, **cmfd)
    dev.initialize()
    dev.set_flux(ora.flux())
    for g in range(hp.n_group):
        dev.set_currents(g, *ora.currents(g))
    return dev


def _check_pair(hp, tight=True):
    cm = TIGHT if tight else {}
    print(f"case F={hp.n_fsr} G={hp.n_group} cells={hp.p.nx * hp.p.ny} tight={tight}", flush=True)
    ora = _oracle_state(hp, **cm)
    dev = _device_from(hp, ora, **cm)
    kd = dev.cmfd_step(1.0)
    ko, _ = ora.cmfd_step(1.0)
    return dev, ora, kd, ko


# coarse-solve instantiations: 1 cell per lane (pins3x3 on one wave, asm17 on
# 320 lanes) and 2 (lat24) on up to 512 lanes; on 1024 lanes, 2 and 3 per lane
# with the rows in LDS (lat34, lat40; lat48, core51, lat55) and 4, 5, 6
# re-reading them from cache (lat64, lat70, lat78: past the 512-lane solve's
# 3072 cells)
@pytest.mark.parametrize("deck", ["pins3x3", "asm17", "lat24", "lat34", "lat40", "lat48", "core51", "lat55",
                                  "lat64", "lat70", "lat78"])
def test_cmfd_step_parity(deck, cmfd_decks):
    hp = cmfd_decks[deck]
    dev, ora, kd, ko = _check_pair(hp, tight=True)
    for g in range(hp.n_group):
        rd, cd, vd = dev.cmfd_matrix(g)
        ro, co, vo = ora.cmfd_matrix(g)
        assert np.array_equal(rd, ro) and np.array_equal(cd, co)
        np.testing.assert_allclose(vd, vo, rtol=1e-12, atol=0)
    assert abs(kd - ko) < 1e-10, (kd, ko)
    np.testing.assert_allclose(dev.cmfd_flux(), ora.cmfd_flux(), rtol=1e-8, atol=0)
    np.testing.assert_allclose(dev.flux(), ora.flux(), rtol=1e-8, atol=0)
    for g in range(hp.n_group):
        c_d, s_d = dev.currents(g)
        c_o, s_o = ora.currents(g)
        scale = np.abs(c_o).max()
        assert np.abs(c_d - c_o).max() <= 1e-8 * scale
        np.testing.assert_allclose(s_d, s_o, rtol=1e-8, atol=1e-14 * np.abs(s_o).max())


@pytest.mark.parametrize("deck", ["pins3x3", "core51"])
def test_cmfd_step_default_tolerances(deck, cmfd_decks):
    hp = cmfd_decks[deck]
    dev, ora, kd, ko = _check_pair(hp, tight=False)
    assert abs(kd - ko) < 1e-6, (kd, ko)
    assert dev.cmfd_iterations() >= 1


def test_cmfd_rejects_without_option(pins3x3):
    from mocc_amd.sweeper import DeviceError, GpuMoCSweeper
    dev = GpuMoCSweeper(pins3x3, n_inner=1, use_cmfd=False)
    dev.initialize()
    with pytest.raises(DeviceError):
        dev.cmfd_step(1.0)


def test_cmfd_negative_fixup_parity(cmfd_decks):
    """<cmfd negative_fixup> (cmfd.cpp:152-154, 175-184): a fine flux with one
    pin made negative, identical on both sides; with the fixup the device
    clamps that coarse flux to zero before the solve as the oracle does, and
    the result differs from the solve without it."""
    hp = cmfd_decks["lat24"]
    out = {}
    for fixup in (1, 0):
        cm = dict(TIGHT, cmfd_negative_fixup=fixup)
        print(f"negative fixup={fixup}", flush=True)
        ora = _oracle_state(hp, **cm)
        phi = ora.flux()
        off = hp.array("pin_fsr_off", hp.p.nx * hp.p.ny + 1, np.int64)
        ip = hp.p.nx * hp.p.ny // 2 + hp.p.nx // 2  # a pin inside the lattice
        phi[off[ip]:off[ip + 1], 0] *= -0.5
        ora.set_flux(phi)
        dev = _device_from(hp, ora, **cm)
        kd = dev.cmfd_step(1.0)
        ko, _ = ora.cmfd_step(1.0)
        assert abs(kd - ko) < 1e-10, (fixup, kd, ko)
        np.testing.assert_allclose(dev.cmfd_flux(), ora.cmfd_flux(), rtol=1e-8, atol=0)
        np.testing.assert_allclose(dev.flux(), ora.flux(), rtol=1e-8, atol=0)
        out[fixup] = kd
    assert abs(out[1] - out[0]) > 1e-9, out


def test_cmfd_second_solve_bitwise(cmfd_decks):
    """mocgpu_initialize drops the currents of an earlier solve: solving one
    CMFD handle twice gives bitwise the same k, iteration count and flux
    (each solve's first CMFD step runs with D-hat = 0)."""
    from mocc_amd.sweeper import GpuMoCSweeper
    hp = cmfd_decks["pins3x3"]
    dev = GpuMoCSweeper(hp, n_inner=5, use_cmfd=True)
    r1, s1 = dev.eigen_solve(1e-7, 1e-6, 200)
    f1 = dev.flux()
    r2, s2 = dev.eigen_solve(1e-7, 1e-6, 200)
    assert r1 == r2 == 0
    assert s1.k == s2.k and s1.iterations == s2.iterations and s1.err_psi == s2.err_psi
    assert np.array_equal(dev.flux(), f1)

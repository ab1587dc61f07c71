"""The 2-D/3-D correction worker and the incoming-flux update on the device
against the FP64 oracle:

  * MoCSweeper_2D3D::sweep (moc_sweeper_2d3d.cpp:48-99) with the
    cmdo::CurrentCorrections worker (correction_worker.hpp:36-292,
    correction_worker.cpp:33-160): alpha, beta, the residuals and the coarse
    currents from identical inputs (the oracle's state loaded into the
    device; the device keeps boundary fluxes in FP32, so agreement is to
    ~1e-7 relative, bar 1e-6), on a lattice with moderator-only pins and
    vacuum faces, in both boundary-update orders;
  * the flat-flux closed form alpha = 1/2, beta = 1 on the infinite medium;
  * MoCSweeper::update_incoming_flux (moc_sweeper.cpp:255-291) after the
    first and after a later CMFD solve.
"""
import numpy as np
import pytest

from conftest import DECKS
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lat4(deck_dir):
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    d = lg.pin_lattice(4, 2, 8, 0.05, "corr4x4_gpu", cmfd=True, bc={"east": "vacuum", "south": "vacuum"})
    rows = d.lattices[1][2]
    rows[1][2] = rows[2][1] = 2  # moderator-only pins
    return HostProblem.load(lg.write_deck(d, deck_dir, 7))


def _pair(hp, gs, n_inner=1, outer=2):
    from mocc_amd.sweeper import GpuMoCSweeper
    ora = Oracle(hp, n_inner=n_inner, use_cmfd=True, gs_boundary=gs)
    ora.initialize()
    ora.calc_fission_source(1.0)
    for _ in range(outer):
        ora.fss_step()
    dev = GpuMoCSweeper(hp, n_inner=n_inner, use_cmfd=True, gs_boundary=gs, correction_worker=1)
    dev.initialize()
    dev.set_flux(ora.flux())
    dev.set_boundary(ora.boundary())
    dev.set_fission_source(ora.calc_fission_source(1.0))
    return dev, ora


@pytest.mark.parametrize("gs", [True, False])
def test_corrections_parity(lat4, gs):
    hp = lat4
    print(f"2D3D correction sweep, 4x4 lattice, gs={gs}", flush=True)
    dev, ora = _pair(hp, gs)
    for g in range(hp.n_group):
        for s in (dev, ora):
            s.build_source(g)
            s.sweep_2d3d(g)
    ad, bd = dev.corrections()
    ao, bo = ora.corrections()
    assert np.all(np.isfinite(ad)) and np.all(np.isfinite(bd))
    np.testing.assert_allclose(ad, ao, rtol=1e-6, atol=0)
    np.testing.assert_allclose(bd, bo, rtol=1e-6, atol=0)
    for g in range(hp.n_group):
        np.testing.assert_allclose(dev.correction_residual(g), ora.correction_residual(g), rtol=1e-5, atol=1e-12)
        cd, sd = dev.currents(g)
        co, so = ora.currents(g)
        assert np.abs(cd - co).max() <= 1e-6 * np.abs(co).max()
        assert np.abs(sd - so).max() <= 1e-6 * np.abs(so).max()
    # the band sweep of the last inner iteration still updated flux and boundary
    np.testing.assert_allclose(dev.flux(), ora.flux(), rtol=1e-4, atol=0)


def test_corrections_flat_flux_device():
    """alpha = 1/2 and beta = 1 on the infinite medium at its exact solution."""
    from mocc_amd.native import HostProblem
    from mocc_amd.sweeper import GpuMoCSweeper
    from test_gpu_eigen import IHM_XML
    hp = HostProblem.from_string(IHM_XML, DECKS)
    G = hp.n_group
    tr = hp.array("xs_tr", G)
    sc = hp.array("xs_sc", G * G).reshape(G, G)
    ch = hp.array("xs_ch", G)
    nf = hp.array("xs_nf", G)
    phi = np.linalg.solve(np.diag(tr) - sc, ch)
    dev = GpuMoCSweeper(hp, n_inner=1, use_cmfd=True, correction_worker=1)
    dev.initialize()
    dev.set_flux(np.tile(phi, (hp.n_fsr, 1)))
    dev.calc_fission_source(float(nf @ phi))
    bc = dev.boundary()
    dev.set_boundary(np.repeat(phi / (4 * np.pi), bc.size // G))
    for g in range(G):
        dev.build_source(g)
        dev.sweep_2d3d(g)
    a, b = dev.corrections()
    np.testing.assert_allclose(a, 0.5, rtol=0, atol=1e-6)
    np.testing.assert_allclose(b, 1.0, rtol=0, atol=1e-6)


def test_update_incoming_flux_parity(lat4):
    """The incoming flux after the first CMFD solve (set from its partial
    currents) and after the second (scaled by the ratio to the first), on
    identical inputs."""
    from mocc_amd.sweeper import GpuMoCSweeper
    hp = lat4
    TIGHT = dict(cmfd_k_tol=1e-12, cmfd_psi_tol=1e-12, cmfd_resid_reduction=1e-10, cmfd_max_iter=500)
    ora = Oracle(hp, n_inner=1, use_cmfd=True, **TIGHT)
    ora.initialize()
    ora.calc_fission_source(1.0)
    for _ in range(3):
        ora.fss_step(with_current=True)
    dev = GpuMoCSweeper(hp, n_inner=1, use_cmfd=True, **TIGHT)
    dev.initialize()
    dev.set_flux(ora.flux())
    dev.set_boundary(ora.boundary())
    for g in range(hp.n_group):
        dev.set_currents(g, *ora.currents(g))
    for step in range(2):
        print(f"CMFD solve {step + 1} then update_incoming_flux", flush=True)
        kd = dev.cmfd_step(1.0)
        ko, _ = ora.cmfd_step(1.0)
        assert abs(kd - ko) < 1e-10
        dev.update_incoming_flux()
        ora.update_incoming_flux()
        bd, bo = dev.boundary(), ora.boundary()
        m = bo != 0
        assert np.array_equal(bd == 0, bo == 0)
        assert np.abs(bd[m] / bo[m] - 1.0).max() < 1e-6
        # the next solve starts from these fluxes: the fission source at the
        # new eigenvalue, then one more transport step on both sides
        for s in (dev, ora):
            s.calc_fission_source(ko)
            s.fss_step(with_current=True)
        # and the oracle's state again, so the second solve has identical
        # inputs while the device keeps its own partial currents of the first
        dev.set_flux(ora.flux())
        dev.set_boundary(ora.boundary())
        for g in range(hp.n_group):
            dev.set_currents(g, *ora.currents(g))


def test_correction_worker_errors(pins3x3):
    from mocc_amd.sweeper import DeviceError, GpuMoCSweeper
    plain = GpuMoCSweeper(pins3x3, n_inner=1, use_cmfd=True)
    with pytest.raises(DeviceError, match="correction worker"):
        plain.sweep_2d3d(0)
    no_cmfd = GpuMoCSweeper(pins3x3, n_inner=1, use_cmfd=False, correction_worker=1)
    no_cmfd.initialize()
    with pytest.raises(DeviceError, match="coarse data"):
        no_cmfd.sweep_2d3d(0)
    with pytest.raises(DeviceError, match="partial currents"):
        no_cmfd.update_incoming_flux()

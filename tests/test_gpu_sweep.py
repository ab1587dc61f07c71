"""Device sweep parity against the FP64 oracle (moc_sweeper_kernel.inc.hpp:36-180).

Both boundary updates are compared: Gauss-Seidel in angle (mocc's default,
two phases on the device) and Jacobi; the oracle runs the same mode, so one
sweep of identical inputs must agree to FP32 accuracy: the bar is 1e-4
relative on FSR scalar flux, observed ~1e-7.
"""
import numpy as np
import pytest

from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

FLUX_RTOL = 1e-4


def _setup_pair(hp, energy_mode, seed=0, gs=False, band_rays=0):
    from mocc_amd.sweeper import GpuMoCSweeper
    F, G = hp.n_fsr, hp.n_group
    rng = np.random.default_rng(seed)
    phi0 = rng.uniform(0.5, 2.0, (F, G))
    dev = GpuMoCSweeper(hp, n_inner=1, energy_mode=energy_mode, gs_boundary=gs, band_rays=band_rays)
    ora = Oracle(hp, n_inner=1, gs_boundary=gs, energy_mode=energy_mode)
    for s in (dev, ora):
        s.initialize()
        s.set_flux(phi0)
        s.calc_fission_source(1.0)
    # random incoming angular flux on the reflective faces
    bc = ora.boundary()
    bc = np.where(bc != 0.0, rng.uniform(0.02, 0.1, bc.size), 0.0)
    ora.set_boundary(bc)
    dev.set_boundary(bc)
    return dev, ora


def _rel(a, b):
    m = np.abs(b) > 0
    return np.max(np.abs(a[m] / b[m] - 1.0))


def test_attenuation_matches_reference_interpolant():
    """The kernel's in-register 1 - exp(-tau) against the FP64 oracle of
    Exponential_Linear<10000> (exponential.hpp:56-112; exp itself past
    tau = 10), at the coordinate w = fl(1000 tau) the kernel interpolates at."""
    import ctypes as C
    import math
    from mocc_amd.native import gpu
    from oracle.binding import exp_table
    rng = np.random.default_rng(7)
    tau = np.concatenate([
        rng.uniform(0, 10, 20000), rng.uniform(0, 0.01, 4000), rng.uniform(0, 1e-5, 500),
        rng.uniform(10, 45, 1000), np.arange(0, 12000) / 1000.0,  # the nodes
        [0.0, 1e-30, 0.2499, 0.25, 0.2501, 9.999, 10.0, 10.001, 39.999, 40.0, 1e3, 1e30],
    ]).astype(np.float32)
    out = np.zeros_like(tau)
    assert gpu().mocgpu_attenuation(tau.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                    tau.size) == 0
    w = np.minimum(tau * np.float32(1000.0), np.float32(39999.0)).astype(np.float64)
    t = w / 1000.0
    ref = np.array([1.0 - exp_table(-v) if v < 10.0 else 1.0 - math.exp(-v) for v in t])
    e = out.astype(np.float64)
    assert e[tau == 0].max() == 0.0
    assert np.abs(e - ref).max() < 1.5e-7
    m = ref > 1e-9  # below, the FP64 interpolant's own cancellation (~1e-16) is not small
    assert np.max(np.abs(e[m] / ref[m] - 1.0)) < 3e-7
    assert e[-1] == 1.0 and e[-2] == 1.0


@pytest.mark.parametrize("gs", [True, False])
@pytest.mark.parametrize("energy_mode", [1, 0])
def test_pins3x3_fss_step_parity(pins3x3, energy_mode, gs):
    dev, ora = _setup_pair(pins3x3, energy_mode, gs=gs)
    for s in (dev, ora):
        s.fss_step(with_current=True)
    phi_d, phi_o = dev.flux(), ora.flux()
    assert _rel(phi_d, phi_o) < FLUX_RTOL
    bd, bo = dev.boundary(), ora.boundary()
    assert np.all((bd == 0) == (bo == 0))
    assert _rel(bd, bo) < 1e-5
    p = pins3x3.p
    n0, n1 = p.nx * p.ny, p.nx * p.ny + (p.nx + 1) * p.ny + (p.ny + 1) * p.nx
    for g in range(pins3x3.n_group):
        cd, sd = dev.currents(g)
        co, so = ora.currents(g)
        scale = np.abs(co[n0:n1]).max()
        assert np.abs(cd[n0:n1] - co[n0:n1]).max() <= 1e-5 * scale
        assert _rel(sd[n0:n1], so[n0:n1]) < 1e-5


@pytest.mark.parametrize("gs", [True, False])
def test_pins3x3_one_wave_workgroup_parity(pins3x3, gs):
    """Bands of at most 32 rays run the one-wave workgroup (32 rays per
    direction, 64 lanes): the same parity against the oracle, currents
    included."""
    dev, ora = _setup_pair(pins3x3, 1, seed=5, gs=gs, band_rays=32)
    assert dev.options().band_rays == 32
    for s in (dev, ora):
        s.fss_step(with_current=True)
    assert _rel(dev.flux(), ora.flux()) < FLUX_RTOL
    assert _rel(dev.boundary(), ora.boundary()) < 1e-5
    p = pins3x3.p
    n0, n1 = p.nx * p.ny, p.nx * p.ny + (p.nx + 1) * p.ny + (p.ny + 1) * p.nx
    for g in range(pins3x3.n_group):
        cd, _ = dev.currents(g)
        co, _ = ora.currents(g)
        assert np.abs(cd[n0:n1] - co[n0:n1]).max() <= 1e-5 * np.abs(co[n0:n1]).max()


def test_pins3x3_reproducible(pins3x3):
    """Integer band tallies: two runs give bitwise identical fluxes."""
    outs = []
    for _ in range(2):
        dev, _ora = _setup_pair(pins3x3, 1, seed=3)
        dev.fss_step(with_current=True)
        dev.fss_step(with_current=False)
        outs.append(dev.flux())
    assert np.array_equal(outs[0], outs[1])


def test_pins3x3_fission_and_pin_flux(pins3x3):
    dev, ora = _setup_pair(pins3x3, 0)
    fd, fo = dev.calc_fission_source(1.1), ora.calc_fission_source(1.1)
    assert np.allclose(fd, fo, rtol=1e-14, atol=0)
    assert dev.total_fission() == pytest.approx(ora.total_fission(), rel=1e-13)
    for g in range(pins3x3.n_group):
        pd, po = dev.get_pin_flux_1g(g), ora.pin_flux(g)
        assert np.allclose(pd, po, rtol=1e-13)
        rd = dev.set_pin_flux_1g(g, pd * 1.5)
        ro = ora.set_pin_flux(g, po * 1.5)
        assert rd == pytest.approx(ro, rel=1e-10)
    assert np.allclose(dev.flux(), ora.flux(), rtol=1e-13)


@pytest.fixture(scope="module")
def assembly17(deck_dir):
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    return HostProblem.load(lg.write_config("assembly17", deck_dir))


@pytest.mark.parametrize("gs", [True, False])
def test_assembly17_fss_step_parity(assembly17, gs):
    """The benchmark workload itself (17x17, 7 groups, 6.7M segments): one
    outer step with currents agrees with the FP64 oracle to 1e-4 (observed
    ~1e-6) in every FSR and group."""
    print(f"assembly17 gs={gs}", flush=True)
    dev, ora = _setup_pair(assembly17, 0, seed=1, gs=gs)
    for s in (dev, ora):
        s.fss_step(with_current=True)
    assert _rel(dev.flux(), ora.flux()) < FLUX_RTOL
    assert _rel(dev.boundary(), ora.boundary()) < 1e-4


@pytest.mark.parametrize("gs", [True, False])
def test_lane_maps_bitwise(pins3x3, assembly17, gs):
    """The three ray-to-lane maps (plain, ray-parity interleave, length
    split) give bitwise the same flux, boundary fluxes and coarse currents:
    the band tallies are integer sums, whose order does not matter. The
    17x17 deck's 64-ray bands exercise the length split's second wave."""
    from mocc_amd.sweeper import GpuMoCSweeper
    for hp in (pins3x3, assembly17):
        rng = np.random.default_rng(11)
        phi0 = rng.uniform(0.5, 2.0, (hp.n_fsr, hp.n_group))
        outs = []
        for m in (0, 1, 2):
            print(f"lane map {m}, {hp.n_fsr} FSRs, gs={gs}", flush=True)
            dev = GpuMoCSweeper(hp, n_inner=1, energy_mode=1, gs_boundary=gs, lane_map=m)
            assert dev.options().lane_map == m
            dev.initialize()
            dev.set_flux(phi0)
            dev.calc_fission_source(1.0)
            dev.fss_step(with_current=True)
            outs.append([dev.flux(), dev.boundary()] + [c for g in range(hp.n_group) for c in dev.currents(g)])
            dev.close()
        for o in outs[1:]:
            for a, b in zip(outs[0], o):
                assert np.array_equal(a, b)


def test_lane_map_checked(pins3x3):
    from mocc_amd.sweeper import DeviceError, GpuMoCSweeper
    with pytest.raises(DeviceError, match="lane_map"):
        GpuMoCSweeper(pins3x3, lane_map=3)


def test_autotuned_lane_map(pins3x3):
    """lane_map = -1: the autotuner picks a band width and lane map by timing
    the candidates; the tuned handle sweeps as accurately as the default."""
    from mocc_amd.sweeper import GpuMoCSweeper
    dev = GpuMoCSweeper(pins3x3, n_inner=1, energy_mode=0, gs_boundary=True, lane_map=-1)
    o = dev.options()
    assert o.band_rays in (32, 48, 64) and o.lane_map in (0, 1, 2)
    ora = Oracle(pins3x3, n_inner=1, gs_boundary=True, energy_mode=0)
    rng = np.random.default_rng(2)
    phi0 = rng.uniform(0.5, 2.0, (pins3x3.n_fsr, pins3x3.n_group))
    for s in (dev, ora):
        s.initialize()
        s.set_flux(phi0)
        s.calc_fission_source(1.0)
        s.fss_step(with_current=False)
    assert _rel(dev.flux(), ora.flux()) < FLUX_RTOL


@pytest.mark.parametrize("gs", [True, False])
def test_flux_residual_parity(pins3x3, gs):
    """TransportSweeper::flux_residual (transport_sweeper.cpp:210-223): the
    2-norm of phi - phi_old after an outer step, against the oracle's."""
    dev, ora = _setup_pair(pins3x3, 0, seed=11, gs=gs)
    for s in (dev, ora):
        s.fss_step()
    rd, ro = dev.flux_residual(), ora.flux_residual()
    print(f"flux residual device {rd:.10e} oracle {ro:.10e}", flush=True)
    assert ro > 0 and abs(rd / ro - 1.0) < 1e-5
    # a second step, closer to the fixed point: the residual shrinks on both
    for s in (dev, ora):
        s.fss_step()
    rd2, ro2 = dev.flux_residual(), ora.flux_residual()
    assert rd2 < rd and abs(rd2 / ro2 - 1.0) < 1e-4


def test_group_index_checked(pins3x3):
    """Entry points taking a group index reject g < 0 and g >= n_group with
    MOCGPU_ERR_ARG before touching device memory."""
    from mocc_amd.native import gpu, ptr
    from mocc_amd.sweeper import GpuMoCSweeper
    dev = GpuMoCSweeper(pins3x3, n_inner=1)
    dev.initialize()
    L = gpu()
    G, npin = pins3x3.n_group, dev.npin
    pin = np.ones(npin)
    cur, sfl = np.zeros(L.mocgpu_n_surf(dev.h)), np.zeros(L.mocgpu_n_surf(dev.h))
    r = np.zeros(1)
    for g in (-1, G):
        assert L.mocgpu_get_pin_flux_1g(dev.h, g, ptr(pin)) == -1
        assert L.mocgpu_set_pin_flux_1g(dev.h, g, ptr(pin), ptr(r)) == -1
        assert L.mocgpu_get_currents(dev.h, g, ptr(cur), ptr(sfl)) == -1
        assert L.mocgpu_set_currents(dev.h, g, ptr(cur), ptr(sfl)) == -1
        assert L.mocgpu_sweep(dev.h, g, 0) == -1
        assert b"group out of range" in L.mocgpu_last_error(dev.h)
    # the handle still works, and a non-positive pin flux is still reported
    assert L.mocgpu_get_pin_flux_1g(dev.h, 0, ptr(pin)) == 0
    pin[0] = 0.0
    assert L.mocgpu_set_pin_flux_1g(dev.h, 0, ptr(pin), ptr(r)) == -4

"""The FP64 oracle pinned against the reference's test values and against
closed forms and invariants. CPU only.

  * exp table: src/core/tests/test_Exponential.cpp:27-92;
  * infinite homogeneous medium: src/sweepers/moc/tests/test_MoC_IHM.cpp:99-185
    (closed form phi = M^-1 chi, k = nu Sigma_f . phi);
  * pin flux get/set identity: src/sweepers/moc/tests/test_MoCSweeper.cpp:58-94;
  * per-pin neutron balance: MoCSweeper::check_balance
    (src/sweepers/moc/moc_sweeper.cpp:479-532) at inner convergence, which
    pins the coarse-surface current tally (moc_current_worker.hpp:202-264),
    on the 3x3 deck and on even lattices whose rays graze pin corners, with
    every ray's coarse data walked against its segments;
  * CMFD on a flat-flux infinite medium reduces to the k-infinity closed form
    (no leakage, D-hat = 0), and the BiCGSTAB solve agrees with a dense solve;
  * CMFD consistency on a heterogeneous lattice with vacuum faces: given a
    converged transport solution and its coarse-surface currents, the
    flux-weighted homogenization (xs_mesh_homogenized.cpp:261-363) and the
    D-hat coupling (cmfd.cpp:479-490) reproduce the transport eigenvalue and
    leave every pin flux unchanged; a wrong flux weight or a sign error in
    D-hat breaks both;
  * the eigenvalue power iteration (eigen_solver.cpp:131-253) converges to
    the infinite-medium k and spectrum (test_MoC_IHM.cpp:152-185), in both
    energy modes.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import DECKS
from mocc_amd.native import HostProblem, gpu
from oracle.binding import Oracle, exp_table, exp_table_node, lib

REAL_FUZZ = 200 * np.finfo(float).eps  # util/fp_utils.hpp:39

IHM_XML = """<mesh id="1" type="rect" pitch="1.26"><sub_x>3</sub_x><sub_y>3</sub_y></mesh>
<pin id="1" mesh="1">1 1 1 1 1 1 1 1 1</pin>
<lattice id="1" nx="3" ny="2">1 1 1 1 1 1</lattice>
<assembly id="1" np="1" hz="0.5"><lattices>1</lattices></assembly>
<core nx="1" ny="1" north="reflect" south="reflect" east="reflect" west="reflect" top="reflect" bottom="reflect">1</core>
<material_lib path="c5g7.xsl"><material id="1" name="UO2-3.3" /></material_lib>
<source scattering="P0" />
<sweeper type="moc" n_inner="{n_inner}"><ang_quad type="ls" order="2" /><rays spacing="0.01" /></sweeper>"""


# --- exponential table ----------------------------------------------------------

def test_exp_table_coarse():
    n, lo, hi = 5, -5.3, 0.0
    for i, x in enumerate([-5.3, -4.24, -3.18, -2.12, -1.06, 0.0]):
        assert abs(exp_table_node(i, n, lo, hi) - math.exp(x)) < REAL_FUZZ
        assert abs(exp_table(x, n, lo, hi) - exp_table_node(i, n, lo, hi)) < REAL_FUZZ
    assert abs(exp_table(-5.088, n, lo, hi) - 6.87479349415065e-03) < REAL_FUZZ
    assert abs(exp_table(-2.756, n, lo, hi) - 7.29640444772866e-02) < REAL_FUZZ


def test_exp_table_10000():
    x = -10.0
    while x < 0.0:
        assert abs(exp_table(x) - math.exp(x)) < 2e-8
        x += 0.1
    # the linear interpolant's bound h^2/8 max|f''| on [-10, 0]
    xs = np.linspace(-10.0, -1e-9, 20011)
    err = max(abs(exp_table(v) - math.exp(v)) for v in xs)
    assert err <= 1.25e-7 * (1 + 1e-9)


# --- sweeps against closed forms ------------------------------------------------

@pytest.fixture(scope="module")
def ihm():
    hp = HostProblem.from_string(IHM_XML.format(n_inner=3000), DECKS)
    G = hp.n_group
    tr = hp.array("xs_tr", G)
    nf = hp.array("xs_nf", G)
    ch = hp.array("xs_ch", G)
    sc = hp.array("xs_sc", G * G).reshape(G, G)
    phi = np.linalg.solve(np.diag(tr) - sc, ch)
    return hp, phi, float(nf @ phi)


def test_ihm_closed_form(ihm):
    hp, phi, k = ihm
    o = Oracle(hp, n_inner=3000)
    o.initialize()
    o.set_flux(np.tile(phi, (hp.n_fsr, 1)))
    fs = o.calc_fission_source(k)
    assert np.abs(fs - 1.0).max() < 1e-13
    for g in range(hp.n_group):
        o.build_source(g)
        o.sweep(g)
    f = o.flux()
    assert (np.abs(f - phi) / phi).max() < 1e-12


def test_ihm_cmfd_k_infinity(ihm):
    """Flat flux, no leakage: one CMFD solve returns the closed-form k and
    leaves the flat flux unchanged."""
    hp, phi, k = ihm
    o = Oracle(hp, n_inner=1, use_cmfd=True, cmfd_k_tol=1e-12, cmfd_psi_tol=1e-12,
               cmfd_resid_reduction=1e-10, cmfd_max_iter=200)
    o.initialize()
    o.set_flux(np.tile(phi, (hp.n_fsr, 1)))
    k1, _ = o.cmfd_step(0.9)
    assert abs(k1 - k) < 1e-9
    ratio = o.flux() / phi[None, :]
    np.testing.assert_allclose(ratio, ratio[0, 0], rtol=1e-8)


def test_ihm_eigen_solve_k_infinity():
    """EigenSolver::solve on the infinite homogeneous medium converges to
    k = nu Sigma_f . M^-1 chi with the flux spectrum M^-1 chi in every FSR;
    Gauss-Seidel and Jacobi energy sweeps reach the same k. A swapped old/new
    total fission in the k update or a missing normalisation fails it."""
    hp = HostProblem.from_string(IHM_XML.format(n_inner=1), DECKS)
    G = hp.n_group
    tr = hp.array("xs_tr", G)
    nf = hp.array("xs_nf", G)
    ch = hp.array("xs_ch", G)
    sc = hp.array("xs_sc", G * G).reshape(G, G)
    phi = np.linalg.solve(np.diag(tr) - sc, ch)
    kinf = float(nf @ phi)
    for energy_mode in (0, 1):
        o = Oracle(hp, n_inner=1, energy_mode=energy_mode)
        rc, st, hist = o.eigen_solve(1e-12, 1e-10, 5000)
        assert rc == 0 and st.converged
        assert abs(st.k - kinf) < 1e-9, (energy_mode, st.k, kinf)
        f = o.flux()
        spec = f / f[:, :1]
        np.testing.assert_allclose(spec, np.tile(phi / phi[0], (hp.n_fsr, 1)), rtol=1e-7, atol=0)
        # the k history approaches k-infinity from the flat start
        assert abs(hist[-1] - kinf) < abs(hist[0] - kinf)


def _vacuum_lattice(deck_dir):
    from mocc_amd import lattice_gen as lg
    d = lg.pin_lattice(3, 2, 16, 0.05, "pin3x3_vac_es", bc={"east": "vacuum", "south": "vacuum"}, cmfd=True)
    return HostProblem.load(lg.write_deck(d, deck_dir, 1))


TIGHT_CMFD = dict(cmfd_k_tol=1e-13, cmfd_psi_tol=1e-13, cmfd_resid_reduction=1e-12, cmfd_max_iter=1000)


def _converged_with_currents(hp, **cmfd):
    """A transport solution converged to k_tol 1e-13 / psi_tol 1e-11 without
    CMFD, then one outer step at that k tallying the coarse currents; returns
    the oracle (CMFD enabled, holding that state) and k."""
    t = Oracle(hp, n_inner=1, use_cmfd=False)
    rc, st, _ = t.eigen_solve(1e-13, 1e-11, 5000)
    assert rc == 0 and st.converged
    o = Oracle(hp, n_inner=1, use_cmfd=True, **cmfd)
    o.initialize()
    o.set_flux(t.flux())
    o.set_boundary(t.boundary())
    o.calc_fission_source(st.k)
    o.fss_step(with_current=True)
    return o, st.k


def test_cmfd_consistency_heterogeneous(deck_dir):
    """On a converged heterogeneous transport solution (3x3 pins, fuel and
    moderator rings, vacuum east and south faces) one CMFD solve returns the
    transport eigenvalue and rescales no pin: the coarse equations with
    flux-weighted cross sections and D-hat from the tallied currents are
    satisfied by the transport pin fluxes themselves."""
    hp = _vacuum_lattice(deck_dir)
    o, k = _converged_with_currents(hp, **TIGHT_CMFD)
    phi0 = o.flux()
    k1, _ = o.cmfd_step(k)
    assert abs(k1 - k) < 1e-9, (k1, k)
    ratio = o.flux() / phi0
    assert np.abs(ratio - 1.0).max() < 1e-8
    # the transport currents matter at this precision: the same solve
    # without them (D-hat = 0, plain diffusion) misses the transport k by
    # about 1e-4, five orders above the agreement with them
    o2 = Oracle(hp, n_inner=1, use_cmfd=True, **TIGHT_CMFD)
    o2.initialize()
    o2.set_flux(phi0)
    k2, _ = o2.cmfd_step(k)
    assert abs(k2 - k) > 1e-5, (k2, k)


def test_cmfd_negative_fixup(deck_dir):
    """<cmfd negative_fixup> clamps negative coarse fluxes to zero before the
    solve (cmfd.cpp:175-184): with one pin's flux made negative, the solves
    with and without the fixup differ, and with positive fluxes the option
    changes nothing."""
    hp = _vacuum_lattice(deck_dir)
    _, k = _converged_with_currents(hp)
    ks = {}
    for fixup in (0, 1):
        for neg in (False, True):
            o, _ = _converged_with_currents(hp, cmfd_negative_fixup=fixup)
            if neg:
                p = o.flux()
                off = hp.array("pin_fsr_off", 10, np.int64)
                p[off[4]:off[5], 0] *= -0.5  # the centre pin, fast group
                o.set_flux(p)
            ks[(fixup, neg)] = o.cmfd_step(k)[0]
    # (the oracle's thread-private tallies meet in an OpenMP critical
    # section, as the reference's do, so two runs can differ in the last bit)
    assert abs(ks[(0, False)] - ks[(1, False)]) < 1e-13, ks
    assert abs(ks[(0, True)] - ks[(1, True)]) > 1e-6, ks


def test_pin_flux_identity(pins3x3):
    o = Oracle(pins3x3, n_inner=2)
    o.initialize()
    o.fss_step()  # a non-flat flux
    for g in range(pins3x3.n_group):
        p1 = o.pin_flux(g)
        o.set_pin_flux(g, p1)
        p2 = o.pin_flux(g)
        np.testing.assert_allclose(p2, p1, rtol=4 * np.finfo(float).eps, atol=0)


def test_set_pin_flux_rejects_nonpositive(pins3x3):
    o = Oracle(pins3x3)
    o.initialize()
    p = o.pin_flux(0)
    p[3] = 0.0
    with pytest.raises(ValueError):
        o.set_pin_flux(0, p)


def _coarse_surf(nx, ny, cell, s):
    """Mesh::coarse_surf numbering for one plane (src/core/mesh.cpp:99-147):
    bottom faces, x-normal faces by rows, y-normal faces by columns, top."""
    ix, iy = cell % nx, cell // nx
    base = nx * ny
    return {0: base + (nx + 1) * iy + ix + 1,          # EAST
            2: base + (nx + 1) * iy + ix,              # WEST
            1: base + (nx + 1) * ny + (ny + 1) * ix + iy + 1,  # NORTH
            3: base + (nx + 1) * ny + (ny + 1) * ix + iy}[s]   # SOUTH


def _pin_balance(hp, n_inner=400):
    """Worst relative pin imbalance over groups after a converged inner
    iteration with the current tally on."""
    p = hp.p
    G, F = hp.n_group, hp.n_fsr
    nx, ny, hz = p.nx, p.ny, p.hz
    o = Oracle(hp, n_inner=n_inner)
    o.initialize()
    o.calc_fission_source(1.0)
    mat = hp.array("fsr_mat", F, np.int64)
    tr = hp.array("xs_tr", p.n_mat * G).reshape(-1, G)
    sc = hp.array("xs_sc", p.n_mat * G * G).reshape(-1, G, G)
    vol = hp.array("fsr_vol", F)
    off = hp.array("pin_fsr_off", nx * ny + 1, np.int64)
    cell = hp.array("pin_cell", nx * ny, np.int64)
    xv = hp.array("x_vec", nx + 1)
    yv = hp.array("y_vec", ny + 1)
    worst = 0.0
    for g in range(G):
        o.build_source(g)
        o.sweep(g, with_current=True)
        src, _ = o.source()
        phi = o.flux()[:, g]
        cur, _ = o.currents(g)
        rm = tr[mat, g] - sc[mat, g, g]
        for ip in range(nx * ny):
            c = int(cell[ip])
            r = slice(off[ip], off[ip + 1])
            gain = float((src[r] * vol[r]).sum())
            b = gain - float((phi[r] * vol[r] * rm[r]).sum())
            dx, dy = xv[c % nx + 1] - xv[c % nx], yv[c // nx + 1] - yv[c // nx]
            b -= cur[_coarse_surf(nx, ny, c, 0)] * dy * hz
            b -= cur[_coarse_surf(nx, ny, c, 1)] * dx * hz
            b += cur[_coarse_surf(nx, ny, c, 2)] * dy * hz
            b += cur[_coarse_surf(nx, ny, c, 3)] * dx * hz
            worst = max(worst, abs(b) / gain)
    return worst


def test_pin_balance_with_currents(pins3x3):
    """At inner convergence every pin balances: sources - removal = net
    outflow from the coarse-surface currents."""
    assert _pin_balance(pins3x3) < 1e-9


def _square_lattice(deck_dir, n, spacing, core=1):
    from mocc_amd import lattice_gen as lg
    d = lg.Deck(f"even_{n}x{core}_{int(round(spacing * 1000))}", 2, 16, spacing)
    d.lattices = {1: (n, n, lg.lattice_rows(n, 0, False))}
    d.core = [[1] * core for _ in range(core)]
    return HostProblem.load(lg.write_deck(d, deck_dir, 4))


def _walk_errors(hp):
    """Rays whose coarse data disagree with their segments: walking the
    surfaces of cm_data from the first cell (coarse_neighbor) must visit,
    for every pin crossing, the pin holding the crossing's segments, and use
    up exactly the ray's segments in both directions (moc_current_worker.hpp:
    220-260 walks them this way)."""
    p = hp.p
    nx, ny, F, R = p.nx, p.ny, p.n_fsr, p.n_ray
    seg_off = hp.array("ray_seg_off", R + 1, np.int64)
    seg_fsr = hp.array("seg_fsr", int(p.n_seg), np.int64)
    cm_cell = hp.array("ray_cm_cell", 2 * R, np.int64)
    cm_off = hp.array("ray_cm_off", R + 1, np.int64)
    cm = hp.array("cm_data", 4 * int(cm_off[-1]), np.int64).reshape(-1, 4)
    off = hp.array("pin_fsr_off", nx * ny + 1, np.int64)
    pin_cell = hp.array("pin_cell", nx * ny, np.int64)
    fsr_cell = np.repeat(pin_cell, np.diff(off))
    step = {0: 1, 1: nx, 2: -1, 3: -nx}
    bad = []
    for r in range(R):
        cells = fsr_cell[seg_fsr[seg_off[r]:seg_off[r + 1]]]
        n = len(cells)
        rows = cm[cm_off[r]:cm_off[r + 1]]
        ok = True
        c, i = int(cm_cell[2 * r]), 0
        for s, k in zip(rows[:, 0], rows[:, 2]):
            if s == 7:
                continue
            ok &= bool(np.all(cells[i:i + k] == c))
            i += k
            c += step.get(int(s), 0)
        ok &= i == n
        c, i = int(cm_cell[2 * r + 1]), n
        for s, k in zip(rows[:, 1], rows[:, 3]):
            if s == 7:
                continue
            ok &= bool(np.all(cells[max(0, i - k):i] == c))
            i -= k
            c += step.get(int(s), 0)
        ok &= i == 0
        if not ok:
            bad.append(r)
    return bad


# Even pin counts put pin corners on the ray grid: rays pass within rounding
# distance of a corner and cross its two lines at points a few ULPs apart.
EVEN = [(6, 0.05, 1), (8, 0.035, 1), (4, 0.02, 2)]


@pytest.mark.parametrize("n,spacing,core", EVEN)
def test_even_lattice_coarse_rays_and_balance(deck_dir, n, spacing, core):
    hp = _square_lattice(deck_dir, n, spacing, core)
    assert _walk_errors(hp) == []
    assert _pin_balance(hp, n_inner=200) < 1e-9


def test_even_lattice_cmfd_converges(deck_dir):
    """With balanced currents CMFD converges on an even lattice as on an odd
    one (tens of outer iterations, not a stall short of the tolerance)."""
    for n in (3, 6):
        hp = _square_lattice(deck_dir, n, 0.05)
        rc, st, _ = Oracle(hp, n_inner=5, use_cmfd=True, energy_mode=1).eigen_solve(1e-7, 1e-6, 40)
        assert rc == 0 and st.converged and st.iterations <= 12, (n, st.iterations)


# --- CMFD linear algebra --------------------------------------------------------

def _five_point(n, m, rng):
    N = n * m
    A = np.zeros((N, N))
    for i in range(N):
        x, y = i % n, i // n
        off = 0.0
        for dx, dy in [(1, 0), (-1, 0), (0, 1), (0, -1)]:
            if 0 <= x + dx < n and 0 <= y + dy < m:
                v = -rng.uniform(0.1, 1.0)
                A[i, i + dx + dy * n] = v
                off += abs(v)
        A[i, i] = off + rng.uniform(0.01, 0.5)
    return A


def test_bicgstab_vs_dense():
    rng = np.random.default_rng(3)
    A = _five_point(9, 7, rng)
    # a non-symmetric perturbation of the off-diagonals (D-hat terms)
    A += np.triu(A, 1) * rng.uniform(-0.3, 0.3, A.shape) * (A != 0)
    n = A.shape[0]
    rp = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        cols.extend(nz)
        vals.extend(A[i, nz])
        rp[i + 1] = len(cols)
    col = np.array(cols, np.int32)
    val = np.array(vals)
    b = rng.uniform(0.5, 1.5, n)
    x = np.zeros(n)
    it = np.zeros(1, np.int32)
    err = np.zeros(1)
    P = C.c_void_p
    lib().oracle_bicgstab(n, rp.ctypes.data_as(P), col.ctypes.data_as(P), val.ctypes.data_as(P),
                          b.ctypes.data_as(P), x.ctypes.data_as(P), 1e-12, 150,
                          it.ctypes.data_as(P), err.ctypes.data_as(P))
    xd = np.linalg.solve(A, b)
    assert err[0] <= 1e-12
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-10
    assert 0 < it[0] <= 150


def test_cmfd_matrix_structure(pins3x3):
    """The assembled CMFD matrix (cmfd.cpp:506-544) with D-hat = 0 on the
    first solve is symmetric with row sums V * Sigma_r plus the boundary
    couplings (zero for reflective faces)."""
    o = Oracle(pins3x3, n_inner=1, use_cmfd=True)
    o.initialize()
    o.fss_step()
    o.cmfd_step(1.0)
    p = pins3x3.p
    nc = p.nx * p.ny
    rp = np.zeros(nc + 1, np.int32)
    col = np.zeros(5 * nc, np.int32)
    val = np.zeros(5 * nc)
    P = C.c_void_p
    for g in range(pins3x3.n_group):
        nnz = lib().oracle_cmfd_matrix(o.h, g, rp.ctypes.data_as(P), col.ctypes.data_as(P),
                                       val.ctypes.data_as(P), 5 * nc)
        assert nnz == rp[-1] and nnz == nc + 2 * (2 * p.nx * p.ny - p.nx - p.ny)
        A = np.zeros((nc, nc))
        for i in range(nc):
            A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
        np.testing.assert_allclose(A, A.T, rtol=1e-13, atol=0)
        assert (np.diag(A) > 0).all()
        off = A - np.diag(np.diag(A))
        assert (off <= 0).all()


# --- 2-D/3-D correction worker and incoming-flux update -------------------------

def _ihm_flat(n_inner=1):
    hp = HostProblem.from_string(IHM_XML.format(n_inner=n_inner), DECKS)
    G = hp.n_group
    tr = hp.array("xs_tr", G)
    sc = hp.array("xs_sc", G * G).reshape(G, G)
    ch = hp.array("xs_ch", G)
    nf = hp.array("xs_nf", G)
    phi = np.linalg.solve(np.diag(tr) - sc, ch)
    return hp, phi, float(nf @ phi)


def _flat_state(o, hp, phi, k):
    o.initialize()
    o.set_flux(np.tile(phi, (hp.n_fsr, 1)))
    o.calc_fission_source(k)
    bc = o.boundary()
    o.set_boundary(np.repeat(phi / (4 * np.pi), bc.size // hp.n_group))


def test_corrections_flat_flux_closed_form():
    """cmdo::CurrentCorrections on the infinite medium at its exact solution
    (isotropic, flat angular flux q): every cell-average and face-average
    angular flux equals q, so alpha_x = alpha_y = 1/2 and beta = 1 for every
    cell, angle and group (correction_worker.cpp:93-153) -- the diamond-
    difference values CorrectionData starts from. The modular rays make the
    face areas exact: each pin face is crossed by a whole number of rays."""
    hp, phi, k = _ihm_flat()
    o = Oracle(hp, n_inner=1)
    _flat_state(o, hp, phi, k)
    for g in range(hp.n_group):
        o.build_source(g)
        o.sweep_2d3d(g)
    a, b = o.corrections()
    assert a.shape == (hp.n_group, 2 * hp.p.n_azi * hp.p.n_polar, hp.p.nx * hp.p.ny, 2)
    np.testing.assert_allclose(a, 0.5, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b, 1.0, rtol=0, atol=1e-13)
    # the flux was already the fixed point: the sweep left it in place
    np.testing.assert_allclose(o.flux(), np.tile(phi, (hp.n_fsr, 1)), rtol=1e-12, atol=0)


def test_corrections_heterogeneous(deck_dir):
    """On a 4x4 lattice with moderator-only pins and vacuum east/south faces:
    beta is exactly 1 in the single-material cells (the flux-weighted
    Sigma_t of a uniform cell is its Sigma_t, which is also its homogenized
    value) and differs from 1 in fuel cells; the coarse currents of the 2D3D
    sweep equal those of the plain sweep with currents from the same state
    (the same angular fluxes at the same crossings); a second sweep moves the
    factors less than the first (residuals shrink)."""
    from mocc_amd import lattice_gen as lg
    d = lg.pin_lattice(4, 2, 8, 0.05, "corr4x4", bc={"east": "vacuum", "south": "vacuum"})
    rows = d.lattices[1][2]
    rows[1][2] = rows[2][1] = 2  # moderator-only pins
    hp = HostProblem.load(lg.write_deck(d, deck_dir, 7))
    G = hp.n_group
    o = Oracle(hp, n_inner=3)
    o.initialize()
    o.calc_fission_source(1.0)
    ref = Oracle(hp, n_inner=3)
    ref.initialize()
    ref.calc_fission_source(1.0)
    for g in range(G):
        o.build_source(g)
        o.sweep_2d3d(g)
        ref.build_source(g)
        ref.sweep(g, with_current=True)
    for g in range(G):
        c2, _ = o.currents(g)
        c1, _ = ref.currents(g)
        np.testing.assert_allclose(c2, c1, rtol=1e-12, atol=1e-14 * np.abs(c1).max())
    a, b = o.corrections()
    mat = hp.array("fsr_mat", hp.n_fsr, np.int64)
    off = hp.array("pin_fsr_off", 17, np.int64)
    cell = hp.array("pin_cell", 16, np.int64)
    uniform = [int(cell[ip]) for ip in range(16) if len(set(mat[off[ip]:off[ip + 1]])) == 1]
    mixed = [c for c in range(16) if c not in uniform]
    assert len(uniform) == 2
    np.testing.assert_allclose(b[:, :, uniform], 1.0, rtol=0, atol=1e-12)
    assert np.abs(b[:, :, mixed] - 1.0).max() > 1e-3
    # alpha grows past 1 in cells on the vacuum faces, where the entering face flux is zero
    assert np.all(np.isfinite(a)) and a.min() > 0.0
    r1 = [o.correction_residual(g) for g in range(G)]
    for g in range(G):
        o.build_source(g)
        o.sweep_2d3d(g)
    r2 = [o.correction_residual(g) for g in range(G)]
    assert all((r2[g] < r1[g]).all() for g in range(G)), (r1, r2)


def test_update_incoming_flux_flat(deck_dir):
    """MoCSweeper::update_incoming_flux (moc_sweeper.cpp:255-291) after CMFD
    solves on the infinite medium at its exact solution: the CMFD surface
    flux is the scalar flux and the net current zero, so the incoming flux
    set from the first solve's partial currents, 2/(4 pi) (J+ + J-) =
    phi/(4 pi), and the ratio of a later solve's to the earlier one's (1)
    both leave the exact angular flux in place."""
    hp, phi, k = _ihm_flat()
    o = Oracle(hp, n_inner=1, use_cmfd=True, cmfd_k_tol=1e-12, cmfd_psi_tol=1e-12,
               cmfd_resid_reduction=1e-10, cmfd_max_iter=200)
    _flat_state(o, hp, phi, k)
    for g in range(hp.n_group):
        o.build_source(g)
        o.sweep(g, with_current=True)
    bc0 = o.boundary()
    for _ in range(2):
        o.cmfd_step(k)
        o.update_incoming_flux()
        np.testing.assert_allclose(o.boundary(), bc0, rtol=1e-10, atol=0)

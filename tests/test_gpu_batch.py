"""Batch engine (mocgpu_batch_*): many problems advanced together by one
captured outer-iteration graph.

* each problem converges to the same eigenvalue as its one-at-a-time device
  solve (same Jacobi energy sweeps; only the reduction trees differ);
* a problem's result is bitwise the same alone and in a mix (the serving
  contract);
* graph replays, with eager device work and synchronizations between them,
  give bitwise the result of eager launches after every replay;
* admission respects the device-memory budget;
* a handle created on a shared geometry solves bitwise like an independent one.
"""
import dataclasses

import numpy as np
import pytest

from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

K_TOL, PSI_TOL, MAX_IT = 1e-7, 1e-6, 3000


@pytest.fixture(scope="module")
def batch_decks(deck_dir):
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    decks = {
        "pin1": lg.write_deck(lg.pin_lattice(1, 2, 16, 0.05, "b_pin1_2g"), deck_dir, 21),
        "pins3x3": lg.write_config("pins3x3", deck_dir),
        "pins3x3_pert": lg.write_deck(dataclasses.replace(lg.pins3x3_deck(), name="b_pins3x3_pert"), deck_dir, 1,
                                rel_perturb=0.1),
        "asm17": lg.write_deck(lg.pin_lattice(17, 2, 4, 0.1, "b_asm17_2g", guide_tubes=True), deck_dir, 2),
    }
    return {k: HostProblem.load(v) for k, v in decks.items()}


def _sweepers(batch_decks, names):
    from mocc_amd.sweeper import GpuMoCSweeper
    return {n: GpuMoCSweeper(batch_decks[n], energy_mode=1) for n in names}


def _solve_batch(sweepers, order, use_graph=True):
    from mocc_amd.sweeper import GpuBatch
    b = GpuBatch(use_graph=use_graph)
    ids = {}
    for n in order:
        ids[b.add(sweepers[n], K_TOL, PSI_TOL, MAX_IT)] = n
    res = b.solve_all(chunk=4)
    out = {ids[cid]: st for cid, st in res.items()}
    b.close()
    return out


def test_batch_matches_single(batch_decks):
    names = ["pin1", "pins3x3", "pins3x3_pert", "asm17"]
    sw = _sweepers(batch_decks, names)
    res = _solve_batch(sw, names)
    assert set(res) == set(names)
    for n in names:
        print(f"case {n}", flush=True)
        single = _sweepers(batch_decks, [n])[n]
        rc, st1 = single.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
        assert rc == 0 and st1.converged
        st = res[n]
        assert st.converged, (n, st.iterations)
        assert abs(st.k - st1.k) < 1e-9, (n, st.k, st1.k)
        assert abs(st.iterations - st1.iterations) <= 1
        # the two solves may stop one outer iteration apart (psi_tol 1e-6)
        f_b, f_s = sw[n].flux(), single.flux()
        assert np.abs(f_b / f_s - 1.0).max() < 1e-5


def test_alone_vs_mixed_bitwise(batch_decks):
    target = "pins3x3_pert"
    alone = _sweepers(batch_decks, [target])
    r_alone = _solve_batch(alone, [target])[target]
    f_alone = alone[target].flux()
    for order in (["asm17", "pin1", target, "pins3x3"], [target, "pins3x3", "asm17"]):
        sw = _sweepers(batch_decks, order)
        r = _solve_batch(sw, order)[target]
        assert r.k == r_alone.k and r.iterations == r_alone.iterations
        assert r.err_psi == r_alone.err_psi
        assert np.array_equal(sw[target].flux(), f_alone)


def test_graph_replays_match_eager(batch_decks):
    from mocc_amd.sweeper import GpuBatch
    names = ["pin1", "pins3x3", "asm17"]
    sg, se = _sweepers(batch_decks, names), _sweepers(batch_decks, names)
    bg, be = GpuBatch(use_graph=True), GpuBatch(use_graph=False)
    for n in names:
        bg.add(sg[n], K_TOL, PSI_TOL, MAX_IT)
        be.add(se[n], K_TOL, PSI_TOL, MAX_IT)
    for rep in range(12):
        bg.run(1)
        be.run(1)
        bg.poll()
        be.poll()
        for n in names:  # eager reads and synchronizations between replays
            assert np.array_equal(sg[n].flux(), se[n].flux()), (rep, n)
            assert sg[n].total_fission() == se[n].total_fission()


def test_batch_set_matches_one_batch(batch_decks):
    """Problems spread over two batches on two streams (GpuBatchSet) give
    bitwise the k, iteration counts and fluxes of one batch."""
    from mocc_amd.sweeper import CapacityError, GpuBatchSet
    names = ["pin1", "pins3x3", "pins3x3_pert", "asm17"]
    one = _sweepers(batch_decks, names)
    r_one = _solve_batch(one, names)
    two = _sweepers(batch_decks, names)
    bs = GpuBatchSet(2)
    ids = {bs.add(two[n], K_TOL, PSI_TOL, MAX_IT): n for n in names}
    assert [b.size() for b in bs.batches] == [2, 2]
    res = bs.solve_all(chunk=4)
    assert set(res) == set(ids)
    for sid, n in ids.items():
        print(f"batch set case {n}", flush=True)
        assert res[sid].k == r_one[n].k and res[sid].iterations == r_one[n].iterations, n
        assert np.array_equal(two[n].flux(), one[n].flux()), n
    bs.close()
    need = two["pins3x3"].state_bytes()
    small = GpuBatchSet(2, budget_bytes=need + need // 2)
    small.add(two["pins3x3"], K_TOL, PSI_TOL, MAX_IT)
    with pytest.raises(CapacityError):
        small.add(two["pins3x3_pert"], K_TOL, PSI_TOL, MAX_IT)
    assert small.size() == 1
    small.close()


def test_admission_budget(batch_decks):
    from mocc_amd.sweeper import CapacityError, GpuBatch
    sw = _sweepers(batch_decks, ["pins3x3", "pins3x3_pert"])
    need = sw["pins3x3"].state_bytes()
    assert need > 0 and sw["pins3x3_pert"].state_bytes() == need
    b = GpuBatch(budget_bytes=need + need // 2)
    i1 = b.add(sw["pins3x3"], K_TOL, PSI_TOL, MAX_IT)
    with pytest.raises(CapacityError):
        b.add(sw["pins3x3_pert"], K_TOL, PSI_TOL, MAX_IT)
    assert b.size() == 1 and b.bytes_used() == need
    b.remove(i1)
    b.add(sw["pins3x3_pert"], K_TOL, PSI_TOL, MAX_IT)
    res = b.solve_all()
    assert len(res) == 1 and list(res.values())[0].converged


def test_create_like_shares_geometry(batch_decks):
    from mocc_amd.sweeper import DeviceError, GpuMoCSweeper
    a = GpuMoCSweeper(batch_decks["pins3x3"], energy_mode=1)
    b = GpuMoCSweeper(batch_decks["pins3x3_pert"], energy_mode=1, like=a)
    c = GpuMoCSweeper(batch_decks["pins3x3_pert"], energy_mode=1)
    print("geometry bytes: shared", b.geometry_bytes(), "own", c.geometry_bytes(), "state", b.state_bytes(),
          flush=True)
    assert b.geometry_bytes() > 0 and c.geometry_bytes() > 0
    rb, sb = b.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
    rc, sc = c.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
    assert rb == rc == 0 and sb.k == sc.k
    assert np.array_equal(b.flux(), c.flux())
    with pytest.raises(DeviceError):
        GpuMoCSweeper(batch_decks["pin1"], energy_mode=1, like=a)


def _cmfd_sweepers(batch_decks, names):
    from mocc_amd.sweeper import GpuMoCSweeper
    return {n: GpuMoCSweeper(batch_decks[n], energy_mode=1, n_inner=5, use_cmfd=True) for n in names}


def test_batch_cmfd_matches_single(batch_decks):
    """Problems with CMFD (5 inner iterations, as mocc's CMFD decks use):
    the batch runs each one's CMFD step (one solve workgroup per problem)
    and its last inner sweep tallies currents; each converges like its
    one-at-a-time device solve."""
    names = ["pins3x3", "pins3x3_pert", "asm17"]
    sw = _cmfd_sweepers(batch_decks, names)
    res = _solve_batch(sw, names)
    for n in names:
        print(f"cmfd case {n}", flush=True)
        single = _cmfd_sweepers(batch_decks, [n])[n]
        rc, st1 = single.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
        assert rc == 0 and st1.converged
        st = res[n]
        assert st.converged, (n, st.iterations)
        assert st.iterations < 40, st.iterations  # CMFD: tens of outer iterations, not hundreds
        assert abs(st.k - st1.k) < 1e-9, (n, st.k, st1.k)
        assert abs(st.iterations - st1.iterations) <= 1
        f_b, f_s = sw[n].flux(), single.flux()
        assert np.abs(f_b / f_s - 1.0).max() < 1e-5


def test_batch_cmfd_alone_vs_mixed_bitwise(batch_decks):
    """A CMFD problem mixed with problems without CMFD: bitwise the same."""
    from mocc_amd.sweeper import GpuMoCSweeper
    target = "pins3x3_pert"
    alone = _cmfd_sweepers(batch_decks, [target])
    r_alone = _solve_batch(alone, [target])[target]
    f_alone = alone[target].flux()
    sw = _sweepers(batch_decks, ["asm17", "pin1"])
    sw[target] = GpuMoCSweeper(batch_decks[target], energy_mode=1, n_inner=5, use_cmfd=True)
    r = _solve_batch(sw, ["asm17", target, "pin1"])[target]
    assert r.k == r_alone.k and r.iterations == r_alone.iterations and r.err_psi == r_alone.err_psi
    assert np.array_equal(sw[target].flux(), f_alone)


def test_batch_cmfd_graph_matches_eager(batch_decks):
    from mocc_amd.sweeper import GpuBatch
    names = ["pins3x3", "asm17"]
    sg, se = _cmfd_sweepers(batch_decks, names), _cmfd_sweepers(batch_decks, names)
    bg, be = GpuBatch(use_graph=True), GpuBatch(use_graph=False)
    for n in names:
        bg.add(sg[n], K_TOL, PSI_TOL, MAX_IT)
        be.add(se[n], K_TOL, PSI_TOL, MAX_IT)
    for rep in range(6):
        bg.run(1)
        be.run(1)
        bg.poll()
        be.poll()
        for n in names:
            assert np.array_equal(sg[n].flux(), se[n].flux()), (rep, n)


@pytest.mark.parametrize("big", ["lat48", "core51"])
def test_batch_cmfd_pins_beside_wide_mesh_bitwise(cmfd_decks, big):
    """A 3x3-pin problem beside the 48x48 lattice or the 51x51 core (the wide
    coarse solve at 5 or 6 cells per lane for the large mesh, the 512-lane
    solve for the 3x3 one, in one launch): the small problem's k, iterations
    and flux are bitwise those of its batch alone."""
    from mocc_amd.sweeper import GpuMoCSweeper
    mk = lambda n: GpuMoCSweeper(cmfd_decks[n], energy_mode=1, n_inner=5, use_cmfd=True)
    print("cmfd 3x3 pins alone", flush=True)
    alone = {"pins3x3": mk("pins3x3")}
    r_alone = _solve_batch(alone, ["pins3x3"])["pins3x3"]
    print(f"cmfd 3x3 pins beside {big}", flush=True)
    mixed = {big: mk(big), "pins3x3": mk("pins3x3")}
    r = _solve_batch(mixed, [big, "pins3x3"])
    assert r["pins3x3"].converged and r[big].converged, (r["pins3x3"].iterations, r[big].iterations)
    assert r["pins3x3"].k == r_alone.k and r["pins3x3"].iterations == r_alone.iterations
    assert np.array_equal(mixed["pins3x3"].flux(), alone["pins3x3"].flux())


def test_batch_plain_problem_beside_cmfd_bitwise(batch_decks):
    """A problem without CMFD sweeps its last inner iteration in the
    current-tallying queue kernel when a CMFD problem shares the batch, and
    in the plain one alone: bitwise the same either way."""
    from mocc_amd.sweeper import GpuMoCSweeper
    target = "asm17"
    alone = _sweepers(batch_decks, [target])
    alone[target] = GpuMoCSweeper(batch_decks[target], energy_mode=1, n_inner=5)
    print("17x17 without CMFD alone", flush=True)
    r_alone = _solve_batch(alone, [target])[target]
    mixed = {target: GpuMoCSweeper(batch_decks[target], energy_mode=1, n_inner=5),
             "pins3x3": GpuMoCSweeper(batch_decks["pins3x3"], energy_mode=1, n_inner=5, use_cmfd=True)}
    print("17x17 without CMFD beside a CMFD problem", flush=True)
    r = _solve_batch(mixed, ["pins3x3", target])[target]
    assert r.k == r_alone.k and r.iterations == r_alone.iterations
    assert np.array_equal(mixed[target].flux(), alone[target].flux())


def test_batch_rejects_gauss_seidel_energy(batch_decks):
    """A batch sweeps Jacobi in energy: a handle in mocc's Gauss-Seidel energy
    mode (energy_mode 0) would follow another iteration path than its own
    solve, so batch_add refuses it (MOCGPU_ERR_UNSUPPORTED) and the batch is
    unchanged."""
    from mocc_amd.sweeper import DeviceError, GpuBatch, GpuMoCSweeper
    gs = GpuMoCSweeper(batch_decks["pins3x3"], energy_mode=0)
    b = GpuBatch()
    with pytest.raises(DeviceError, match="-6"):
        b.add(gs, K_TOL, PSI_TOL, MAX_IT)
    assert b.size() == 0 and b.bytes_used() == 0
    ok = _sweepers(batch_decks, ["pins3x3"])["pins3x3"]
    b.add(ok, K_TOL, PSI_TOL, MAX_IT)
    assert b.size() == 1
    b.close()


def test_member_handle_is_owned_by_the_batch(batch_decks):
    """From mocgpu_batch_add to mocgpu_batch_remove the batch's launches own a
    handle's device state: single-problem calls that change it are refused,
    reads wait for the batch, a second batch cannot take it, and destroying
    it takes it out of the batch first, so the others still solve as alone."""
    from mocc_amd.sweeper import DeviceError, GpuBatch
    sw = _sweepers(batch_decks, ["pin1", "pins3x3"])
    b = GpuBatch()
    i1 = b.add(sw["pin1"], K_TOL, PSI_TOL, MAX_IT)
    b.add(sw["pins3x3"], K_TOL, PSI_TOL, MAX_IT)
    with pytest.raises(DeviceError, match="member of a batch"):
        sw["pin1"].initialize()
    with pytest.raises(DeviceError, match="member of a batch"):
        sw["pin1"].eigen_solve(K_TOL, PSI_TOL, MAX_IT)
    b.run(2)  # no poll: the read below waits for the batch's queued work
    assert np.all(np.isfinite(sw["pin1"].flux())) and sw["pin1"].total_fission() > 0
    other = GpuBatch()
    with pytest.raises(DeviceError, match="another batch"):
        other.add(sw["pin1"], K_TOL, PSI_TOL, MAX_IT)
    assert other.size() == 0
    other.close()
    print("destroy a member handle, then solve the rest", flush=True)
    sw["pins3x3"].close()
    assert b.size() == 1
    res = b.solve_all(chunk=4)
    (cid, st), = list(res.items())
    assert cid == i1 and st.converged
    b.remove(i1)
    rc, st1 = sw["pin1"].eigen_solve(K_TOL, PSI_TOL, MAX_IT)
    assert rc == 0 and abs(st.k - st1.k) < 1e-9
    b.close()


@pytest.mark.parametrize("big", ["lat24", "lat34", "lat40", "lat55"])
def test_batch_cmfd_mixed_mesh_sizes_bitwise(cmfd_decks, big):
    """A 3x3-pin problem beside a larger mesh: 24x24 (576 cells, the 512-lane
    solve at 2 cells per lane for both) or 34x34, 40x40, 55x55 (1156, 1600,
    3025 cells: the wide solve at 3, 4 and 6 cells per lane for the large
    mesh, the 512-lane solve for the 3x3 one, in one launch). Each problem's
    k, iterations and flux are bitwise those of its batch alone, and the large
    mesh's k agrees with its single-problem solve."""
    from mocc_amd.sweeper import GpuMoCSweeper
    mk = lambda n: GpuMoCSweeper(cmfd_decks[n], energy_mode=1, n_inner=5, use_cmfd=True)
    alone = {"pins3x3": mk("pins3x3"), big: mk(big)}
    r_alone = {n: _solve_batch(alone, [n])[n] for n in alone}
    print(f"cmfd 3x3 pins beside {big}", flush=True)
    mixed = {big: mk(big), "pins3x3": mk("pins3x3")}
    r = _solve_batch(mixed, [big, "pins3x3"])
    for n in mixed:
        assert r[n].converged, (n, r[n].iterations)
        assert r[n].k == r_alone[n].k and r[n].iterations == r_alone[n].iterations, n
        assert np.array_equal(mixed[n].flux(), alone[n].flux()), n
    single = mk(big)
    rc, st = single.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
    assert rc == 0 and abs(st.k - r[big].k) < 1e-9 and abs(st.iterations - r[big].iterations) <= 1


def test_branch_batch_vs_oracle(tmp_path):
    """Sixteen cross-section branch cases of the batched workload in one
    batch (4 single pins, 8 3x3 lattices, 4 17x17 assemblies with guide
    tubes; 2 and 7 groups; CY 16x3 at 0.02, 0.035 and 0.05 cm, the 17x17
    cases at full density; cross sections x U[0.9, 1.1]; CMFD with 5 inner
    iterations): every case's keff within 1e-5 of the FP64 oracle run in the
    same iteration scheme (Jacobi energy sweeps, Gauss-Seidel boundary
    update, CMFD)."""
    from mocc_amd.branch import BranchFactory, c4_cases
    from mocc_amd.sweeper import GpuBatch
    cases = c4_cases(16)
    big = [i for i, c in enumerate(cases) if c.kind == "17x17"]
    for i, sp in zip(big, (0.02, 0.02, 0.035, 0.05)):
        cases[i] = dataclasses.replace(cases[i], spacing=sp)
    assert {cases[i].groups for i in big} == {2, 7}
    factory = BranchFactory(str(tmp_path), energy_mode=1, n_inner=5, use_cmfd=True)
    handles = [factory.make(c) for c in cases]
    b = GpuBatch()
    ids = [b.add(h, K_TOL, PSI_TOL, MAX_IT) for h in handles]
    res = b.solve_all(chunk=4)
    b.close()
    worst = 0.0
    for i, c in enumerate(cases):
        st = res[ids[i]]
        assert st.converged, (i, c, st.iterations)
        hp = factory.host_problem(c)
        o = Oracle(hp, n_inner=5, use_cmfd=True, energy_mode=1, gs_boundary=True)
        rc, so, _ = o.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
        assert rc == 0
        dk = abs(st.k - so.k)
        worst = max(worst, dk)
        print(f"case {i} {c.kind} {c.groups}g {c.spacing} cm: k {st.k:.8f} oracle {so.k:.8f} |dk| {dk:.1e} "
              f"outer {st.iterations}/{so.iterations}", flush=True)
        assert dk < 1e-5, (i, c, st.k, so.k)
    for h in handles:
        h.close()
    factory.close()
    print(f"worst |dk| over 16 branch cases: {worst:.2e}", flush=True)


def test_batch_one_wave_queue_bitwise(batch_decks):
    """A batch whose problems all run bands of at most 32 rays sweeps with
    one-wave queue workgroups; a problem's result is bitwise the same as in a
    batch that also holds 64-ray bands (two-wave workgroups)."""
    from mocc_amd.sweeper import GpuMoCSweeper
    mk = lambda n, br: GpuMoCSweeper(batch_decks[n], energy_mode=1, band_rays=br)
    narrow = {"pins3x3": mk("pins3x3", 32)}
    assert narrow["pins3x3"].options().band_rays == 32
    print("3x3 pins, 32-ray bands, one-wave queue", flush=True)
    r1 = _solve_batch(narrow, ["pins3x3"])["pins3x3"]
    mixed = {"pins3x3": mk("pins3x3", 32), "asm17": mk("asm17", 64)}
    print("3x3 pins, 32-ray bands, beside 64-ray bands", flush=True)
    r2 = _solve_batch(mixed, ["asm17", "pins3x3"])["pins3x3"]
    assert r1.converged and r1.k == r2.k and r1.iterations == r2.iterations
    assert np.array_equal(narrow["pins3x3"].flux(), mixed["pins3x3"].flux())


def test_autotuned_branch_batch(tmp_path):
    """The batch-level autotuner (BranchFactory.tune: every band width x
    lane map timed on a batch of sample cases, the fastest kept for all
    geometries) picks a valid layout, and the cases solve in one batch to the
    eigenvalues of the same cases at the default layout (within the rounding
    of the fixed-point tally scale, which follows the band layout)."""
    from mocc_amd.branch import BranchFactory, c4_cases
    from mocc_amd.sweeper import GpuBatch
    cases = c4_cases(8)
    res = {}
    for tune in (True, False):
        f = BranchFactory(str(tmp_path / f"t{int(tune)}"), energy_mode=1, n_inner=5, use_cmfd=True)
        if tune:
            w, m = f.tune(cases[:4], n_outer=2)
            print("tuned (band rays, lane map):", (w, m), "candidates:", f.tuning, flush=True)
            assert w in (32, 48, 64) and m in (0, 1, 2) and len(f.tuning) == 9
            assert (w, m) == min(f.tuning, key=lambda r: r[2])[:2]
        hs = [f.make(c) for c in cases]
        if tune:
            assert all(v == (w, m) for v in f.tuned().values())
        b = GpuBatch()
        ids = [b.add(h, K_TOL, PSI_TOL, MAX_IT) for h in hs]
        out = b.solve_all(chunk=4)
        b.close()
        res[tune] = [out[i] for i in ids]
        for h in hs:
            h.close()
        f.close()
    for a, b in zip(res[True], res[False]):
        assert a.converged and b.converged
        assert abs(a.k - b.k) < 1e-9, (a.k, b.k)


def test_batch_rejects_oversized_cmfd_mesh(cmfd_decks):
    """A single problem's CMFD takes up to 6144 coarse cells; a batched one up
    to 3072 (the batch's 512-lane solve): adding a 4096-cell problem fails
    with a message and leaves the batch usable."""
    from mocc_amd.sweeper import DeviceError, GpuBatch, GpuMoCSweeper
    big = GpuMoCSweeper(cmfd_decks["lat64"], energy_mode=1, n_inner=1, use_cmfd=True)
    b = GpuBatch()
    with pytest.raises(DeviceError, match="at most 3072 coarse cells"):
        b.add(big, 1e-6, 1e-5, 10)
    assert b.size() == 0

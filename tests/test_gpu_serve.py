"""Serving loop (continuous batching over the batch engine): every request
is answered, and each answer is bitwise the answer of the same request solved
alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def factory(tmp_path_factory):
    from mocc_amd.branch import BranchFactory
    f = BranchFactory(str(tmp_path_factory.mktemp("branch")))
    yield f
    f.close()


def _requests(n, seed):
    from mocc_amd.branch import Case
    rng = np.random.default_rng(seed)
    kinds = ["pin", "3x3", "17x17"]
    cases = [Case(kinds[int(rng.integers(0, 3))], int(rng.choice([2, 7])), 0.1, 900 + i) for i in range(n)]
    return cases


def _solve_alone(factory, case, k_tol=1e-7, psi_tol=1e-6):
    from mocc_amd.sweeper import GpuBatch
    s = factory.make(case)
    b = GpuBatch()
    b.add(s, k_tol, psi_tol, 3000)
    (cid, st), = list(b.solve_all().items())
    flux = s.flux()
    b.close()
    s.close()
    return st, flux


def test_serve_matches_alone(factory):
    from mocc_amd.serve import Server, poisson_arrivals, summarize
    cases = _requests(16, 3)
    arrivals = poisson_arrivals(len(cases), rate=200.0, seed=3)
    srv = Server(factory, chunk=2, keep_flux=True)
    recs = srv.serve(arrivals, cases)
    srv.close()
    summ = summarize(recs)
    print(summ, flush=True)
    assert summ["converged"] == len(cases)
    for r in recs[::5]:
        print(f"request {r.index} {r.case}", flush=True)
        st, flux = _solve_alone(factory, r.case)
        assert st.k == r.k and st.iterations == r.iterations
        assert np.array_equal(flux, r.flux)


def test_serve_under_memory_budget(factory):
    """A budget that holds about two problems at a time still serves all."""
    from mocc_amd.serve import Server, poisson_arrivals
    cases = _requests(6, 5)
    probe = factory.make(cases[0])
    budget = 2 * max(probe.state_bytes(), 1)
    probe.close()
    # every request fits alone: use cases no larger than the probe's
    cases = [c for c in cases if c.kind == cases[0].kind and c.groups <= cases[0].groups] or cases[:1]
    srv = Server(factory, budget_bytes=budget, chunk=2)
    recs = srv.serve(poisson_arrivals(len(cases), rate=1000.0, seed=5), cases)
    srv.close()
    assert all(r.converged for r in recs)


@pytest.fixture(scope="module")
def cmfd_factory(tmp_path_factory):
    from mocc_amd.branch import BranchFactory
    f = BranchFactory(str(tmp_path_factory.mktemp("branch_cmfd")), n_inner=5, use_cmfd=True)
    yield f
    f.close()


def test_serve_cmfd_matches_alone(cmfd_factory):
    """Serving with CMFD (5 inner iterations), including 2x2 assemblies of
    17x17 pins: an even pin count, whose rays graze pin corners. Every request
    converges in tens of outer iterations, and a request's answer is bitwise
    the answer of the same request solved alone."""
    from mocc_amd.branch import Case
    from mocc_amd.serve import Server, poisson_arrivals, summarize
    kinds = ["2x2x17x17", "pin", "3x3", "17x17", "2x2x17x17", "3x3"]
    cases = [Case(k, 2 if i % 2 == 0 else 7, 0.1, 700 + i) for i, k in enumerate(kinds)]
    srv = Server(cmfd_factory, chunk=2, keep_flux=True)
    recs = srv.serve(poisson_arrivals(len(cases), rate=100.0, seed=7), cases)
    srv.close()
    print(summarize(recs), flush=True)
    assert all(r.converged and r.iterations < 40 for r in recs), [(r.case.kind, r.iterations) for r in recs]
    for r in recs[::2]:
        print(f"request {r.index} {r.case}", flush=True)
        st, flux = _solve_alone(cmfd_factory, r.case)
        assert st.k == r.k and st.iterations == r.iterations
        assert np.array_equal(flux, r.flux)


def test_recycled_handle_matches_fresh(tmp_path):
    """A finished case's handle given the next case of its geometry
    (mocgpu_reload_xs through BranchFactory.release/make) answers that case
    bitwise as a newly created handle does, in a batch and alone, with and
    without CMFD."""
    from mocc_amd.branch import BranchFactory, Case
    from mocc_amd.sweeper import GpuBatch
    for use_cmfd in (False, True):
        ni = 5 if use_cmfd else 1
        f_old = BranchFactory(str(tmp_path / f"r{int(use_cmfd)}"), n_inner=ni, use_cmfd=use_cmfd)
        f_new = BranchFactory(str(tmp_path / f"n{int(use_cmfd)}"), n_inner=ni, use_cmfd=use_cmfd)
        first, second = Case("3x3", 7, 0.1, 71), Case("3x3", 7, 0.1, 72)
        s = f_old.make(first)
        rc, st = s.eigen_solve(1e-7, 1e-6, 3000)
        assert rc == 0 and st.converged
        f_old.release(first, s)
        r = f_old.make(second)
        assert r is s, "the released handle is reused"
        fresh = f_new.make(second)
        print(f"cmfd={use_cmfd}: recycled vs fresh, alone", flush=True)
        (rc1, st1), (rc2, st2) = r.eigen_solve(1e-7, 1e-6, 3000), fresh.eigen_solve(1e-7, 1e-6, 3000)
        assert rc1 == rc2 == 0 and st1.k == st2.k and st1.iterations == st2.iterations
        assert np.array_equal(r.flux(), fresh.flux())
        print(f"cmfd={use_cmfd}: recycled vs fresh, in a batch", flush=True)
        f_old.release(second, r)
        r = f_old.make(second)
        out = []
        for h in (r, fresh):
            b = GpuBatch()
            b.add(h, 1e-7, 1e-6, 3000)
            (cid, st), = list(b.solve_all().items())
            out.append((st.k, st.iterations, h.flux()))
            b.close()
        assert out[0][0] == out[1][0] and out[0][1] == out[1][1] and np.array_equal(out[0][2], out[1][2])
        f_old.release(second, r)
        fresh.close()
        f_old.close()
        f_new.close()


def test_reload_xs_refuses_other_geometry(tmp_path):
    """mocgpu_reload_xs takes only a problem of the handle's geometry and
    group count: another lattice or group count is refused with a message
    and leaves the handle usable."""
    from mocc_amd.branch import BranchFactory, Case
    from mocc_amd.sweeper import DeviceError
    f = BranchFactory(str(tmp_path))
    s = f.make(Case("3x3", 2, 0.1, 81))
    for other in (Case("pin", 2, 0.1, 82), Case("3x3", 7, 0.1, 83)):
        with pytest.raises(DeviceError, match="geometry or group count differs"):
            s.reload_xs(f.host_problem(other))
    s.reload_xs(f.host_problem(Case("3x3", 2, 0.1, 84)))
    rc, st = s.eigen_solve(1e-7, 1e-6, 3000)
    assert rc == 0 and st.converged
    s.close()
    f.close()

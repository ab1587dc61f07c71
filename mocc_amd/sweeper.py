"""Python mirror of mocc's TransportSweeper / MoCSweeper interface over the
gfx950 library (include/mocgpu.h).

Method names and argument meaning follow src/core/transport_sweeper.hpp and
src/sweepers/moc/moc_sweeper.hpp; errors the reference raises as exceptions
come back as status codes and are re-raised here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .native import EigenState, HostProblem, MocgpuOptions, gpu, ptr


class DeviceError(RuntimeError):
    pass


class NonPositiveFluxError(ValueError):
    pass


class GpuMoCSweeper:
    def __init__(self, prob: HostProblem, n_inner: int | None = None, energy_mode: int = 0,
                 use_cmfd: bool | None = None, band_rays: int = 0, like: "GpuMoCSweeper | None" = None,
                 gs_boundary: bool | None = None, lane_map: int | None = None, **cmfd):
        """Options default to the deck's. like: another sweeper on the same
        geometry whose device ray plan is shared (mocgpu_create_like);
        gs_boundary: Gauss-Seidel (mocc's default) or Jacobi boundary update."""
        L = gpu()
        opt = prob.options()
        if n_inner is not None:
            opt.n_inner = n_inner
        if gs_boundary is not None:
            opt.gs_boundary = 1 if gs_boundary else 0
        if lane_map is not None:
            opt.lane_map = lane_map
        opt.energy_mode = energy_mode
        opt.band_rays = band_rays
        if use_cmfd is not None:
            opt.use_cmfd = 1 if use_cmfd else 0
        for k, v in cmfd.items():
            setattr(opt, k, v)
        self.prob = prob
        self.opt = opt
        self.h = C.c_void_p()
        if like is not None:
            rc = L.mocgpu_create_like(like.h, prob.view, C.byref(opt), C.byref(self.h))
        else:
            rc = L.mocgpu_create(prob.view, C.byref(opt), C.byref(self.h))
        if rc != 0:
            raise DeviceError(L.mocgpu_last_error(None).decode())
        self.F, self.G = prob.n_fsr, prob.n_group
        self.npin = prob.p.nx * prob.p.ny

    def _check(self, rc):
        if rc != 0:
            msg = gpu().mocgpu_last_error(self.h).decode()
            if rc == -4:
                raise NonPositiveFluxError(msg)
            raise DeviceError(f"mocgpu error {rc}: {msg}")

    def close(self):
        if self.h:
            gpu().mocgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reload_xs(self, prob: HostProblem):
        """Take prob's cross sections (same geometry and group count) and
        start over from a new handle's state, reusing the device buffers
        (mocgpu_reload_xs)."""
        self._check(gpu().mocgpu_reload_xs(self.h, prob.view))

    # --- TransportSweeper -------------------------------------------------
    def initialize(self):
        self._check(gpu().mocgpu_initialize(self.h))

    def flux(self) -> np.ndarray:
        out = np.zeros((self.F, self.G))
        self._check(gpu().mocgpu_get_flux(self.h, ptr(out)))
        return out

    def set_flux(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.shape == (self.F, self.G)
        self._check(gpu().mocgpu_set_flux(self.h, ptr(phi)))

    def boundary(self) -> np.ndarray:
        n = gpu().mocgpu_boundary_size(self.h)
        out = np.zeros(n)
        self._check(gpu().mocgpu_get_boundary(self.h, ptr(out), n))
        return out

    def set_boundary(self, bc):
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        self._check(gpu().mocgpu_set_boundary(self.h, ptr(bc), bc.size))

    def store_old_flux(self):
        self._check(gpu().mocgpu_store_old_flux(self.h))

    def calc_fission_source(self, k: float) -> np.ndarray:
        out = np.zeros(self.F)
        self._check(gpu().mocgpu_calc_fission_source(self.h, k, ptr(out)))
        return out

    def set_fission_source(self, fs):
        fs = np.ascontiguousarray(fs, dtype=np.float64)
        self._check(gpu().mocgpu_set_fission_source(self.h, ptr(fs)))

    def build_source(self, g: int):
        self._check(gpu().mocgpu_build_source(self.h, g))

    def sweep(self, g: int, with_current: bool = False):
        self._check(gpu().mocgpu_sweep(self.h, g, 1 if with_current else 0))

    def sweep_all_groups(self, with_current: bool = False):
        """One inner sweep of all groups, enqueued without waiting."""
        self._check(gpu().mocgpu_sweep_all(self.h, 1 if with_current else 0))

    def synchronize(self):
        self._check(gpu().mocgpu_synchronize(self.h))

    def fss_step(self, with_current: bool = False):
        self._check(gpu().mocgpu_fss_step(self.h, 1 if with_current else 0))

    def total_fission(self, old: bool = False) -> float:
        out = np.zeros(1)
        self._check(gpu().mocgpu_total_fission(self.h, 1 if old else 0, ptr(out)))
        return float(out[0])

    def flux_residual(self) -> float:
        out = np.zeros(1)
        self._check(gpu().mocgpu_flux_residual(self.h, ptr(out)))
        return float(out[0])

    def get_pin_flux_1g(self, g: int) -> np.ndarray:
        out = np.zeros(self.npin)
        self._check(gpu().mocgpu_get_pin_flux_1g(self.h, g, ptr(out)))
        return out

    def set_pin_flux_1g(self, g: int, pin) -> float:
        pin = np.ascontiguousarray(pin, dtype=np.float64)
        r = np.zeros(1)
        self._check(gpu().mocgpu_set_pin_flux_1g(self.h, g, ptr(pin), ptr(r)))
        return float(r[0])

    def currents(self, g: int):
        n = gpu().mocgpu_n_surf(self.h)
        c, s = np.zeros(n), np.zeros(n)
        self._check(gpu().mocgpu_get_currents(self.h, g, ptr(c), ptr(s)))
        return c, s

    def set_currents(self, g: int, current, surface_flux):
        c = np.ascontiguousarray(current, dtype=np.float64)
        s = np.ascontiguousarray(surface_flux, dtype=np.float64)
        self._check(gpu().mocgpu_set_currents(self.h, g, ptr(c), ptr(s)))

    # --- 2-D/3-D correction worker (cmdo::MoCSweeper_2D3D) ------------------
    def sweep_2d3d(self, g: int):
        """MoCSweeper_2D3D::sweep(g): the last inner sweep tallies the
        correction factors and the coarse currents of group g."""
        self._check(gpu().mocgpu_sweep_2d3d(self.h, g))

    def corrections(self):
        """(alpha [G][4nd][n_cell][2], beta [G][4nd][n_cell])."""
        p = self.prob.p
        nang, n_cell = 2 * p.n_azi * p.n_polar, p.nx * p.ny
        a = np.zeros((self.G, nang, n_cell, 2))
        b = np.zeros((self.G, nang, n_cell))
        self._check(gpu().mocgpu_get_corrections(self.h, ptr(a), ptr(b)))
        return a, b

    def correction_residual(self, g: int) -> np.ndarray:
        out = np.zeros(3)
        self._check(gpu().mocgpu_correction_residual(self.h, g, ptr(out)))
        return out

    def update_incoming_flux(self):
        """MoCSweeper::update_incoming_flux from the CMFD partial currents."""
        self._check(gpu().mocgpu_update_incoming_flux(self.h))

    # --- CMFD / EigenSolver -------------------------------------------------
    def cmfd_step(self, k: float) -> float:
        """EigenSolver::do_cmfd: returns the CMFD eigenvalue."""
        kk = np.array([k])
        self._check(gpu().mocgpu_cmfd_step(self.h, ptr(kk)))
        return float(kk[0])

    def cmfd_iterations(self) -> int:
        return int(gpu().mocgpu_cmfd_iterations(self.h))

    def cmfd_bicgstab_iterations(self) -> int:
        """BiCGSTAB iterations of the last CMFD solve (all groups)."""
        return int(gpu().mocgpu_cmfd_bicgstab_iterations(self.h))

    def cmfd_flux(self) -> np.ndarray:
        """CMFD::flux(): coarse flux [n_cell, G] of the last solve."""
        out = np.zeros((self.npin, self.G))
        self._check(gpu().mocgpu_cmfd_flux(self.h, ptr(out)))
        return out

    def cmfd_matrix(self, g: int):
        """Assembled CMFD matrix of group g (CSR with sorted columns)."""
        n = self.npin
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(5 * n, np.int32)
        val = np.zeros(5 * n)
        nnz = gpu().mocgpu_cmfd_matrix(self.h, g, ptr(rp), ptr(col), ptr(val), 5 * n)
        if nnz < 0:
            self._check(nnz)
        return rp, col[:nnz], val[:nnz]

    def eigen_solve(self, k_tol: float, psi_tol: float, max_iter: int, min_iter: int = 0):
        st = EigenState()
        rc = gpu().mocgpu_eigen_solve(self.h, k_tol, psi_tol, max_iter, min_iter, C.byref(st))
        if rc not in (0, -5):
            self._check(rc)
        return rc, st

    def eigen_iterate(self, st: EigenState):
        self._check(gpu().mocgpu_eigen_iterate(self.h, C.byref(st)))
        return st

    def options(self):
        """Effective options (band_rays, lane_map after autotuning)."""
        from .native import MocgpuOptions
        o = MocgpuOptions()
        self._check(gpu().mocgpu_get_options(self.h, C.byref(o)))
        return o

    def state_bytes(self) -> int:
        return int(gpu().mocgpu_state_bytes(self.h))

    def geometry_bytes(self) -> int:
        return int(gpu().mocgpu_geometry_bytes(self.h))

    # --- measurement --------------------------------------------------------
    def updates_per_sweep(self) -> int:
        return int(gpu().mocgpu_updates_per_sweep(self.h))

    def time_sweeps(self, n: int):
        a, b = np.zeros(1), np.zeros(1)
        self._check(gpu().mocgpu_time_sweeps(self.h, n, ptr(a), ptr(b)))
        return float(a[0]), float(b[0])


class CapacityError(RuntimeError):
    pass


class GpuBatch:
    """Many problems advanced together (mocgpu_batch_*). Handles stay owned by
    the caller and must outlive their membership."""

    def __init__(self, budget_bytes: int = 0, use_graph: bool = True):
        self.b = C.c_void_p()
        rc = gpu().mocgpu_batch_create(int(budget_bytes), 1 if use_graph else 0, C.byref(self.b))
        if rc != 0:
            raise DeviceError(gpu().mocgpu_batch_last_error(None).decode())
        self.members = {}

    def _check(self, rc):
        if rc < 0:
            msg = gpu().mocgpu_batch_last_error(self.b).decode()
            if rc == -7:
                raise CapacityError(msg)
            raise DeviceError(f"mocgpu batch error {rc}: {msg}")
        return rc

    def add(self, s: GpuMoCSweeper, k_tol: float, psi_tol: float, max_iter: int, min_iter: int = 0) -> int:
        cid = C.c_int32()
        self._check(gpu().mocgpu_batch_add(self.b, s.h, k_tol, psi_tol, max_iter, min_iter, C.byref(cid)))
        self.members[cid.value] = s
        return cid.value

    def run(self, n_outer: int = 1):
        self._check(gpu().mocgpu_batch_run(self.b, n_outer))

    def poll(self, cap: int = 4096):
        ids = (C.c_int32 * cap)()
        sts = (EigenState * cap)()
        n = self._check(gpu().mocgpu_batch_poll(self.b, ids, sts, cap))
        return [(ids[i], sts[i]) for i in range(n)]

    def remove(self, cid: int):
        self._check(gpu().mocgpu_batch_remove(self.b, cid))
        self.members.pop(cid, None)

    def size(self) -> int:
        return int(gpu().mocgpu_batch_size(self.b))

    def bytes_used(self) -> int:
        return int(gpu().mocgpu_batch_bytes_used(self.b))

    def solve_all(self, chunk: int = 4, max_outer: int = 100000, compact: float = 0.25):
        """Run until every member has finished; returns {id: EigenState}.
        Finished problems are taken out of the batch (their handles keep the
        results) once they are at least a fraction `compact` of the problems
        still in it, so the launches shrink with the work."""
        done = {}
        pending = []
        outer = 0
        total = self.size()  # members whose handles were destroyed have left the batch
        while len(done) < total and outer < max_outer:
            self.run(chunk)
            outer += chunk
            for cid, st in self.poll():
                done[cid] = st
                pending.append(cid)
            if pending and len(done) < total and len(pending) >= compact * self.size():
                for cid in pending:
                    gpu().mocgpu_batch_remove(self.b, cid)
                pending = []
        return done

    def close(self):
        if self.b:
            gpu().mocgpu_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBatchSet:
    """Problems spread over several batches, each with its own stream, run
    together: one batch's coarse solves (a workgroup per problem) and the
    tails of its launches overlap with the other batches' sweeps. A new
    problem joins the batch with the fewest members. A problem's result does
    not depend on which batch holds it or what else is there (bitwise), so
    the set gives the same answers as one batch. Same methods as GpuBatch;
    ids are the set's own."""

    def __init__(self, n: int = 2, budget_bytes: int = 0, use_graph: bool = True):
        if n < 1:
            raise ValueError("a batch set needs at least one batch")
        self.batches = [GpuBatch(use_graph=use_graph) for _ in range(n)]
        self.budget = int(budget_bytes)
        self._where = {}  # set id -> (batch index, batch id)
        self._next = 0
        self._live = [0] * n  # unfinished members of each batch
        self._done = set()    # set ids reported finished and still members

    def add(self, s: GpuMoCSweeper, k_tol: float, psi_tol: float, max_iter: int, min_iter: int = 0) -> int:
        if self.budget > 0 and self.bytes_used() + s.state_bytes() > self.budget:
            raise CapacityError("device memory budget exhausted")
        j = min(range(len(self.batches)), key=lambda i: self.batches[i].size())
        cid = self.batches[j].add(s, k_tol, psi_tol, max_iter, min_iter)
        sid = self._next
        self._next += 1
        self._where[sid] = (j, cid)
        self._live[j] += 1
        return sid

    def run(self, n_outer: int = 1):
        """Enqueue n_outer outer iterations on every batch with unfinished
        members (returns at once; poll waits)."""
        for j, b in enumerate(self.batches):
            if self._live[j] > 0:
                b.run(n_outer)

    def poll(self, cap: int = 4096):
        back = {w: sid for sid, w in self._where.items()}
        out = []
        for j, b in enumerate(self.batches):
            if b.size() > 0:
                for cid, st in b.poll(cap):
                    sid = back[(j, cid)]
                    if sid not in self._done:
                        self._done.add(sid)
                        self._live[j] -= 1
                    out.append((sid, st))
        return out

    def remove(self, sid: int):
        j, cid = self._where.pop(sid)
        if sid in self._done:
            self._done.discard(sid)
        else:
            self._live[j] -= 1
        self.batches[j].remove(cid)

    def size(self) -> int:
        return sum(b.size() for b in self.batches)

    def bytes_used(self) -> int:
        return sum(b.bytes_used() for b in self.batches)

    def solve_all(self, chunk: int = 4, max_outer: int = 100000, compact: float = 0.25):
        """GpuBatch.solve_all over the set; returns {id: EigenState}."""
        done = {}
        pending = []
        outer = 0
        total = self.size()
        while len(done) < total and outer < max_outer:
            self.run(chunk)
            outer += chunk
            for sid, st in self.poll():
                done[sid] = st
                pending.append(sid)
            if pending and len(done) < total and len(pending) >= compact * self.size():
                for sid in pending:
                    self.remove(sid)
                pending = []
        return done

    def close(self):
        for b in self.batches:
            b.close()
        self._where.clear()
        self._done.clear()
        self._live = [0] * len(self.batches)

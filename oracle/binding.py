"""ctypes binding of the FP64 oracle (oracle/_build/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from mocc_amd.native import EigenState, MocgpuProblem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

i32, i64, f64, P = C.c_int32, C.c_int64, C.c_double, C.c_void_p


class OracleOptions(C.Structure):
    _fields_ = [("n_inner", i32), ("gs_boundary", i32), ("energy_mode", i32), ("exact_exp", i32),
                ("use_cmfd", i32), ("cmfd_max_iter", i32), ("cmfd_k_tol", f64),
                ("cmfd_psi_tol", f64), ("cmfd_resid_reduction", f64), ("n_threads", i32),
                ("cmfd_negative_fixup", i32)]


_L = None


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(LIB)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_default_options.argtypes = [C.POINTER(OracleOptions)]
        L.oracle_create.argtypes = [C.POINTER(MocgpuProblem), C.POINTER(OracleOptions), C.POINTER(P)]
        L.oracle_destroy.argtypes = [P]
        for fn in ("oracle_initialize", "oracle_store_old_flux"):
            getattr(L, fn).argtypes = [P]
        for fn in ("oracle_set_flux", "oracle_get_flux", "oracle_get_boundary", "oracle_set_boundary",
                   "oracle_set_fission_source"):
            getattr(L, fn).argtypes = [P, P]
        L.oracle_boundary_size.restype = i64
        L.oracle_boundary_size.argtypes = [P]
        L.oracle_calc_fission_source.argtypes = [P, f64, P]
        L.oracle_build_source.argtypes = [P, i32]
        L.oracle_get_source.argtypes = [P, P, P]
        L.oracle_sweep.argtypes = [P, i32, i32]
        L.oracle_fss_step.argtypes = [P, i32]
        L.oracle_total_fission.argtypes = [P, i32, P]
        L.oracle_flux_residual.argtypes = [P, P]
        L.oracle_get_currents.argtypes = [P, i32, P, P]
        L.oracle_n_surf.restype = i32
        L.oracle_n_surf.argtypes = [P]
        L.oracle_get_pin_flux_1g.argtypes = [P, i32, P]
        L.oracle_set_pin_flux_1g.argtypes = [P, i32, P, P]
        L.oracle_cmfd_step.argtypes = [P, P, P]
        L.oracle_cmfd_matrix.argtypes = [P, i32, P, P, P, i32]
        L.oracle_cmfd_flux.argtypes = [P, P]
        L.oracle_eigen_solve.argtypes = [P, f64, f64, i32, i32, C.POINTER(EigenState), P, i32]
        L.oracle_eigen_iterate.argtypes = [P, C.POINTER(EigenState)]
        L.oracle_exp_table.restype = f64
        L.oracle_exp_table.argtypes = [f64, i32, f64, f64]
        L.oracle_exp_table_node.restype = f64
        L.oracle_exp_table_node.argtypes = [i32, i32, f64, f64]
        L.oracle_time_sweeps.argtypes = [P, i32, P]
        L.oracle_bicgstab.argtypes = [i32, P, P, P, P, P, f64, i32, P, P]
        L.oracle_sweep_2d3d.argtypes = [P, i32]
        L.oracle_get_corrections.argtypes = [P, P, P]
        L.oracle_correction_residual.argtypes = [P, i32, P]
        L.oracle_update_incoming_flux.argtypes = [P]
        _L = L
    return _L


def _p(a):
    return a.ctypes.data_as(P)


class Oracle:
    """FP64 restatement of MoCSweeper/Source/CMFD/EigenSolver on one problem."""

    def __init__(self, prob, n_inner=1, gs_boundary=True, energy_mode=0, exact_exp=False,
                 use_cmfd=False, n_threads=0, **cmfd):
        L = lib()
        o = OracleOptions()
        L.oracle_default_options(C.byref(o))
        o.n_inner = n_inner
        o.gs_boundary = 1 if gs_boundary else 0
        o.energy_mode = energy_mode
        o.exact_exp = 1 if exact_exp else 0
        o.use_cmfd = 1 if use_cmfd else 0
        o.n_threads = n_threads
        for k, v in cmfd.items():
            setattr(o, k, v)
        self.prob = prob
        self.h = P()
        if L.oracle_create(prob.view, C.byref(o), C.byref(self.h)) != 0:
            raise RuntimeError(L.oracle_last_error().decode())
        self.F, self.G = prob.n_fsr, prob.n_group

    def __del__(self):
        try:
            if self.h:
                lib().oracle_destroy(self.h)
        except Exception:
            pass

    def initialize(self):
        lib().oracle_initialize(self.h)

    def flux(self):
        out = np.zeros((self.F, self.G))
        lib().oracle_get_flux(self.h, _p(out))
        return out

    def set_flux(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        lib().oracle_set_flux(self.h, _p(phi))

    def boundary(self):
        out = np.zeros(lib().oracle_boundary_size(self.h))
        lib().oracle_get_boundary(self.h, _p(out))
        return out

    def set_boundary(self, bc):
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        lib().oracle_set_boundary(self.h, _p(bc))

    def store_old_flux(self):
        lib().oracle_store_old_flux(self.h)

    def calc_fission_source(self, k):
        out = np.zeros(self.F)
        lib().oracle_calc_fission_source(self.h, k, _p(out))
        return out

    def set_fission_source(self, fs):
        fs = np.ascontiguousarray(fs, dtype=np.float64)
        lib().oracle_set_fission_source(self.h, _p(fs))

    def build_source(self, g):
        lib().oracle_build_source(self.h, g)

    def source(self):
        s, q = np.zeros(self.F), np.zeros(self.F)
        lib().oracle_get_source(self.h, _p(s), _p(q))
        return s, q

    def sweep(self, g, with_current=False):
        rc = lib().oracle_sweep(self.h, g, 1 if with_current else 0)
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())

    def fss_step(self, with_current=False):
        rc = lib().oracle_fss_step(self.h, 1 if with_current else 0)
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())

    def total_fission(self, old=False):
        out = np.zeros(1)
        lib().oracle_total_fission(self.h, 1 if old else 0, _p(out))
        return float(out[0])

    def flux_residual(self):
        out = np.zeros(1)
        lib().oracle_flux_residual(self.h, _p(out))
        return float(out[0])

    def currents(self, g):
        n = lib().oracle_n_surf(self.h)
        c, s = np.zeros(n), np.zeros(n)
        lib().oracle_get_currents(self.h, g, _p(c), _p(s))
        return c, s

    def pin_flux(self, g):
        out = np.zeros(self.prob.p.nx * self.prob.p.ny)
        lib().oracle_get_pin_flux_1g(self.h, g, _p(out))
        return out

    def set_pin_flux(self, g, pin):
        pin = np.ascontiguousarray(pin, dtype=np.float64)
        r = np.zeros(1)
        rc = lib().oracle_set_pin_flux_1g(self.h, g, _p(pin), _p(r))
        if rc:
            raise ValueError(lib().oracle_last_error().decode())
        return float(r[0])

    def cmfd_step(self, k):
        kk = np.array([k], dtype=np.float64)
        it = np.zeros(1, np.int32)
        rc = lib().oracle_cmfd_step(self.h, _p(kk), _p(it))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        return float(kk[0]), int(it[0])

    def cmfd_flux(self):
        p = self.prob.p
        out = np.zeros((p.nx * p.ny, self.G))
        lib().oracle_cmfd_flux(self.h, _p(out))
        return out

    def cmfd_matrix(self, g):
        n = self.prob.p.nx * self.prob.p.ny
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(5 * n, np.int32)
        val = np.zeros(5 * n)
        nnz = lib().oracle_cmfd_matrix(self.h, g, _p(rp), _p(col), _p(val), 5 * n)
        if nnz < 0:
            raise RuntimeError("no CMFD matrix assembled yet")
        return rp, col[:nnz], val[:nnz]

    def eigen_solve(self, k_tol, psi_tol, max_iter, min_iter=0):
        st = EigenState()
        hist = np.zeros(max_iter)
        rc = lib().oracle_eigen_solve(self.h, k_tol, psi_tol, max_iter, min_iter, C.byref(st), _p(hist),
                                      max_iter)
        return rc, st, hist[: st.iterations]

    def sweep_2d3d(self, g):
        """MoCSweeper_2D3D::sweep(g): the last inner sweep tallies the 2-D/3-D
        correction factors (and the coarse currents) of group g."""
        if lib().oracle_sweep_2d3d(self.h, g) != 0:
            raise ValueError("group out of range")

    def corrections(self):
        """(alpha [G][4nd][n_cell][2], beta [G][4nd][n_cell])."""
        p = self.prob.p
        n_cell, nang = p.nx * p.ny, 2 * p.n_azi * p.n_polar
        a = np.zeros((self.G, nang, n_cell, 2))
        b = np.zeros((self.G, nang, n_cell))
        lib().oracle_get_corrections(self.h, _p(a), _p(b))
        return a, b

    def correction_residual(self, g):
        out = np.zeros(3)
        lib().oracle_correction_residual(self.h, g, _p(out))
        return out

    def update_incoming_flux(self):
        lib().oracle_update_incoming_flux(self.h)

    def time_sweeps(self, n):
        out = np.zeros(1)
        lib().oracle_time_sweeps(self.h, n, _p(out))
        return float(out[0])


def exp_table(v, n=10000, xmin=-10.0, xmax=0.0):
    return lib().oracle_exp_table(v, n, xmin, xmax)


def exp_table_node(i, n=10000, xmin=-10.0, xmax=0.0):
    return lib().oracle_exp_table_node(i, n, xmin, xmax)

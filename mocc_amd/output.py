"""mocc's HDF5 output file for a problem solved on the device.

``solve(sweeper, ...)`` runs EigenSolver::solve on the device and keeps the
convergence history EigenSolver::output writes (eigen_solver.cpp:181-187,
293-315); ``write(...)`` produces the file mocc's driver writes
(driver.cpp:51-57) through libmocc_h5 (include/mocc_h5.h lists the datasets);
``read`` / ``read_string`` / ``datasets`` read one back.

libmocc_h5 links the HDF5 C library of this image. Its run path holds only
its own directory, so libhdf5 is loaded here first, by its full path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .native import LIB_DIR, EigenState, HostOutputInfo, HostProblem, gpu, host, ptr

HDF5_PREFIX = os.environ.get("MOCC_HDF5_PREFIX", "/opt/conda")
P, i32, i64 = C.c_void_p, C.c_int32, C.c_int64


class MoccOutput(C.Structure):
    _fields_ = [("fsr_flux", P), ("n_iter", i32), ("history", P), ("current", P), ("n_surf", i32)]


_h5 = None


def h5lib() -> C.CDLL:
    global _h5
    if _h5 is None:
        hdf5 = os.path.join(HDF5_PREFIX, "lib", "libhdf5.so.103")
        if not os.path.exists(hdf5):
            raise RuntimeError(f"the HDF5 C library is not at {hdf5} (set MOCC_HDF5_PREFIX)")
        C.CDLL(hdf5, mode=C.RTLD_GLOBAL)
        host()  # libmocc_host, which libmocc_h5 calls
        path = os.path.join(LIB_DIR, "libmocc_h5.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run mocc_amd.build.build_all() first")
        L = C.CDLL(path)
        L.mocc_h5_last_error.restype = C.c_char_p
        L.mocc_h5_write.argtypes = [P, C.c_char_p, C.POINTER(MoccOutput)]
        L.mocc_h5_read.restype = i64
        L.mocc_h5_read.argtypes = [C.c_char_p, C.c_char_p, P, i64, P, i32, C.POINTER(i32)]
        L.mocc_h5_read_string.restype = i32
        L.mocc_h5_read_string.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, i32]
        L.mocc_h5_list.restype = i32
        L.mocc_h5_list.argtypes = [C.c_char_p, C.c_char_p, i32]
        _h5 = L
    return _h5


def _err() -> str:
    return h5lib().mocc_h5_last_error().decode()


def output_info(hp: HostProblem) -> HostOutputInfo:
    out = HostOutputInfo()
    if host().mocc_host_get_output_info(hp.h, C.byref(out)) != 0:
        raise RuntimeError(host().mocc_host_last_error().decode())
    return out


def solve(sweeper, k_tol: float, psi_tol: float, max_iter: int, min_iter: int = 0):
    """EigenSolver::solve on the device with its convergence history:
    returns (rc, EigenState, history[n_iter, 4] of k, error_k, error_psi,
    seconds). rc is 0 or MOCGPU_ERR_MAXITER (-5); other failures raise."""
    st = EigenState()
    hist = np.zeros((max(int(max_iter), 1), 4))
    rc = gpu().mocgpu_eigen_solve_history(sweeper.h, k_tol, psi_tol, max_iter, min_iter, C.byref(st),
                                          ptr(hist), int(max_iter))
    if rc not in (0, -5):
        sweeper._check(rc)
    return rc, st, hist[: st.iterations].copy()


def write(hp: HostProblem, path: str, fsr_flux, history, currents=None):
    """Write mocc's output file for problem hp. fsr_flux: [F, G]; history:
    [n_iter, 4] (k, error_k, error_psi, seconds); currents: [G, n_surf] coarse
    surface currents, needed when the deck has CMFD on and sets
    <cmfd dump_current="true">."""
    phi = np.ascontiguousarray(fsr_flux, dtype=np.float64)
    if phi.shape != (hp.n_fsr, hp.n_group):
        raise ValueError(f"fsr_flux must be [{hp.n_fsr}, {hp.n_group}], got {phi.shape}")
    h = np.ascontiguousarray(history, dtype=np.float64).reshape(-1, 4)
    o = MoccOutput()
    o.fsr_flux = ptr(phi)
    o.n_iter = h.shape[0]
    o.history = ptr(h) if h.size else None
    if currents is not None:
        cur = np.ascontiguousarray(currents, dtype=np.float64)
        if cur.ndim != 2 or cur.shape[0] != hp.n_group:
            raise ValueError("currents must be [G, n_surf]")
        o.current = ptr(cur)
        o.n_surf = cur.shape[1]
    if h5lib().mocc_h5_write(hp.h, path.encode(), C.byref(o)) != 0:
        raise RuntimeError(_err())


def solve_and_write(sweeper, path: str, k_tol: float, psi_tol: float, max_iter: int, min_iter: int = 0):
    """The driver's run for a device sweeper: solve, then write the file."""
    rc, st, hist = solve(sweeper, k_tol, psi_tol, max_iter, min_iter)
    cur = None
    info = output_info(sweeper.prob)
    if info.dump_current and info.cmfd:
        cur = np.stack([sweeper.currents(g)[0] for g in range(sweeper.G)])
    write(sweeper.prob, path, sweeper.flux(), hist, cur)
    return rc, st, hist


def read(path: str, name: str) -> np.ndarray:
    """A numeric dataset as float64 with its shape."""
    L = h5lib()
    dims = np.zeros(8, dtype=np.int64)
    nd = i32(0)
    n = L.mocc_h5_read(path.encode(), name.encode(), None, 0, ptr(dims), 8, C.byref(nd))
    if n < 0 and "buffer too small" not in _err():
        raise KeyError(f"{name}: {_err()}")
    shape = tuple(int(d) for d in dims[: nd.value])
    out = np.zeros(int(np.prod(shape)) if shape else 1)
    n = L.mocc_h5_read(path.encode(), name.encode(), ptr(out), out.size, ptr(dims), 8, C.byref(nd))
    if n < 0:
        raise KeyError(f"{name}: {_err()}")
    return out.reshape(shape)


def read_string(path: str, name: str) -> str:
    L = h5lib()
    n = L.mocc_h5_read_string(path.encode(), name.encode(), None, 0)
    if n < 0:
        raise KeyError(f"{name}: {_err()}")
    buf = C.create_string_buffer(n + 1)
    L.mocc_h5_read_string(path.encode(), name.encode(), buf, n + 1)
    return buf.value.decode()


def datasets(path: str) -> list[str]:
    L = h5lib()
    n = L.mocc_h5_list(path.encode(), None, 0)
    if n < 0:
        raise RuntimeError(_err())
    buf = C.create_string_buffer(n + 1)
    L.mocc_h5_list(path.encode(), buf, n + 1)
    return [s for s in buf.value.decode().split("\n") if s]

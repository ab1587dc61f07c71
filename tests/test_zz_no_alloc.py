"""mocgpu_attenuation_noalloc, the entry point that allocates no device
memory, against mocgpu_attenuation (which stages its values in a buffer it
allocates on every call): bitwise the same results on two calls in a row, the
second with fewer values, and sizes past the fixed scratch are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from mocc_amd.native import gpu
    L = gpu()
    L.mocgpu_attenuation_noalloc.restype = C.c_int32
    L.mocgpu_attenuation_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    return L


def _both(L, tau):
    a = np.full_like(tau, np.nan)
    b = np.full_like(tau, np.nan)
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    assert L.mocgpu_attenuation(p(tau), p(a), tau.size) == 0
    assert L.mocgpu_attenuation_noalloc(p(tau), p(b), tau.size) == 0
    return a, b


def test_noalloc_matches_allocating_entry_point():
    L = _lib()
    n_max = 262144  # MOCGPU_ATTENUATION_NOALLOC_MAX
    rng = np.random.default_rng(17)
    big = np.concatenate([rng.uniform(0, 12, n_max - 6), [0.0, 1e-30, 0.25, 10.0, 40.0, 1e30]]).astype(np.float32)
    small = rng.uniform(0, 0.3, 1000).astype(np.float32)
    print(f"noalloc call 1: n={big.size}", flush=True)
    a, b = _both(L, big)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    print(f"noalloc call 2: n={small.size}", flush=True)
    a, b = _both(L, small)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # past the scratch: refused before any device work, output untouched
    over = np.zeros(n_max + 1, np.float32)
    out = np.full_like(over, 7.0)
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    assert L.mocgpu_attenuation_noalloc(p(over), p(out), over.size) == -7
    assert (out == 7.0).all()
    assert L.mocgpu_attenuation_noalloc(None, p(out), 4) == -1

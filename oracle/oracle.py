"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this module.  It restates the reference's single-rank FFTW Poisson path; see
dnsp_oracle.c for the reference lines each function follows.

Arrays are numpy float64 in C order with shape (ny, nz, nx), which is the same
memory as the reference's Fortran ``P(0:nx-1, 0:nz-1, 0:ny-1)`` (x fastest).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdnsp_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle shared library in place (gcc, OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


_BASELINE_PATH = os.path.join(_HERE, "libdnsp_baseline.so")
_baseline = None


def baseline_lib():
    """The same source built for speed (-O3 -march=x86-64-v3, contraction
    allowed): the CPU baseline bench.py times.  Never the correctness oracle."""
    global _baseline
    if _baseline is None:
        if not os.path.exists(_BASELINE_PATH):
            build()
        _baseline = _bind(ctypes.CDLL(_BASELINE_PATH))
    return _baseline


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    """Argument and result types of the exported functions."""
    L.orc_r2hc.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_hc2r.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_cfft_inplace.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
    L.orc_yp_ext.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_mk_P_abc.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]
    L.orc_eigen.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _ip]
    L.orc_get_lo_n_bounds_1d.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip]
    L.orc_poisson_forward.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.orc_poisson_backward.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    bands = [_dp] * 6
    L.orc_poisson_tridiag.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + bands + [_dp]
    L.orc_poisson_solve.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + bands + [_dp]
    L.orc_poisson_solve.restype = ctypes.c_int
    L.orc_min_pivot.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + bands
    L.orc_min_pivot.restype = ctypes.c_double
    L.orc_apply_operator.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + bands + [_dp, _dp]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_get_threads.restype = ctypes.c_int
    return L


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def set_threads(n, L=None):
    (L or lib()).orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def r2hc(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_r2hc(x.size, _p(x), _p(out))
    return out


def hc2r(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_hc2r(x.size, _p(x), _p(out))
    return out


def cfft(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    v = x.view(np.float64)
    lib().orc_cfft_inplace(x.size, _p(v), int(bool(inverse)))
    return x


def yp_ext_from_yu(yu):
    yu = np.ascontiguousarray(yu, dtype=np.float64)
    ny = yu.size - 1
    out = np.empty(ny + 2)
    lib().orc_yp_ext(ny, _p(yu), _p(out))
    return out


def mk_P_abc(yp_ext, is_dirich_first=False):
    """Returns (a, b, c), each of length ny = len(yp_ext) - 2."""
    yp_ext = np.ascontiguousarray(yp_ext, dtype=np.float64)
    ny = yp_ext.size - 2
    a, b, c = (np.empty(ny) for _ in range(3))
    lib().orc_mk_P_abc(ny, _p(yp_ext), int(bool(is_dirich_first)), _p(a), _p(b), _p(c))
    return a, b, c


def bands_from_yu(yu):
    """(a, b, c, a00, b00, c00) for faces yu, as calc_bands.py:282-283 writes them."""
    ype = yp_ext_from_yu(yu)
    return mk_P_abc(ype, False) + mk_P_abc(ype, True)


def eigen(n, inv_d):
    lam = np.empty(n)
    is00 = np.empty(n, dtype=np.int32)
    lib().orc_eigen(n, float(inv_d), _p(lam), _pi(is00))
    return lam, is00


def get_lo_n_bounds_1d(mpi_divs, n_glob):
    lo = np.empty(mpi_divs, dtype=np.int32)
    n = np.empty(mpi_divs, dtype=np.int32)
    lib().orc_get_lo_n_bounds_1d(int(mpi_divs), int(n_glob), _pi(lo), _pi(n))
    return lo, n


def _bands(bands):
    return [_p(np.ascontiguousarray(b, dtype=np.float64)) for b in bands]


def forward(P):
    """Stages S1-S3: (ny, nz, nx) field -> real-real halfcomplex spectrum, same shape."""
    S = np.ascontiguousarray(P, dtype=np.float64).copy()
    ny, nz, nx = S.shape
    lib().orc_poisson_forward(nx, nz, ny, _p(S))
    return S


def backward(S):
    """Stages S5-S8 including the 1/(nx*nz) scale."""
    P = np.ascontiguousarray(S, dtype=np.float64).copy()
    ny, nz, nx = P.shape
    lib().orc_poisson_backward(nx, nz, ny, _p(P))
    return P


def tridiag(S, inv_dx, inv_dz, bands):
    S = np.ascontiguousarray(S, dtype=np.float64).copy()
    ny, nz, nx = S.shape
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bands]
    lib().orc_poisson_tridiag(nx, nz, ny, float(inv_dx), float(inv_dz), *_bands(bs), _p(S))
    return S


def solve(rhs, inv_dx, inv_dz, bands):
    """Whole solve: returns the solution for an (ny, nz, nx) RHS."""
    P = np.ascontiguousarray(rhs, dtype=np.float64).copy()
    ny, nz, nx = P.shape
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bands]
    rc = lib().orc_poisson_solve(nx, nz, ny, float(inv_dx), float(inv_dz), *_bands(bs), _p(P))
    if rc != 0:
        raise ValueError("oracle solve rejected the sizes")
    return P


def solve_inplace(P, inv_dx, inv_dz, bands, L=None):
    """In-place solve; L: the library to run (default the oracle)."""
    ny, nz, nx = P.shape
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bands]
    rc = (L or lib()).orc_poisson_solve(nx, nz, ny, float(inv_dx), float(inv_dz), *_bands(bs), _p(P))
    if rc != 0:
        raise ValueError("oracle solve rejected the sizes")


def min_pivot(nx, nz, ny, inv_dx, inv_dz, bands):
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bands]
    return float(lib().orc_min_pivot(nx, nz, ny, float(inv_dx), float(inv_dz), *_bands(bs)))


def apply_operator(p, inv_dx, inv_dz, bands):
    p = np.ascontiguousarray(p, dtype=np.float64)
    ny, nz, nx = p.shape
    out = np.empty_like(p)
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bands]
    lib().orc_apply_operator(nx, nz, ny, float(inv_dx), float(inv_dz), *_bands(bs), _p(p), _p(out))
    return out

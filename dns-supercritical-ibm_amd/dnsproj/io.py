"""The reference's on-disk formats over libdnsproj.so (host side): field
files (src/mod_io.f90), coefficient text files (calc_bands.py writer,
read_column_r8 in src/mod_utils.f90) and yu.txt faces."""
import ctypes

import numpy as np

from . import _check, _dp, _ip, lib

_i, _c = ctypes.c_int, ctypes.c_char_p

_SIGS = [
    ("dnsp_io_field_shape", [_c, _ip, _ip, _ip]),
    ("dnsp_io_read_field", [_c, _i, _i, _i, _i, _dp]),
    ("dnsp_io_write_field", [_c, _i, _i, _i, _i, _dp]),
    ("dnsp_io_read_column", [_c, _i, _i, _i, _dp]),
    ("dnsp_io_write_table", [_c, _i, _i, _dp]),
    ("dnsp_io_read_faces", [_c, _ip, _dp]),
]
_ready = False


def _lib():
    global _ready
    L = lib()
    if not _ready:
        for name, args in _SIGS:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _ready = True
    return L


def exported_names():
    return [s[0] for s in _SIGS]


def _p(path):
    return str(path).encode()


def field_shape(path):
    a, b, c = _i(), _i(), _i()
    _check(_lib().dnsp_io_field_shape(_p(path), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value, c.value


def read_field(path, nx, nz, ny, halo=False):
    """(ny, nz, nx) array, or the (ny+2, nz+2, nx+2) halo array (zero halo) when halo."""
    shape = (ny + 2, nz + 2, nx + 2) if halo else (ny, nz, nx)
    A = np.zeros(shape)
    _check(_lib().dnsp_io_read_field(_p(path), int(nx), int(nz), int(ny), int(bool(halo)), A.ctypes.data_as(_dp)))
    return A


def write_field(path, A, halo=False):
    A = np.ascontiguousarray(A, dtype=np.float64)
    ny, nz, nx = A.shape
    if halo:
        nx, nz, ny = nx - 2, nz - 2, ny - 2
    _check(_lib().dnsp_io_write_field(_p(path), nx, nz, ny, int(bool(halo)), A.ctypes.data_as(_dp)))


def read_column(path, icol, n_prev, n):
    out = np.zeros(max(int(n), 0))
    _check(_lib().dnsp_io_read_column(_p(path), int(icol), int(n_prev), int(n), out.ctypes.data_as(_dp)))
    return out


def write_table(path, rows):
    rows = np.ascontiguousarray(np.atleast_2d(rows), dtype=np.float64)
    if rows.size == 0:
        rows = rows.reshape(0, 1)
    _check(_lib().dnsp_io_write_table(_p(path), rows.shape[0], rows.shape[1], rows.ctypes.data_as(_dp)))


def read_faces(path):
    ny = _i()
    _check(_lib().dnsp_io_read_faces(_p(path), ctypes.byref(ny), None))
    yu = np.zeros(ny.value + 1)
    _check(_lib().dnsp_io_read_faces(_p(path), ctypes.byref(ny), yu.ctypes.data_as(_dp)))
    return yu

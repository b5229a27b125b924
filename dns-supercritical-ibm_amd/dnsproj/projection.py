"""Projection stencils and IBM forcing over libdnsproj.so.

Names follow the reference (src/mod_cfd_subroutines.f90, src/mod_utils.f90,
src/diezdecomp_api_ibm.f90, src/ibm_setup/).  Halo arrays are torch float64
tensors of shape (ny+2, nz+2, nx+2) on a HIP device (the reference's
(-1:nx, -1:nz, -1:ny) fields); P is (ny, nz, nx).  1-D coefficients, rho_min
and wall values are device tensors too.
"""
import ctypes

import numpy as np

from . import DnspError, _check, _dp, _ip, _stream_ptr, _vp, lib

_i, _d = ctypes.c_int, ctypes.c_double

_SIGS = [
    ("dnsp_halo_periodic_xz", [_vp, _i, _i, _i, _vp]),
    ("dnsp_set_val_slice", [_vp, _i, _i, _i, _i, _i, _d, _vp]),
    ("dnsp_neumann_ghosts_y", [_vp, _i, _i, _i, _vp]),
    ("dnsp_build_p_rhs", [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _vp]),
    ("dnsp_build_p_rhs_varprop", [_vp] * 11 + [_i, _i, _i] + [_d] * 5 + [_vp]),
    ("dnsp_rho_min", [_vp, _i, _i, _i, _vp, _vp]),
    ("dnsp_update_vars", [_vp] * 15 + [_i, _i, _i, _i, _d, _d, _d, _i, _vp]),
    ("dnsp_geom_coeffs", [_i, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("dnsp_ibm_create", [_i, _i, _i, _i, _i, _i, _ip, _dp, ctypes.POINTER(_vp)]),
    ("dnsp_ibm_exec", [_vp, _vp, _vp]),
    ("dnsp_ibm_info", [_vp, _ip, _ip, _ip]),
    ("dnsp_ibm_free", [_vp]),
    ("dnsp_ibm_read_coeffs", [ctypes.c_char_p, _i, _i, _i, _ip, _ip, _dp]),
    ("dnsp_ibm_div_coeffs", [_i, _ip, _dp, _i, _i, _i, _d, _d, _dp]),
    ("dnsp_ibm_read_slab", [ctypes.c_char_p, _i, _i, _i, _i, _i, _ip]),
    ("dnsp_ibm_part2", [_vp, _vp, _vp, _i, _i, _i, _i, _d, _d, _vp]),
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


def _dev(t, shape=None):
    import torch
    if t is None:
        return _vp(None)
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise ValueError("expected a contiguous float64 tensor on a HIP device")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
    return _vp(t.data_ptr())


def _idev(t, n):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.int32 or t.numel() != n:
        raise ValueError("expected a contiguous int32 device tensor of the right size")
    return _vp(t.data_ptr())


def _halo_shape(A):
    ny2, nz2, nx2 = A.shape
    return nx2 - 2, nz2 - 2, ny2 - 2


def halo_periodic_xz(A, stream=None):
    nx, nz, ny = _halo_shape(A)
    _check(_lib().dnsp_halo_periodic_xz(_dev(A), nx, nz, ny, _vp(_stream_ptr(stream))))
    return A


def set_val_slice(A, k0, k1, val, stream=None):
    nx, nz, ny = _halo_shape(A)
    _check(_lib().dnsp_set_val_slice(_dev(A), nx, nz, ny, int(k0), int(k1), float(val), _vp(_stream_ptr(stream))))
    return A


def neumann_ghosts_y(A, stream=None):
    nx, nz, ny = _halo_shape(A)
    _check(_lib().dnsp_neumann_ghosts_y(_dev(A), nx, nz, ny, _vp(_stream_ptr(stream))))
    return A


def build_p_rhs(P, U_hat, W_hat, V_hat, inv_dy_p, inv_dtmax, inv_dx, inv_dz, stream=None):
    ny, nz, nx = P.shape
    hs = (ny + 2, nz + 2, nx + 2)
    _check(_lib().dnsp_build_p_rhs(_dev(P), _dev(U_hat, hs), _dev(W_hat, hs), _dev(V_hat, hs), _dev(inv_dy_p, (ny,)),
                                   nx, nz, ny, float(inv_dtmax), float(inv_dx), float(inv_dz),
                                   _vp(_stream_ptr(stream))))
    return P


def build_p_rhs_varprop(P, U_hat, W_hat, V_hat, P_now, P_old, rho, div_now, inv_dy_p, inv_dy_p_pad, rho_min,
                        inv_dtmax, inv_dx, inv_dz, inv_dx2, inv_dz2, stream=None):
    ny, nz, nx = P.shape
    hs = (ny + 2, nz + 2, nx + 2)
    _check(_lib().dnsp_build_p_rhs_varprop(
        _dev(P), _dev(U_hat, hs), _dev(W_hat, hs), _dev(V_hat, hs), _dev(P_now, hs), _dev(P_old, hs), _dev(rho, hs),
        _dev(div_now, hs), _dev(inv_dy_p, (ny,)), _dev(inv_dy_p_pad, (ny + 1,)), _dev(rho_min, (1,)), nx, nz, ny,
        float(inv_dtmax), float(inv_dx), float(inv_dz), float(inv_dx2), float(inv_dz2), _vp(_stream_ptr(stream))))
    return P


def rho_min(rho, out, stream=None):
    nx, nz, ny = _halo_shape(rho)
    _check(_lib().dnsp_rho_min(_dev(rho), nx, nz, ny, _dev(out, (1,)), _vp(_stream_ptr(stream))))
    return out


def update_vars(P, P_per, P_now, P_old, U, V, W, U_hat, V_hat, W_hat, rho, dPdy_ddy_hi, dPdy_ddy_lo, rho_min,
                wall_bc, kmin_Vhat, dtmax, inv_dx, inv_dz, use_incompressible=True, stream=None):
    ny, nz, nx = P.shape
    hs = (ny + 2, nz + 2, nx + 2)
    _check(_lib().dnsp_update_vars(
        _dev(P, (ny, nz, nx)), _dev(P_per, hs), _dev(P_now, hs), _dev(P_old, hs), _dev(U, hs), _dev(V, hs),
        _dev(W, hs), _dev(U_hat, hs), _dev(V_hat, hs), _dev(W_hat, hs), _dev(rho, hs), _dev(dPdy_ddy_hi, (ny,)),
        _dev(dPdy_ddy_lo, (ny,)), _dev(rho_min, (1,)), _dev(wall_bc, (6,)), nx, nz, ny, int(kmin_Vhat),
        float(dtmax), float(inv_dx), float(inv_dz), int(bool(use_incompressible)), _vp(_stream_ptr(stream))))


def geom_coeffs(yu):
    """Host: inv_dy_p, inv_dy_u, inv_dy_p_pad, dPdy_ddy_lo, dPdy_ddy_hi for faces yu."""
    yu = np.ascontiguousarray(yu, dtype=np.float64)
    ny = yu.size - 1
    out = dict(inv_dy_p=np.empty(ny), inv_dy_u=np.empty(ny + 1), inv_dy_p_pad=np.empty(ny + 1),
               dPdy_ddy_lo=np.empty(ny), dPdy_ddy_hi=np.empty(ny))
    _check(_lib().dnsp_geom_coeffs(ny, yu.ctypes.data_as(_dp), *[out[k].ctypes.data_as(_dp) for k in
                                                                  ("inv_dy_p", "inv_dy_u", "inv_dy_p_pad",
                                                                   "dPdy_ddy_lo", "dPdy_ddy_hi")]))
    return out


# ---------------------------------------------------------------------------
# IBM
# ---------------------------------------------------------------------------
class IbmForcing:
    """diezDecomp_ibm_type on one GPU: create once, exec on a padded field."""

    def __init__(self, nx, nz, ny, nhalo, nskip, ijk, c):
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 9)
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3)
        if ijk.shape[0] != c.shape[0]:
            raise ValueError("ijk and c must have the same number of records")
        h = _vp()
        _check(_lib().dnsp_ibm_create(int(nx), int(nz), int(ny), int(nhalo), int(nskip), ijk.shape[0],
                                      ijk.ctypes.data_as(_ip), c.ctypes.data_as(_dp), ctypes.byref(h)))
        self._h = h
        self.nx, self.nz, self.ny, self.nside = int(nx), int(nz), int(ny), int(nhalo) + int(nskip)

    @property
    def padded_shape(self):
        s = 2 * self.nside
        return (self.ny + s, self.nz + s, self.nx + s)

    @property
    def info(self):
        a, b, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib().dnsp_ibm_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(d)))
        return {"n_forcings": a.value, "n_images": b.value, "disjoint": bool(d.value)}

    def exec(self, A, stream=None):
        _check(_lib().dnsp_ibm_exec(self._h, _dev(A, self.padded_shape), _vp(_stream_ptr(stream))))
        return A

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib().dnsp_ibm_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ibm_read_coeffs(path, nx, nz, ny):
    """Returns ijk (n, 9) int32 and c (n, 3) = (cw, c1, c2)."""
    n = ctypes.c_int()
    p = str(path).encode()
    _check(_lib().dnsp_ibm_read_coeffs(p, nx, nz, ny, ctypes.byref(n), None, None))
    ijk = np.zeros((n.value, 9), dtype=np.int32)
    c = np.zeros((n.value, 3))
    _check(_lib().dnsp_ibm_read_coeffs(p, nx, nz, ny, ctypes.byref(n), ijk.ctypes.data_as(_ip),
                                       c.ctypes.data_as(_dp)))
    return ijk, c


def ibm_div_coeffs(ijk, yp, nx, nz, inv_dx, inv_dz):
    ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 9)
    yp = np.ascontiguousarray(yp, dtype=np.float64)
    c = np.zeros((ijk.shape[0], 3))
    _check(_lib().dnsp_ibm_div_coeffs(ijk.shape[0], ijk.ctypes.data_as(_ip), yp.ctypes.data_as(_dp), int(nx),
                                      int(nz), yp.size, float(inv_dx), float(inv_dz), c.ctypes.data_as(_dp)))
    return c


def ibm_read_slab(path, nx, nz, ny, k_first=0, is_top=False):
    out = np.zeros((nz, nx), dtype=np.int32)
    _check(_lib().dnsp_ibm_read_slab(str(path).encode(), int(nx), int(nz), int(ny), int(k_first), int(bool(is_top)),
                                     out.ctypes.data_as(_ip)))
    return out


def ibm_part2(A, bot, top, kmax, wall_bot, wall_top, stream=None):
    nx, nz, ny = _halo_shape(A)
    _check(_lib().dnsp_ibm_part2(_dev(A), _idev(bot, nx * nz), _idev(top, nx * nz), nx, nz, ny, int(kmax),
                                 float(wall_bot), float(wall_top), _vp(_stream_ptr(stream))))
    return A


__all__ = ["DnspError"] + [n for n in dir() if not n.startswith("_")]


# ---------------------------------------------------------------------------
# projection step object (include/dnsproj.h: dnsp_step_*)
# ---------------------------------------------------------------------------
STEP_FIELDS = ("U_hat", "V_hat", "W_hat", "U", "V", "W", "rho", "P_now", "P_old", "div_now", "P_per", "P")

_STEP_SIGS = [
    ("dnsp_step_create", [_i, _i, _i, _dp, _d, _d, _d, _i, _i, ctypes.POINTER(_vp)]),
    ("dnsp_step_field", [_vp, _i, ctypes.POINTER(_vp)]),
    ("dnsp_step_write", [_vp, _i, _vp, _vp]),
    ("dnsp_step_read", [_vp, _i, _vp, _vp]),
    ("dnsp_step_set_walls", [_vp, _dp]),
    ("dnsp_step_set_ibm", [_vp, _i, _vp]),
    ("dnsp_step_set_columns", [_vp, _ip, _ip]),
    ("dnsp_step_set_component_columns", [_vp, ctypes.c_int, _ip, _ip]),
    ("dnsp_step_poisson", [_vp, ctypes.POINTER(_vp)]),
    ("dnsp_step_footprint", [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    ("dnsp_step_enqueue", [_vp, _i, _vp]),
    ("dnsp_step_capture", [_vp, _i, _vp]),
    ("dnsp_step_launch", [_vp, _vp]),
    ("dnsp_step_free", [_vp]),
]
_SIGS.extend(_STEP_SIGS)


class IbmForcingSet:
    """IBM handles of U, V, W for one case (records from workloads.rough_wall_ibm)."""

    def __init__(self, nx, nz, ny, records):
        self.handles = {}
        for comp in ("U", "V", "W"):
            ijk, c = records[comp]
            self.handles[comp] = IbmForcing(nx, nz, ny, 1, 0, ijk, c)

    def close(self):
        for h in self.handles.values():
            h.close()
        self.handles = {}


class ProjectionStep:
    """Resident fields of one channel case and its projection step on the device."""

    def __init__(self, nx, nz, ny, yu, inv_dx, inv_dz, dtmax, incompressible=False, precision=0):
        L = _lib()
        yu = np.ascontiguousarray(yu, dtype=np.float64)
        if yu.size != ny + 1:
            raise ValueError("yu must hold ny+1 faces")
        h = _vp()
        _check(L.dnsp_step_create(int(nx), int(nz), int(ny), yu.ctypes.data_as(_dp), float(inv_dx), float(inv_dz),
                                  float(dtmax), int(bool(incompressible)), int(precision), ctypes.byref(h)))
        self._h = h
        self.nx, self.nz, self.ny = int(nx), int(nz), int(ny)
        self._ibm = None

    @classmethod
    def from_case(cls, c, precision=0, with_ibm=True):
        s = cls(c["nx"], c["nz"], c["ny"], c["yu"], c["inv_dx"], c["inv_dz"], c["dtmax"], precision=precision)
        s.set_walls(c["walls"])
        if with_ibm and c.get("ibm") is not None:
            s.set_ibm(IbmForcingSet(c["nx"], c["nz"], c["ny"], c["ibm"]))
            s.set_columns(c["bot"], c["top"])
        return s

    def shape(self, name):
        if name == "P":
            return (self.ny, self.nz, self.nx)
        return (self.ny + 2, self.nz + 2, self.nx + 2)

    def write(self, name, t, stream=None):
        _check(_lib().dnsp_step_write(self._h, STEP_FIELDS.index(name), _dev(t, self.shape(name)),
                                      _vp(_stream_ptr(stream))))

    def read(self, name, stream=None):
        import torch
        out = torch.empty(self.shape(name), dtype=torch.float64, device="cuda")
        _check(_lib().dnsp_step_read(self._h, STEP_FIELDS.index(name), _vp(out.data_ptr()), _vp(_stream_ptr(stream))))
        return out

    def load(self, fields, stream=None):
        """Writes host arrays (dict name -> numpy array) into the resident fields."""
        import torch
        for name, a in fields.items():
            self.write(name, torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda"), stream)

    def set_walls(self, walls):
        w = np.ascontiguousarray(walls, dtype=np.float64)
        _check(_lib().dnsp_step_set_walls(self._h, w.ctypes.data_as(_dp)))

    def set_ibm(self, ibm_set):
        self._ibm = ibm_set  # the handles must outlive the step
        for k, comp in enumerate(("U", "V", "W")):
            h = ibm_set.handles.get(comp) if ibm_set else None
            _check(_lib().dnsp_step_set_ibm(self._h, k, h._h if h is not None else _vp(None)))

    def set_columns(self, bot, top, component=None):
        """Column slabs (nz, nx) of U, V and W alike, or of one component
        (0, 1, 2 or "U", "V", "W"); None removes them."""
        if bot is None:
            b = t = _ip()
        else:
            b = np.ascontiguousarray(bot, dtype=np.int32)
            t = np.ascontiguousarray(top, dtype=np.int32)
            b, t = b.ctypes.data_as(_ip), t.ctypes.data_as(_ip)
        if component is None:
            _check(_lib().dnsp_step_set_columns(self._h, b, t))
        else:
            k = "UVW".index(component) if isinstance(component, str) else int(component)
            _check(_lib().dnsp_step_set_component_columns(self._h, k, b, t))

    @property
    def footprint(self):
        v = ctypes.c_size_t()
        _check(_lib().dnsp_step_footprint(self._h, ctypes.byref(v)))
        return v.value

    def poisson_handle(self):
        v = _vp()
        _check(_lib().dnsp_step_poisson(self._h, ctypes.byref(v)))
        return v

    def autotune(self, reps=5, stream=None):
        """Times the kernel variants of the step's Poisson solver (FP64 and,
        for a mixed-precision step, mixed) on a temporary array and keeps the
        fastest (dnsp_poisson_autotune); field contents are untouched."""
        _check(lib().dnsp_poisson_autotune(self.poisson_handle(), _vp(None), int(reps), _vp(_stream_ptr(stream)),
                                           ctypes.cast(None, _dp)))

    def enqueue(self, refinements=-1, stream=None):
        _check(_lib().dnsp_step_enqueue(self._h, int(refinements), _vp(_stream_ptr(stream))))

    def capture(self, refinements=-1, stream=None):
        _check(_lib().dnsp_step_capture(self._h, int(refinements), _vp(_stream_ptr(stream))))

    def launch(self, stream=None):
        _check(_lib().dnsp_step_launch(self._h, _vp(_stream_ptr(stream))))

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib().dnsp_step_free(self._h)
            self._h = None
        if self._ibm is not None:
            self._ibm.close()
            self._ibm = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# rough-wall IBM coefficient generation (include/dnsproj.h: dnsp_ibm_generate)
# ---------------------------------------------------------------------------
class _Surface(ctypes.Structure):
    _fields_ = [("n_modes", _i), ("mnab", _dp), ("Lx", _d), ("Lz", _d), ("Ly", _d)]


_GEN_SIGS = [
    ("dnsp_ibm_generate", [ctypes.POINTER(_Surface), _i, _i, _i, _dp, _dp, _dp, _i, _i, _i, _d, _d, _i, _ip, _ip,
                           _dp, _ip, _ip]),
    ("dnsp_ibm_write_coeffs", [ctypes.c_char_p, _i, _ip, _dp]),
    ("dnsp_ibm_write_slab", [ctypes.c_char_p, _i, _i, _ip]),
]
_SIGS.extend(_GEN_SIGS)


def read_mnab(path):
    """Fourier modes of a rough surface file (count, then rows m n a b), with a
    and b halved as the examples' init_surf_obj does on reading."""
    from . import io as IO
    n = int(open(path).readline().split()[0])
    cols = [IO.read_column(path, c, 0, n) for c in range(4)]
    return np.stack([cols[0], cols[1], 0.5 * cols[2], 0.5 * cols[3]], axis=1)


def ibm_generate(mnab, Lx, Lz, Ly, x, z, y, cc_x=1, cc_z=1, quady=False, bc_bot=0.0, bc_top=0.0):
    """Records (ijk (n, 9), c (n, 3): cw, c1, c2) sorted by target (k, j, i),
    and the column slabs (bottom, top) as (nz*cc_z, nx*cc_x) int arrays."""
    m = np.ascontiguousarray(mnab, dtype=np.float64).reshape(-1, 4)
    surf = _Surface(m.shape[0], m.ctypes.data_as(_dp), float(Lx), float(Lz), float(Ly))
    x, z, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, z, y))
    args = (ctypes.byref(surf), x.size, z.size, y.size, x.ctypes.data_as(_dp), z.ctypes.data_as(_dp),
            y.ctypes.data_as(_dp), int(cc_x), int(cc_z), int(bool(quady)), float(bc_bot), float(bc_top))
    n = _i()
    _check(_lib().dnsp_ibm_generate(*args, 0, ctypes.byref(n), None, None, None, None))
    ijk = np.zeros((n.value, 9), dtype=np.int32)
    c = np.zeros((n.value, 3))
    NX, NZ = x.size * int(cc_x), z.size * int(cc_z)
    sb = np.zeros((NZ, NX), dtype=np.int32)
    st = np.zeros((NZ, NX), dtype=np.int32)
    _check(_lib().dnsp_ibm_generate(*args, n.value, ctypes.byref(n), ijk.ctypes.data_as(_ip), c.ctypes.data_as(_dp),
                                    sb.ctypes.data_as(_ip), st.ctypes.data_as(_ip)))
    return ijk, c, sb, st


def ibm_write_coeffs(path, ijk, c):
    ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 9)
    c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3)
    _check(_lib().dnsp_ibm_write_coeffs(str(path).encode(), ijk.shape[0], ijk.ctypes.data_as(_ip),
                                        c.ctypes.data_as(_dp)))


def ibm_write_slab(path, slab):
    s = np.ascontiguousarray(slab, dtype=np.int32)
    nz, nx = s.shape
    _check(_lib().dnsp_ibm_write_slab(str(path).encode(), nx, nz, s.ctypes.data_as(_ip)))

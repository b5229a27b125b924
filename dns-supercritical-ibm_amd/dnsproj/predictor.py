"""The time-step stages before the projection (momentum predictor, T,
thermophysical properties, energy right-hand side) over libdnsproj.so; names
follow src/mod_cfd_subroutines.f90 (build_vars, update_props,
build_rhs_energy).  Fields are torch float64 tensors on a HIP device: halo
arrays (ny+2, nz+2, nx+2), the Adams-Bashforth histories and RHS_energy
(ny, nz, nx).  The predictor works on a ProjectionStep's U, V, W, hats, rho
and div_now; Predictor.enqueue then step.enqueue is one time step.
"""
import ctypes

import numpy as np

from . import _check, _dp, _stream_ptr, _vp, lib
from .projection import _dev

_i, _d = ctypes.c_int, ctypes.c_double

PRED_FIELDS = ("T", "Ru_old", "Rv_old", "Rw_old", "Rt_old", "mu", "cond", "buoy", "cp", "enth", "divjac", "T_ext",
               "RHS_energy")
INTERIOR = ("Ru_old", "Rv_old", "Rw_old", "Rt_old", "RHS_energy")
QUANTITIES = ("enth", "rho", "mu", "cond", "cp", "drho_dT")
LAW_CONST, LAW_POWER, LAW_CUBIC = 0, 1, 2


class PredParams(ctypes.Structure):
    _fields_ = [("nu", _d), ("cond", _d), ("Sf_x", _d), ("Sf_z", _d), ("Sq", _d), ("Gb_x", _d), ("Gb_y", _d),
                ("Gb_z", _d), ("walls", _d * 8), ("kmin_Vhat", _i)]


_SIGS = [
    ("dnsp_step_info", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_d),
                        ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_i)]),
    ("dnsp_predictor_geom", [_i, _dp, _dp]),
    ("dnsp_predictor_create", [_vp, _dp, ctypes.POINTER(PredParams), ctypes.POINTER(_vp)]),
    ("dnsp_predictor_field", [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t)]),
    ("dnsp_predictor_write", [_vp, _i, _vp, _vp]),
    ("dnsp_predictor_read", [_vp, _i, _vp, _vp]),
    ("dnsp_predictor_set_law", [_vp, _i, _i, _dp, _i]),
    ("dnsp_predictor_set_derived", [_vp, _i, _d, _i, _d, _d]),
    ("dnsp_predictor_set_columns", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    ("dnsp_predictor_set_component_columns", [_vp, ctypes.c_int, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    ("dnsp_predictor_set_quady", [_vp, ctypes.c_int]),
    ("dnsp_predictor_set_ibm", [_vp, _i, _vp]),
    ("dnsp_predictor_enqueue", [_vp, _vp]),
    ("dnsp_predictor_capture", [_vp, _i, _i, _vp]),
    ("dnsp_predictor_launch", [_vp, _vp]),
    ("dnsp_predictor_free", [_vp]),
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
        L.dnsp_predictor_geom_size.argtypes = [_i]
        L.dnsp_predictor_geom_size.restype = ctypes.c_size_t
        _ready = True
    return L


def exported_names():
    return [s[0] for s in _SIGS] + ["dnsp_predictor_geom_size"]


def predictor_geom(yu):
    """The y-derivative coefficients (dict as oracle/predictor.py's predictor_geom)."""
    yu = np.ascontiguousarray(yu, dtype=np.float64)
    ny = yu.size - 1
    L = _lib()
    out = np.zeros(L.dnsp_predictor_geom_size(ny))
    _check(L.dnsp_predictor_geom(ny, yu.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
    o = 0

    def take(n, cols):
        nonlocal o
        a = out[o:o + n * cols].reshape(n, cols) if cols > 1 else out[o:o + n]
        o += n * cols
        return a

    g = dict(Ruw_ddy=take(ny, 3), Ruw_d2dy2=take(ny, 3), Rv_ddy=take(ny + 1, 3), Rv_d2dy2=take(ny + 1, 3),
             Rv_interp=take(ny + 1, 2), inv_dy_u=take(ny + 1, 1), extrap=take(4, 1))
    return g


class Predictor:
    """The predictor stages of one channel case, on a ProjectionStep's fields."""

    def __init__(self, step, yu, nu, cond, walls, Sf_x=0.0, Sf_z=0.0, Sq=0.0, Gb=(0.0, 0.0, 0.0), kmin_Vhat=1):
        L = _lib()
        yu = np.ascontiguousarray(yu, dtype=np.float64)
        p = PredParams()
        p.nu, p.cond, p.Sf_x, p.Sf_z, p.Sq = float(nu), float(cond), float(Sf_x), float(Sf_z), float(Sq)
        p.Gb_x, p.Gb_y, p.Gb_z = (float(v) for v in Gb)
        keys = ("U_bot", "V_bot", "W_bot", "T_bot", "U_top", "V_top", "W_top", "T_top")
        w = [walls[k] for k in keys] if isinstance(walls, dict) else list(walls)
        for i in range(8):
            p.walls[i] = float(w[i])
        p.kmin_Vhat = int(kmin_Vhat)
        h = _vp()
        _check(L.dnsp_predictor_create(step._h, yu.ctypes.data_as(_dp), ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.step = step  # the step must outlive the predictor
        self.nx, self.nz, self.ny = step.nx, step.nz, step.ny

    def shape(self, name):
        if name in INTERIOR:
            return (self.ny, self.nz, self.nx)
        return (self.ny + 2, self.nz + 2, self.nx + 2)

    def write(self, name, t, stream=None):
        _check(_lib().dnsp_predictor_write(self._h, PRED_FIELDS.index(name), _dev(t, self.shape(name)),
                                           _vp(_stream_ptr(stream))))

    def read(self, name, stream=None):
        import torch
        out = torch.empty(self.shape(name), dtype=torch.float64, device="cuda")
        _check(_lib().dnsp_predictor_read(self._h, PRED_FIELDS.index(name), _vp(out.data_ptr()),
                                          _vp(_stream_ptr(stream))))
        return out

    def load(self, fields, stream=None):
        import torch
        for name, a in fields.items():
            self.write(name, torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda"), stream)

    def set_law(self, quantity, law):
        """law: ("const", c) | ("power", c, p) | ("cubic", [(hi, x0, c0, c1, c2, c3), ...])."""
        q = QUANTITIES.index(quantity)
        if law[0] == "const":
            data, kind, n = np.array([law[1]], dtype=np.float64), LAW_CONST, 1
        elif law[0] == "power":
            data, kind, n = np.array([law[1], law[2]], dtype=np.float64), LAW_POWER, 1
        elif law[0] == "cubic":
            segs = [[np.inf if s[0] is None else s[0], *s[1:]] for s in law[1]]
            data, kind, n = np.ascontiguousarray(segs, dtype=np.float64).ravel(), LAW_CUBIC, len(segs)
        else:
            raise ValueError(law[0])
        _check(_lib().dnsp_predictor_set_law(self._h, q, kind, data.ctypes.data_as(_dp), n))

    def set_derived(self, divjac=("const", 0.0), buoy=("none",)):
        f = 1 if divjac[0] == "formula" else 0
        c = float(divjac[1]) if divjac[0] == "const" else 0.0
        lin = 1 if buoy[0] == "linear" else 0
        b1, b0 = (float(buoy[1]), float(buoy[2])) if lin else (0.0, 0.0)
        _check(_lib().dnsp_predictor_set_derived(self._h, f, c, lin, b1, b0))

    def set_columns(self, bot, top, component=None):
        """Column slabs (nz, nx) of U, V, W and T alike, or of one component
        (0..3 or "U", "V", "W", "T"); None removes them."""
        if component is None:
            fn, args = _lib().dnsp_predictor_set_columns, ()
        else:
            k = "UVWT".index(component) if isinstance(component, str) else int(component)
            fn, args = _lib().dnsp_predictor_set_component_columns, (k,)
        if bot is None:
            _check(fn(self._h, *args, None, None))
            return
        b = np.ascontiguousarray(bot, dtype=np.int32)
        t = np.ascontiguousarray(top, dtype=np.int32)
        ip = ctypes.POINTER(_i)
        _check(fn(self._h, *args, b.ctypes.data_as(ip), t.ctypes.data_as(ip)))

    def set_quady(self, quady=True):
        """The quadratic-wall (use_rough_surf_quady) ordering of the rough-wall branch."""
        _check(_lib().dnsp_predictor_set_quady(self._h, int(bool(quady))))

    def set_ibm(self, forcings):
        """forcings: {"U"|"V"|"W"|"T"|"div": IbmForcing or None}; the handles must outlive the predictor."""
        self._ibm = dict(forcings)
        for k, name in enumerate(("U", "V", "W", "T", "div")):
            h = self._ibm.get(name)
            _check(_lib().dnsp_predictor_set_ibm(self._h, k, h._h if h is not None else _vp(None)))

    def enqueue(self, stream=None):
        _check(_lib().dnsp_predictor_enqueue(self._h, _vp(_stream_ptr(stream))))

    def capture(self, with_step=True, refinements=-1, stream=None):
        _check(_lib().dnsp_predictor_capture(self._h, int(bool(with_step)), int(refinements),
                                             _vp(_stream_ptr(stream))))

    def launch(self, stream=None):
        _check(_lib().dnsp_predictor_launch(self._h, _vp(_stream_ptr(stream))))

    def close(self):
        if getattr(self, "_h", None):
            _lib().dnsp_predictor_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

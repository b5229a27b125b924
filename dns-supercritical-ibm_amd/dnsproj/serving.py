"""Serving engine over libdnsproj.so (include/dnsproj.h: dnsp_server_*).

Projection-step requests for an ensemble of channel cases, continuously
batched on one GPU: cases become resident on first admission within an HBM
budget (least-recently-used idle cases evicted), requests run on a pool of
stream slots, and each request's inputs depend only on (case seed, request
seed), so its result is the same alone and in the mix.
"""
import ctypes

import numpy as np

from . import _check, _dp, _ip, _vp, lib

_i, _d = ctypes.c_int, ctypes.c_double


class CaseDesc(ctypes.Structure):
    _fields_ = [("nx", _i), ("nz", _i), ("ny", _i), ("yu", _dp), ("inv_dx", _d), ("inv_dz", _d),
                ("dtmax", _d), ("incompressible", _i), ("walls", _d * 6), ("n_ibm", _i * 3),
                ("ibm_ijk", _ip * 3), ("ibm_c", _dp * 3), ("bot", _ip), ("top", _ip),
                ("seed", ctypes.c_ulonglong)]


_SIGS = [
    ("dnsp_server_create", [ctypes.c_size_t, _i, _i, _i, ctypes.POINTER(_vp)]),
    ("dnsp_server_add_case", [_vp, ctypes.POINTER(CaseDesc), ctypes.POINTER(_i)]),
    ("dnsp_server_case_footprint", [_vp, _i, ctypes.POINTER(ctypes.c_size_t)]),
    ("dnsp_server_case_form", [_vp, _i, ctypes.POINTER(_i), _dp, _dp]),
    ("dnsp_server_run", [_vp, _i, _ip, _dp, ctypes.POINTER(ctypes.c_ulonglong), _dp, _dp, _dp, _dp]),
    ("dnsp_server_free", [_vp]),
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


PICK_FORM = -2  # DNSP_SERVER_PICK_FORM


class Server:
    """hbm_budget in bytes; refinements -1 (FP64), 0/1 (mixed, precision 1) or
    PICK_FORM (precision 1: per case the faster of FP64 and mixed with one
    refinement, timed at its first admission)."""

    def __init__(self, hbm_budget, max_slots=4, precision=1, refinements=1):
        h = _vp()
        _check(_lib().dnsp_server_create(int(hbm_budget), int(max_slots), int(precision), int(refinements),
                                         ctypes.byref(h)))
        self._h = h
        self._keep = []

    def add_case(self, c, seed):
        """c: a dict as workloads.projection_case makes (fields are not used:
        request inputs are generated on the device)."""
        keep = []
        d = CaseDesc()
        d.nx, d.nz, d.ny = int(c["nx"]), int(c["nz"]), int(c["ny"])
        yu = np.ascontiguousarray(c["yu"], dtype=np.float64)
        keep.append(yu)
        d.yu = yu.ctypes.data_as(_dp)
        d.inv_dx, d.inv_dz, d.dtmax = float(c["inv_dx"]), float(c["inv_dz"]), float(c["dtmax"])
        d.incompressible = int(bool(c.get("incompressible", False)))
        for k, v in enumerate(np.asarray(c["walls"], dtype=np.float64)):
            d.walls[k] = float(v)
        ibm = c.get("ibm")
        for k, comp in enumerate(("U", "V", "W")):
            if ibm is None:
                d.n_ibm[k] = 0
                continue
            ijk = np.ascontiguousarray(ibm[comp][0], dtype=np.int32)
            cc = np.ascontiguousarray(ibm[comp][1], dtype=np.float64)
            keep += [ijk, cc]
            d.n_ibm[k] = ijk.shape[0]
            d.ibm_ijk[k] = ijk.ctypes.data_as(_ip)
            d.ibm_c[k] = cc.ctypes.data_as(_dp)
        if c.get("bot") is not None:
            b = np.ascontiguousarray(c["bot"], dtype=np.int32)
            t = np.ascontiguousarray(c["top"], dtype=np.int32)
            keep += [b, t]
            d.bot, d.top = b.ctypes.data_as(_ip), t.ctypes.data_as(_ip)
        d.seed = int(seed)
        cid = _i()
        _check(_lib().dnsp_server_add_case(self._h, ctypes.byref(d), ctypes.byref(cid)))
        self._keep.append(keep)
        return cid.value

    def footprint(self, case_id):
        v = ctypes.c_size_t()
        _check(_lib().dnsp_server_case_footprint(self._h, int(case_id), ctypes.byref(v)))
        return v.value

    def case_form(self, case_id):
        """(refinements of the case's form, FP64 step ms, mixed step ms); the
        times are -1 unless the server picks forms and the case was admitted."""
        r, a, b = _i(), _d(), _d()
        _check(_lib().dnsp_server_case_form(self._h, int(case_id), ctypes.byref(r), ctypes.byref(a),
                                            ctypes.byref(b)))
        return r.value, a.value, b.value

    def run(self, case_ids, arrivals_s, seeds):
        """Replays the trace; returns dict(latency_ms, start_ms, checksum (n, 4), stats)."""
        n = len(case_ids)
        cid = np.ascontiguousarray(case_ids, dtype=np.int32)
        arr = np.ascontiguousarray(arrivals_s, dtype=np.float64)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        lat, start = np.zeros(n), np.zeros(n)
        cks = np.zeros((n, 4))
        stats = np.zeros(4)
        _check(_lib().dnsp_server_run(self._h, n, cid.ctypes.data_as(_ip), arr.ctypes.data_as(_dp),
                                      sd.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)),
                                      lat.ctypes.data_as(_dp), start.ctypes.data_as(_dp),
                                      cks.ctypes.data_as(_dp), stats.ctypes.data_as(_dp)))
        return {"latency_ms": lat, "start_ms": start, "checksum": cks,
                "stats": {"admissions": int(stats[0]), "evictions": int(stats[1]),
                          "peak_resident_bytes": float(stats[2]), "wall_s": float(stats[3])}}

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib().dnsp_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

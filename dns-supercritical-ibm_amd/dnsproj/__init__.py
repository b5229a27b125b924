"""Host-side mirror of the reference's pressure-Poisson interface over libdnsproj.so.

The reference's Fortran module ``poisson_solver_multigpu`` exposes
``init_spec_poisson_multigpu`` / ``run_spec_poisson_multigpu``
(src/poisson_solver_mgpu_dtdma/poisson_solver_multigpu.f90:33-40, :328-335) and
``get_lo_n_bounds_1d`` (:556-575); its band generator is ``mk_P_abc``
(src/pycode/calc_bands.py:181-201).  This package exposes the same names with
the same argument meaning, calling the HIP library through its C ABI
(include/dnsproj.h).  Device arrays are torch tensors (float64, contiguous,
on a HIP device); torch is only the allocator and stream provider here.

There is no CPU fallback: when the library is missing every call raises.
"""
import ctypes
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libdnsproj.so")

DNSP_OK = 0
DNSP_ERR_ARG = 1
DNSP_ERR_HIP = 2
DNSP_ERR_NOMEM = 3
DNSP_ERR_SINGULAR = 4
DNSP_ERR_UNSUPPORTED = 5
DNSP_ERR_NOT_CONVERGED = 6
DNSP_ERR_NONFINITE = 7

FFT_AUTO, FFT_GENERIC, FFT_FIXED = 0, 1, 2
TRIDIAG_AUTO, TRIDIAG_SERIAL, TRIDIAG_PARTITION = 0, 1, 2

PRECISION_FP64, PRECISION_MIXED = 0, 1
MAX_REFINEMENTS = 8
TUNE_SLOTS = 32

STAGE_X_FWD, STAGE_Z_FWD, STAGE_TRIDIAG, STAGE_Z_BWD, STAGE_X_BWD = range(5)
STAGE_NAMES = ("x_fwd", "z_fwd", "tridiag", "z_bwd", "x_bwd")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

_lib = None


class DnspError(RuntimeError):
    """A nonzero status from the C ABI; ``code`` is the DNSP_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"dnsproj error {code}: {msg}")
        self.code = code


# (name, argtypes, restype) of every exported entry point
_SIGNATURES = [
    ("dnsp_last_error", [], ctypes.c_char_p),
    ("dnsp_version", [], ctypes.c_char_p),
    ("dnsp_stream_synchronize", [_vp], ctypes.c_int),
    ("dnsp_poisson_init", [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                           _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    ("dnsp_poisson_run", [_vp, _vp, _vp], ctypes.c_int),
    ("dnsp_poisson_solve", [_vp, _vp, _vp, _vp], ctypes.c_int),
    ("dnsp_poisson_run_host", [_vp, _dp], ctypes.c_int),
    ("dnsp_poisson_forward", [_vp, _vp, _vp], ctypes.c_int),
    ("dnsp_poisson_tridiag", [_vp, _vp, _vp], ctypes.c_int),
    ("dnsp_poisson_backward", [_vp, _vp, _vp], ctypes.c_int),
    ("dnsp_poisson_stage", [_vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    ("dnsp_poisson_set_strategy", [_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("dnsp_poisson_get_strategy", [_vp, _ip, _ip, _ip], ctypes.c_int),
    ("dnsp_poisson_set_variants", [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("dnsp_poisson_get_variants", [_vp, _ip, _ip, _ip], ctypes.c_int),
    ("dnsp_poisson_set_mixed_variants", [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("dnsp_poisson_get_mixed_variants", [_vp, _ip, _ip, _ip], ctypes.c_int),
    ("dnsp_poisson_autotune", [_vp, _vp, ctypes.c_int, _vp, _dp], ctypes.c_int),
    ("dnsp_poisson_solve_mixed", [_vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    ("dnsp_poisson_set_form", [_vp, ctypes.c_int], ctypes.c_int),
    ("dnsp_poisson_run_mixed", [_vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    ("dnsp_poisson_solve_refine", [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _ip, _dp, _vp], ctypes.c_int),
    ("dnsp_poisson_residual", [_vp, _vp, _vp, _dp, _vp], ctypes.c_int),
    ("dnsp_poisson_solve_stage", [_vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    ("dnsp_poisson_set_blocked", [_vp, ctypes.c_int], ctypes.c_int),
    ("dnsp_poisson_get_blocked", [_vp, _ip, _ip], ctypes.c_int),
    ("dnsp_poisson_get_solve_variants", [_vp, _ip], ctypes.c_int),
    ("dnsp_poisson_set_mixed_fp32", [_vp, ctypes.c_int], ctypes.c_int),
    ("dnsp_poisson_mixed_info", [_vp, _ip, _dp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    ("dnsp_poisson_residual_noalloc", [_vp, _vp, _vp, _dp, _vp], ctypes.c_int),
    ("dnsp_poisson_check", [_vp, _vp, ctypes.POINTER(ctypes.c_longlong), _vp], ctypes.c_int),
    ("dnsp_batch_create", [ctypes.c_int, _ip, _ip, _ip, _dp, _dp, _dp, ctypes.c_int, ctypes.POINTER(_vp)],
     ctypes.c_int),
    ("dnsp_batch_bind", [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)], ctypes.c_int),
    ("dnsp_batch_run", [_vp, _vp], ctypes.c_int),
    ("dnsp_batch_run_persistent", [_vp, _vp], ctypes.c_int),
    ("dnsp_batch_run_mixed", [_vp, ctypes.c_int, _vp], ctypes.c_int),
    ("dnsp_batch_autotune", [_vp, ctypes.c_int, _vp, _dp], ctypes.c_int),
    ("dnsp_batch_set_variants", [_vp] + [ctypes.c_int] * 6, ctypes.c_int),
    ("dnsp_batch_get_variants", [_vp, _ip], ctypes.c_int),
    ("dnsp_batch_size", [_vp, _ip], ctypes.c_int),
    ("dnsp_batch_free", [_vp], ctypes.c_int),
    ("dnsp_poisson_min_pivot", [_vp, _dp], ctypes.c_int),
    ("dnsp_poisson_shape", [_vp, _ip, _ip, _ip], ctypes.c_int),
    ("dnsp_poisson_free", [_vp], ctypes.c_int),
    ("dnsp_get_lo_n_bounds_1d", [ctypes.c_int, ctypes.c_int, _ip, _ip], ctypes.c_int),
    ("dnsp_mk_P_abc", [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
]


def lib():
    """Loads libdnsproj.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {_PKG_ROOT}` "
                               "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args, res in _SIGNATURES:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_names():
    return [s[0] for s in _SIGNATURES]


def _check(rc):
    if rc != DNSP_OK:
        raise DnspError(rc, lib().dnsp_last_error().decode(errors="replace"))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def mk_P_abc(yu, is_dirich_first=False):
    """Poisson y bands for the faces ``yu`` (length ny+1): returns (a, b, c).

    calc_bands.py:181-201 takes ``yp_ext`` = [yu0, cell centres, yu_ny]
    (:226-227); this takes the faces and forms yp_ext itself.
    """
    yu, pyu = _f64(yu)
    ny = yu.size - 1
    a, b, c = (np.empty(ny) for _ in range(3))
    _check(lib().dnsp_mk_P_abc(ny, pyu, int(bool(is_dirich_first)), a.ctypes.data_as(_dp),
                               b.ctypes.data_as(_dp), c.ctypes.data_as(_dp)))
    return a, b, c


def get_lo_n_bounds_1d(mpi_divs, n_glob):
    lo = np.zeros(mpi_divs, dtype=np.int32)
    n = np.zeros(mpi_divs, dtype=np.int32)
    _check(lib().dnsp_get_lo_n_bounds_1d(int(mpi_divs), int(n_glob), lo.ctypes.data_as(_ip),
                                         n.ctypes.data_as(_ip)))
    return lo, n


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _device_ptr(P, numel):
    import torch
    if not isinstance(P, torch.Tensor):
        raise TypeError("P must be a torch tensor on a HIP device")
    if P.dtype != torch.float64 or not P.is_cuda or not P.is_contiguous():
        raise ValueError("P must be a contiguous float64 tensor on a HIP device")
    if P.numel() != numel:
        raise ValueError(f"P has {P.numel()} elements, the solver expects {numel}")
    return P.data_ptr()


class PoissonSolver:
    """One single-rank Poisson solver instance (the C handle).

    ``bands`` = (P_band_a, P_band_b, P_band_c, P_band_a_00, P_band_b_00, P_band_c_00).
    Arrays are (ny, nz, nx) C-order tensors, i.e. the reference's x-fastest P.
    ``precision=PRECISION_MIXED`` adds the mixed-precision form
    (``solve_mixed``, ``solve_refine``) to the FP64 one.
    """

    def __init__(self, nx, nz, ny, inv_dx, inv_dz, bands, precision=0):
        L = lib()
        bs = [_f64(b) for b in bands]
        if len(bs) != 6 or any(b[0].size != ny for b in bs):
            raise ValueError("six band arrays of length ny are required")
        h = _vp()
        _check(L.dnsp_poisson_init(int(nx), int(nz), int(ny), float(inv_dx), float(inv_dz),
                                   *[b[1] for b in bs], int(precision), ctypes.byref(h)))
        self._h = h
        self._bands = bs  # keep alive
        self.nx, self.nz, self.ny = int(nx), int(nz), int(ny)
        self.inv_dx, self.inv_dz = float(inv_dx), float(inv_dz)
        self.precision = int(precision)

    @property
    def shape(self):
        return (self.ny, self.nz, self.nx)

    @property
    def numel(self):
        return self.nx * self.nz * self.ny

    def set_strategy(self, fft=FFT_AUTO, tridiag=TRIDIAG_AUTO):
        _check(lib().dnsp_poisson_set_strategy(self._h, int(fft), int(tridiag)))

    @property
    def strategy(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().dnsp_poisson_get_strategy(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"fft_ct_x": a.value, "fft_ct_z": b.value, "part_rows": c.value}

    def set_variants(self, xvar, zvar, tvar=0):
        _check(lib().dnsp_poisson_set_variants(self._h, int(xvar), int(zvar), int(tvar)))

    @property
    def variants(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().dnsp_poisson_get_variants(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return (a.value, b.value, c.value)

    def set_mixed_variants(self, xvar, zvar, tvar=0):
        _check(lib().dnsp_poisson_set_mixed_variants(self._h, int(xvar), int(zvar), int(tvar)))

    @property
    def mixed_variants(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().dnsp_poisson_get_mixed_variants(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return (a.value, b.value, c.value)

    def solve_stage(self, rhs, P, stage, stream=None):
        """Stage 0..4 of solve(rhs, P) (x fwd, z fwd, tridiagonal, z inv, x inv)."""
        _check(lib().dnsp_poisson_solve_stage(self._h, _vp(_device_ptr(rhs, self.numel)),
                                              _vp(_device_ptr(P, self.numel)), int(stage),
                                              _vp(_stream_ptr(stream))))
        return P

    def set_blocked(self, on=True):
        """The blocked solve (x-blocked / mode-major layouts through a workspace)
        where supported (default on), or the in-place stages."""
        _check(lib().dnsp_poisson_set_blocked(self._h, int(bool(on))))

    @property
    def blocked(self):
        """(supported, in use)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().dnsp_poisson_get_blocked(self._h, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), bool(b.value)

    @property
    def solve_variants(self):
        """Variants of the five kernels solve() runs (stage order)."""
        v = (ctypes.c_int * 5)()
        _check(lib().dnsp_poisson_get_solve_variants(self._h, v))
        return tuple(v)

    def set_mixed_fp32(self, on=True):
        """FP32 arithmetic in the mixed sweeps' well-conditioned tridiagonal
        solves (default on); off: FP64 arithmetic for every mode."""
        _check(lib().dnsp_poisson_set_mixed_fp32(self._h, int(bool(on))))

    @property
    def mixed_info(self):
        """(fp32 sweeps on, |lambda| from which a mode runs in FP32, modes that do)."""
        on, lam, n = ctypes.c_int(), ctypes.c_double(), ctypes.c_longlong()
        _check(lib().dnsp_poisson_mixed_info(self._h, ctypes.byref(on), ctypes.byref(lam), ctypes.byref(n)))
        return bool(on.value), lam.value, n.value

    def autotune(self, work=None, reps=5, stream=None):
        """Picks the fastest kernel variants; returns the timing table (ms per pair)."""
        times = np.zeros(TUNE_SLOTS)
        wp = _vp(_device_ptr(work, self.numel)) if work is not None else _vp(None)
        _check(lib().dnsp_poisson_autotune(self._h, wp, int(reps), _vp(_stream_ptr(stream)),
                                           times.ctypes.data_as(_dp)))
        out = {"x_ms": times[:4].tolist(), "z_ms": times[4:8].tolist(), "tri_ms": times[8:16].tolist(),
               "picked": self.variants}
        if self.precision == PRECISION_MIXED:
            out.update({"mixed_x_ms": times[16:20].tolist(), "mixed_z_ms": times[20:24].tolist(),
                        "mixed_tri_ms": times[24:32].tolist(), "mixed_picked": self.mixed_variants})
        return out

    def solve_mixed(self, rhs, P, refinements=1, stream=None):
        """Mixed-precision solve (FP32 transforms + FP64 refinement): rhs -> P (rhs may be P)."""
        if self._h is None:
            raise RuntimeError("solver has been freed")
        _check(lib().dnsp_poisson_solve_mixed(self._h, _vp(_device_ptr(rhs, self.numel)),
                                              _vp(_device_ptr(P, self.numel)), int(refinements),
                                              _vp(_stream_ptr(stream))))
        return P

    def set_form(self, refinements=-1):
        """The form run/run_host use: -1 FP64, >= 0 the mixed form with that many refinements."""
        _check(lib().dnsp_poisson_set_form(self._h, int(refinements)))

    def run_mixed(self, P, refinements=1, stream=None):
        return self.solve_mixed(P, P, refinements, stream)

    def solve_refine(self, rhs, P, tol=1e-10, max_refinements=4, stream=None):
        """Refines until ||rhs - L P||/||rhs|| <= tol; returns (refinements done, relative residual).
        Raises DnspError with code DNSP_ERR_NOT_CONVERGED when tol is not met
        after max_refinements, DNSP_ERR_NONFINITE for a NaN or Inf residual."""
        done, rel = ctypes.c_int(), ctypes.c_double()
        _check(lib().dnsp_poisson_solve_refine(self._h, _vp(_device_ptr(rhs, self.numel)),
                                               _vp(_device_ptr(P, self.numel)), float(tol),
                                               int(max_refinements), ctypes.byref(done), ctypes.byref(rel),
                                               _vp(_stream_ptr(stream))))
        return done.value, rel.value

    def residual(self, rhs, P, stream=None):
        """||rhs - L P|| / ||rhs|| of the discrete operator (FP64, on the device)."""
        rel = ctypes.c_double()
        _check(lib().dnsp_poisson_residual(self._h, _vp(_device_ptr(rhs, self.numel)),
                                           _vp(_device_ptr(P, self.numel)), ctypes.byref(rel),
                                           _vp(_stream_ptr(stream))))
        return rel.value

    def residual_noalloc(self, rhs, P, stream=None):
        """As residual(), through the entry point that allocates nothing (its
        scratch is a static device array of the library, shared by every call
        on the GPU: overlapping calls race)."""
        rel = ctypes.c_double()
        _check(lib().dnsp_poisson_residual_noalloc(self._h, _vp(_device_ptr(rhs, self.numel)),
                                                   _vp(_device_ptr(P, self.numel)), ctypes.byref(rel),
                                                   _vp(_stream_ptr(stream))))
        return rel.value

    def check(self, P, stream=None):
        """Number of NaN/Inf values in the device solution P (0, or raises
        DnspError with code DNSP_ERR_NONFINITE); synchronises the stream."""
        cnt = ctypes.c_longlong()
        _check(lib().dnsp_poisson_check(self._h, _vp(_device_ptr(P, self.numel)), ctypes.byref(cnt),
                                        _vp(_stream_ptr(stream))))
        return cnt.value

    @property
    def min_pivot(self):
        out = ctypes.c_double()
        _check(lib().dnsp_poisson_min_pivot(self._h, ctypes.byref(out)))
        return out.value

    def _call(self, fn, P, stream):
        if self._h is None:
            raise RuntimeError("solver has been freed")
        _check(fn(self._h, _vp(_device_ptr(P, self.numel)), _vp(_stream_ptr(stream))))
        return P

    def run(self, P, stream=None):
        """In-place solve of the device array P (RHS in, solution out)."""
        return self._call(lib().dnsp_poisson_run, P, stream)

    def solve(self, rhs, P, stream=None):
        """Out-of-place solve: rhs (device, unchanged) -> P (device)."""
        if self._h is None:
            raise RuntimeError("solver has been freed")
        _check(lib().dnsp_poisson_solve(self._h, _vp(_device_ptr(rhs, self.numel)),
                                        _vp(_device_ptr(P, self.numel)), _vp(_stream_ptr(stream))))
        return P

    def forward(self, P, stream=None):
        return self._call(lib().dnsp_poisson_forward, P, stream)

    def tridiag(self, S, stream=None):
        return self._call(lib().dnsp_poisson_tridiag, S, stream)

    def backward(self, S, stream=None):
        return self._call(lib().dnsp_poisson_backward, S, stream)

    def stage(self, P, stage, stream=None):
        """Launches one kernel of the solve in place (STAGE_* constants)."""
        return self.stage_oop(P, P, stage, stream)

    def stage_oop(self, src, P, stage, stream=None):
        """One kernel reading src and writing P (src != P only for the x stages)."""
        if self._h is None:
            raise RuntimeError("solver has been freed")
        _check(lib().dnsp_poisson_stage(self._h, _vp(_device_ptr(src, self.numel)),
                                        _vp(_device_ptr(P, self.numel)), int(stage), _vp(_stream_ptr(stream))))
        return P

    def run_host(self, P):
        """Solve a host numpy array in place (copies to and from the device)."""
        P = np.asarray(P)
        if P.dtype != np.float64 or not P.flags["C_CONTIGUOUS"] or P.size != self.numel:
            raise ValueError("P must be a C-contiguous float64 array of nx*nz*ny elements")
        _check(lib().dnsp_poisson_run_host(self._h, P.ctypes.data_as(_dp)))
        return P

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().dnsp_poisson_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PoissonBatch:
    """Many independent solves of uneven sizes in one batched launch sequence.

    ``cases``: sequence of dicts with nx, nz, ny, inv_dx, inv_dz and bands (the
    six band arrays).  ``bind`` fixes the device arrays (torch float64
    tensors, (ny, nz, nx) each); ``run`` / ``run_mixed`` then only enqueue
    kernels, so they can be captured in a graph.
    """

    def __init__(self, cases, precision=PRECISION_FP64):
        L = lib()
        n = len(cases)
        if n < 1:
            raise ValueError("at least one instance")
        self.shapes = [(int(c["ny"]), int(c["nz"]), int(c["nx"])) for c in cases]
        ar = lambda key, dt: np.ascontiguousarray([c[key] for c in cases], dtype=dt)
        nx, nz, ny = ar("nx", np.int32), ar("nz", np.int32), ar("ny", np.int32)
        idx, idz = ar("inv_dx", np.float64), ar("inv_dz", np.float64)
        bands = []
        for c in cases:
            bs = [np.ascontiguousarray(b, dtype=np.float64) for b in c["bands"]]
            if len(bs) != 6 or any(b.size != int(c["ny"]) for b in bs):
                raise ValueError("six band arrays of length ny per instance")
            bands.extend(bs)
        bcat = np.ascontiguousarray(np.concatenate(bands))
        h = _vp()
        _check(L.dnsp_batch_create(n, nx.ctypes.data_as(_ip), nz.ctypes.data_as(_ip), ny.ctypes.data_as(_ip),
                                   idx.ctypes.data_as(_dp), idz.ctypes.data_as(_dp), bcat.ctypes.data_as(_dp),
                                   int(precision), ctypes.byref(h)))
        self._h = h
        self.n = n
        self.precision = int(precision)
        self._bound = None

    def bind(self, rhs, P):
        if len(rhs) != self.n or len(P) != self.n:
            raise ValueError(f"{self.n} rhs and P arrays required")
        pr = (_vp * self.n)(*[_device_ptr(r, int(np.prod(s))) for r, s in zip(rhs, self.shapes)])
        pp = (_vp * self.n)(*[_device_ptr(p, int(np.prod(s))) for p, s in zip(P, self.shapes)])
        _check(lib().dnsp_batch_bind(self._h, pr, pp))
        self._bound = (list(rhs), list(P))  # keep the tensors alive

    def run(self, stream=None):
        _check(lib().dnsp_batch_run(self._h, _vp(_stream_ptr(stream))))

    def run_persistent(self, stream=None):
        """The FP64 solve of every instance in one cooperative launch (x and z
        variant 2, two-pass tridiagonal), results as run() with those
        variants."""
        _check(lib().dnsp_batch_run_persistent(self._h, _vp(_stream_ptr(stream))))

    def run_mixed(self, refinements=1, stream=None):
        _check(lib().dnsp_batch_run_mixed(self._h, int(refinements), _vp(_stream_ptr(stream))))

    def autotune(self, reps=3, stream=None):
        times = np.zeros(TUNE_SLOTS)
        _check(lib().dnsp_batch_autotune(self._h, int(reps), _vp(_stream_ptr(stream)), times.ctypes.data_as(_dp)))
        # tridiagonal: variants 0..3 (32 chunks) and 4, 5 (16 chunks)
        out = {"x_ms": times[:4].tolist(), "z_ms": times[4:8].tolist(), "tri_ms": times[8:14].tolist()}
        if self.precision == PRECISION_MIXED:
            out.update({"mixed_x_ms": times[16:20].tolist(), "mixed_z_ms": times[20:24].tolist(),
                        "mixed_tri_ms": times[24:30].tolist()})
        out["picked"] = self.variants
        return out

    @property
    def variants(self):
        v = np.zeros(6, dtype=np.int32)
        _check(lib().dnsp_batch_get_variants(self._h, v.ctypes.data_as(_ip)))
        return tuple(int(x) for x in v)

    def set_variants(self, xvar, zvar, tvar, mxvar=0, mzvar=0, mtvar=1):
        _check(lib().dnsp_batch_set_variants(self._h, int(xvar), int(zvar), int(tvar), int(mxvar), int(mzvar),
                                             int(mtvar)))

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().dnsp_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_spec_poisson_multigpu(P_band_a, P_band_b, P_band_c, P_band_a_00, P_band_b_00, P_band_c_00,
                               nx_glob, nz_glob, ny_loc, inv_dx, inv_dz):
    """Single-rank init (poisson_solver_multigpu.f90:33-326): returns the solver handle."""
    return PoissonSolver(nx_glob, nz_glob, ny_loc, inv_dx, inv_dz,
                         (P_band_a, P_band_b, P_band_c, P_band_a_00, P_band_b_00, P_band_c_00))


def run_spec_poisson_multigpu(solver, P, stream=None):
    """Single-rank solve (poisson_solver_multigpu.f90:328-554): P (device) RHS in, solution out."""
    return solver.run(P, stream)

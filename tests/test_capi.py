"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/dnsproj.h, and its host-only helpers agree with the oracle.
Argument validation that happens before any HIP call is checked too."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import dnsproj as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "dnsproj.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dnsp_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = D.lib()
    names = header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    import dnsproj.projection as DP
    import dnsproj.serving as SV
    import dnsproj.io as IO
    import dnsproj.predictor as PR
    bound = (set(D.exported_names()) | set(DP.exported_names()) | set(SV.exported_names()) |
             set(IO.exported_names()) | set(PR.exported_names()))
    assert set(names) == bound, (set(names) - bound, bound - set(names))


def test_version_and_error_strings():
    L = D.lib()
    assert b"gfx950" in L.dnsp_version()
    assert isinstance(L.dnsp_last_error(), bytes)


@pytest.mark.parametrize("ny", [2, 3, 7, 64, 384])
def test_mk_P_abc_bitexact_vs_oracle(ny):
    rng = np.random.default_rng(ny)
    yu = np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 0.2, size=ny))]) + rng.uniform(-0.2, 0.2)
    for dirich in (False, True):
        got = D.mk_P_abc(yu, dirich)
        ref = O.mk_P_abc(O.yp_ext_from_yu(yu), dirich)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)


def test_mk_P_abc_rejects_coincident_faces():
    yu = np.array([0.0, 1.0, 1.0, 2.0])
    yu[2] = yu[1]
    # coincident centres are fine, coincident faces give a zero width only when
    # the mirrored ghost coincides; force it with a degenerate first cell
    yu = np.array([0.0, 0.0, 1.0])
    with pytest.raises(D.DnspError) as e:
        D.mk_P_abc(yu)
    assert e.value.code == D.DNSP_ERR_ARG


@pytest.mark.parametrize("divs,n", [(1, 7), (3, 10), (4, 4), (5, 3), (8, 386), (7, 0)])
def test_get_lo_n_bounds_1d_vs_oracle(divs, n):
    lo, cnt = D.get_lo_n_bounds_1d(divs, n)
    lo2, cnt2 = O.get_lo_n_bounds_1d(divs, n)
    assert np.array_equal(lo, lo2) and np.array_equal(cnt, cnt2)


def test_init_rejects_bad_sizes_before_touching_the_device():
    L = D.lib()
    ny = 1
    b = np.zeros(4)
    p = b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    h = ctypes.c_void_p()
    rc = L.dnsp_poisson_init(8, 8, ny, 1.0, 1.0, p, p, p, p, p, p, 0, ctypes.byref(h))
    assert rc == D.DNSP_ERR_ARG and b"ny" in L.dnsp_last_error()
    rc = L.dnsp_poisson_init(0, 8, 4, 1.0, 1.0, p, p, p, p, p, p, 0, ctypes.byref(h))
    assert rc == D.DNSP_ERR_ARG
    bad = np.array([0.0, np.nan, 0.0, 0.0])
    q = bad.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = L.dnsp_poisson_init(8, 8, 4, 1.0, 1.0, p, q, p, p, p, p, 0, ctypes.byref(h))
    assert rc == D.DNSP_ERR_ARG and b"finite" in L.dnsp_last_error()
    rc = L.dnsp_poisson_init(8, 8, 4, 1.0, 1.0, p, p, p, p, p, p, 7, ctypes.byref(h))
    assert rc == D.DNSP_ERR_ARG
    assert not h.value


def test_null_handle_calls_fail_cleanly():
    L = D.lib()
    assert L.dnsp_poisson_run(None, None, None) == D.DNSP_ERR_ARG
    out = ctypes.c_double()
    assert L.dnsp_poisson_min_pivot(None, ctypes.byref(out)) == D.DNSP_ERR_ARG
    assert L.dnsp_poisson_free(None) == 0
    assert L.dnsp_poisson_residual_noalloc(None, None, None, ctypes.byref(out), None) == D.DNSP_ERR_ARG
    assert L.dnsp_poisson_set_mixed_fp32(None, 1) == D.DNSP_ERR_ARG
    assert L.dnsp_poisson_mixed_info(None, None, None, None) == D.DNSP_ERR_ARG

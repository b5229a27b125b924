"""GPU checks of the blocked FP64 solve (x-blocked and mode-major layouts
through the handle's workspace).  It runs the same transforms and the same
per-mode arithmetic as the in-place stages; where its z kernels hold a
different number of columns per workgroup than the in-place ones the
compiler may fuse a different product of a complex multiply into a
multiply-add, so the two agree to 1e-13 (relative to the max-abs), and the
blocked solve run out of place, in place and stage by stage agrees with
itself bit for bit.  Against the oracle it keeps the FP64 bar (1e-12
relative, mean removed)."""
import sys

import numpy as np
import pytest

import oracle as O
import dnsproj as D
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def case(nx, nz, ny, seed):
    rng = np.random.default_rng(seed)
    yu = W.tanh_faces(ny, float(rng.uniform(1.2, 2.2)))
    return dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / W.LX, inv_dz=nz / W.LZ, yu=yu,
                rhs=W.rhs_uniform((ny, nz, nx), seed, yu))


def rel_mean_removed(got, ref):
    g = got - got.mean()
    r = ref - ref.mean()
    return np.abs(g - r).max() / max(np.abs(r).max(), 1e-300)


BLOCKED_CASES = [(64, 32, 64), (32, 96, 40), (128, 64, 33), (192, 128, 48), (256, 32, 100), (768, 32, 36),
                 (96, 768, 32), (384, 192, 512), (512, 96, 70), (128, 64, 1024)]


@pytest.mark.parametrize("nx,nz,ny", BLOCKED_CASES)
def test_blocked_equals_in_place_and_oracle(cuda, nx, nz, ny):
    say("blocked", nx, nz, ny)
    c = case(nx, nz, ny, nx + 7 * nz + ny)
    bands = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
    s = D.PoissonSolver(nx, nz, ny, c["inv_dx"], c["inv_dz"], bands)
    assert s.blocked == (True, True)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    f = to_dev(c["rhs"])
    for xv, zv, tv in ((0, 1, 4), (2, 0, 0), (3, 3, 6)):
        s.set_variants(xv, zv, tv)
        P = to_dev(np.zeros_like(c["rhs"]))
        s.solve(f, P)
        got = to_host(P)
        if ny <= 512:
            assert rel_mean_removed(got, ref) <= 1e-12, (xv, zv, tv)
        else:
            # 1024 rows on a strongly stretched grid: the partitioned
            # tridiagonal's bar there is the residual (as for the in-place
            # stages of test_every_tridiag_variant_tall_columns)
            r = O.apply_operator(got, c["inv_dx"], c["inv_dz"], bands) - c["rhs"]
            assert np.linalg.norm(r) / np.linalg.norm(c["rhs"]) <= 1e-10, (xv, zv, tv)
        s.set_blocked(False)
        assert s.blocked == (True, False)
        Q = to_dev(np.zeros_like(c["rhs"]))
        s.solve(f, Q)
        q = to_host(Q)
        assert np.abs(q - got).max() <= 1e-13 * np.abs(q).max(), (xv, zv, tv)
        s.set_blocked(True)
        # in place, and stage by stage
        R = to_dev(c["rhs"])
        s.run(R)
        assert np.array_equal(to_host(R), got), (xv, zv, tv)
        T = to_dev(np.zeros_like(c["rhs"]))
        for st in range(5):
            s.solve_stage(f, T, st)
        assert np.array_equal(to_host(T), got), (xv, zv, tv)
    assert np.array_equal(to_host(f), c["rhs"])  # rhs untouched
    s.close()


def test_blocked_support_and_tuner(cuda):
    """Sizes outside the blocked solve (nx not a multiple of 16, runtime-sized
    transforms, the serial Thomas for ny < 32) report it unsupported and still
    solve; the tuner leaves the blocked solve's result unchanged in accuracy."""
    say("blocked support")
    for nx, nz, ny in ((96 + 8, 32, 40), (40, 32, 40), (64, 32, 16)):
        c = case(nx, nz, ny, 3)
        bands = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
        s = D.PoissonSolver(nx, nz, ny, c["inv_dx"], c["inv_dz"], bands)
        assert s.blocked == (False, False), (nx, nz, ny)
        P = to_dev(c["rhs"])
        s.run(P)
        ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
        assert rel_mean_removed(to_host(P), ref) <= 1e-12
        s.close()
    c = case(256, 128, 96, 5)
    bands = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
    s = D.PoissonSolver(256, 128, 96, c["inv_dx"], c["inv_dz"], bands)
    s.autotune(reps=2)
    P = to_dev(c["rhs"])
    s.run(P)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    assert rel_mean_removed(to_host(P), ref) <= 1e-12
    s.close()

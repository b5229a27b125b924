"""GPU check of dnsp_poisson_residual_noalloc, the residual entry point whose
scratch is a static device array of the library: through the C ABI it must
give dnsp_poisson_residual's result bit for bit, on two calls in a row (the
second on a smaller grid, so stale scratch of the first would show), and
reject grids beyond its scratch."""
import sys

import numpy as np
import pytest

import dnsproj as D
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def solver(nx, nz, ny, gamma=1.6):
    yu = W.tanh_faces(ny, gamma)
    bands = D.mk_P_abc(yu, False) + D.mk_P_abc(yu, True)
    return D.PoissonSolver(nx, nz, ny, nx / W.LX, nz / W.LZ, bands)


def test_noalloc_residual_matches_allocating_one(cuda):
    import torch
    got = []
    for nx, nz, ny, seed in ((96, 64, 40, 1), (64, 32, 24, 2)):
        say("noalloc residual", nx, nz, ny)
        s = solver(nx, nz, ny)
        rng = np.random.default_rng(seed)
        f = torch.from_numpy(rng.uniform(-1, 1, (ny, nz, nx))).to(cuda)
        p = torch.from_numpy(rng.uniform(-1, 1, (ny, nz, nx)) * 1e-3).to(cuda)
        # an arbitrary field and a solved one (small residual)
        a = s.residual_noalloc(f, p)
        b = s.residual(f, p)
        assert a == b, (a, b)
        x = f.clone()
        s.run(x)
        torch.cuda.synchronize()
        a2 = s.residual_noalloc(f, x)
        b2 = s.residual(f, x)
        assert a2 == b2, (a2, b2)
        assert a2 < 1e-10
        got.append((a, a2))
        s.close()
    assert got[0] != got[1]


def test_noalloc_residual_rejects_oversized_grid(cuda):
    import torch
    say("noalloc residual size limit")
    nx, nz, ny = 32, 512, 520  # nz*ny = 266240 lines > 262144
    s = solver(nx, nz, ny)
    f = torch.zeros((ny, nz, nx), dtype=torch.float64, device=cuda)
    with pytest.raises(D.DnspError) as e:
        s.residual_noalloc(f, f)
    assert e.value.code == D.DNSP_ERR_ARG
    assert "static scratch" in str(e.value)
    s.close()

"""Solves on the wall-normal grids of the reference's four example cases
(tests/golden/yu_*.txt, copied from examples/*/yu.txt): the oracle at a
reduced periodic size, and the device residual at the cases' own sizes
(512x512x288, 560x280x280, 624x312x770 with the radix-13 lengths,
576x576x624)."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
import dnsproj as D
from dnsproj import io as IO
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRIDS = json.load(open(os.path.join(GOLDEN, "reference_grids.json")))


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


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_reference_grid_reduced_size_matches_oracle(cuda, name):
    g = GRIDS[name]
    yu = IO.read_faces(os.path.join(GOLDEN, f"yu_{name}.txt"))
    ny = len(yu) - 1
    nx, nz = 64, 32
    say("reference grid", name, nx, nz, ny)
    bands = D.mk_P_abc(yu, False) + D.mk_P_abc(yu, True)
    f = W.rhs_uniform((ny, nz, nx), 11, yu)
    idx, idz = nx / g["Lx"], nz / g["Lz"]
    ref = O.solve(f, idx, idz, bands)
    s = D.PoissonSolver(nx, nz, ny, idx, idz, bands, precision=D.PRECISION_MIXED)
    P = to_dev(f)
    s.run(P)
    got = to_host(P)
    gm, rm = got - got.mean(), ref - ref.mean()
    assert np.abs(gm - rm).max() <= 1e-12 * np.abs(rm).max()
    Q = to_dev(np.zeros_like(f))
    done, rel = s.solve_refine(to_dev(f), Q, tol=1e-10, max_refinements=4)
    assert rel <= 1e-10
    r = O.apply_operator(to_host(Q), idx, idz, bands) - f
    assert np.linalg.norm(r) / np.linalg.norm(f) <= 1e-10
    s.close()


@pytest.mark.slow
@pytest.mark.parametrize("name", sorted(GRIDS))
def test_reference_grid_full_size_residual(cuda, name):
    import torch
    g = GRIDS[name]
    yu = IO.read_faces(os.path.join(GOLDEN, f"yu_{name}.txt"))
    nx, nz, ny = g["nx"], g["nz"], g["ny"]
    say("reference grid full", name, nx, nz, ny)
    bands = D.mk_P_abc(yu, False) + D.mk_P_abc(yu, True)
    s = D.PoissonSolver(nx, nz, ny, nx / g["Lx"], nz / g["Lz"], bands)
    gen = torch.Generator(device="cuda").manual_seed(5)
    f = torch.rand((ny, nz, nx), dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    f -= f.mean()  # the residual check below needs no compatibility
    P = torch.empty_like(f)
    s.solve(f, P)
    rel = s.residual(f, P)
    say("reference grid full", name, "rel residual", rel)
    assert rel <= 1e-10
    s.close()

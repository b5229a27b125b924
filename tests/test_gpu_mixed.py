"""GPU checks of the mixed-precision Poisson solve (FP32 transforms and
spectra, FP64 tridiagonal arithmetic and (0,0) mode, FP64 residual refinement)
and of the device residual norm, against the CPU oracle.

Tolerances (relative residual = ||rhs - L p||_2 / ||rhs||_2 with L the
discrete operator of the FP64 solve, evaluated by the oracle):
  * one refinement sweep: relative residual <= 1e-10 (the FP64 target), and
    the solution within 1e-9 of the oracle's FP64 solution (mean-removed
    max-abs, relative to the oracle's max-abs);
  * no refinement (FP32 sweep only): relative residual < 1e-3;
  * device residual norm of an arbitrary field: within 1e-12 (relative) of
    the oracle's.
"""
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


def make_solver(c, precision=D.PRECISION_MIXED):
    bands = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
    s = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], bands, precision=precision)
    return s, bands


def rel_mean_removed(got, ref):
    g = got - got.mean()
    r = ref - ref.mean()
    return np.abs(g - r).max() / max(np.abs(r).max(), 1e-300)


def oracle_rel_resid(p, c, bands):
    r = O.apply_operator(p, c["inv_dx"], c["inv_dz"], bands) - c["rhs"]
    return np.linalg.norm(r) / np.linalg.norm(c["rhs"])


def case_grid(nx, nz, ny, seed):
    rng = np.random.default_rng(seed)
    yu = W.tanh_faces(ny, float(rng.uniform(1.2, 2.2)))
    return dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / W.LX, inv_dz=nz / W.LZ, yu=yu,
                rhs=W.rhs_uniform((ny, nz, nx), seed, yu))


MIXED_CASES = [(32, 32, 2), (96, 64, 7), (128, 96, 33), (192, 128, 40), (256, 32, 130), (384, 64, 17),
               (512, 32, 12), (768, 32, 24), (32, 768, 5), (64, 384, 64)]


def test_config1_mixed_one_refinement(cuda):
    say("mixed config1")
    c = W.config1()
    s, bands = make_solver(c)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    f = to_dev(c["rhs"])
    P = to_dev(np.zeros_like(c["rhs"]))
    s.solve_mixed(f, P, refinements=1)
    got = to_host(P)
    assert oracle_rel_resid(got, c, bands) <= 1e-10
    assert rel_mean_removed(got, ref) <= 1e-9
    assert np.array_equal(to_host(f), c["rhs"])  # rhs untouched
    s.solve_mixed(f, P, refinements=0)
    got0 = to_host(P)
    assert oracle_rel_resid(got0, c, bands) < 1e-3
    s.solve_mixed(f, P, refinements=2)
    assert oracle_rel_resid(to_host(P), c, bands) <= 1e-10
    s.close()


@pytest.mark.parametrize("nx,nz,ny", MIXED_CASES)
def test_mixed_matches_oracle(cuda, nx, nz, ny):
    say("mixed", nx, nz, ny)
    c = case_grid(nx, nz, ny, 3 * nx + 5 * nz + ny)
    s, bands = make_solver(c)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    f = to_dev(c["rhs"])
    P = to_dev(np.zeros_like(c["rhs"]))
    done, rel = s.solve_refine(f, P, tol=1e-10, max_refinements=4)
    got = to_host(P)
    assert 1 <= done <= 2 and rel <= 1e-10
    assert oracle_rel_resid(got, c, bands) <= 1e-10
    assert rel_mean_removed(got, ref) <= 1e-9
    # fixed-sweep entry point, same sweeps: identical result
    P2 = to_dev(np.zeros_like(c["rhs"]))
    s.solve_mixed(f, P2, refinements=done)
    assert np.array_equal(to_host(P2), got)
    s.close()


def test_fp32_sweeps_split_and_accuracy(cuda):
    """The sweeps run FP32 arithmetic on the mode groups whose modes all have
    |lambda| >= max_k(|a_k| + |c_k|) / 128 and FP64 on the rest (the (0,0)
    and lowest modes); with either arithmetic one refinement meets 1e-10 on
    a wall-refined 384-row grid (the production y grid at a smaller x-z
    box), and the two agree with the oracle's FP64 solution to 1e-9."""
    say("mixed fp32 sweeps split")
    nx, nz, ny = 192, 96, 384
    yu = W.wall_refined_faces(ny, 0.05, 1.015)
    c = dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / W.LX, inv_dz=nz / W.LZ, yu=yu,
             rhs=W.rhs_band_limited((ny, nz, nx), 1, yu=yu))
    s, bands = make_solver(c)
    on, lam_min, n32 = s.mixed_info
    a, cc = np.asarray(bands[0]), np.asarray(bands[2])
    assert on and lam_min == np.max(np.abs(a) + np.abs(cc)) / 128.0
    ax, _ = O.eigen(nx, c["inv_dx"])
    az, _ = O.eigen(nz, c["inv_dz"])
    lam = np.abs((az[:, None] + ax[None, :]).ravel())  # mode a + nx*b
    M = 16  # modes per group of the default variant
    want = 0
    for g in range(1, (nx * nz + M - 1) // M):
        blk = lam[g * M:(g + 1) * M]
        if np.all(blk >= lam_min):
            want += blk.size
    assert n32 == want and 0 < n32 < nx * nz  # 30% of the modes on this small x-z box
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    f = to_dev(c["rhs"])
    got = {}
    for fp32 in (True, False):
        s.set_mixed_fp32(fp32)
        P = to_dev(np.zeros_like(c["rhs"]))
        s.solve_mixed(f, P, refinements=1)
        got[fp32] = to_host(P)
        assert oracle_rel_resid(got[fp32], c, bands) <= 1e-10, fp32
        assert rel_mean_removed(got[fp32], ref) <= 1e-9, fp32
    assert s.mixed_info[0] is False and s.mixed_info[2] == 0
    assert not np.array_equal(got[True], got[False])
    s.close()


def test_mixed_in_place_equals_out_of_place(cuda):
    say("mixed in place")
    c = case_grid(192, 96, 24, 11)
    s, _ = make_solver(c)
    f = to_dev(c["rhs"])
    for refinements in (0, 1, 2, 3):
        P = to_dev(np.zeros_like(c["rhs"]))
        s.solve_mixed(f, P, refinements)
        Q = to_dev(c["rhs"])
        s.run_mixed(Q, refinements)
        assert np.array_equal(to_host(P), to_host(Q)), refinements
    s.close()


def test_mixed_is_deterministic(cuda):
    say("mixed determinism")
    c = case_grid(256, 128, 48, 5)
    s, _ = make_solver(c)
    f = to_dev(c["rhs"])
    outs = []
    for _ in range(3):
        P = to_dev(np.zeros_like(c["rhs"]))
        s.solve_mixed(f, P, 1)
        outs.append(to_host(P))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    s.close()


def test_every_mixed_variant_and_the_tuner(cuda):
    say("mixed variants")
    c = case_grid(384, 96, 40, 21)
    s, bands = make_solver(c)
    f = to_dev(c["rhs"])
    for v in range(4):
        for tv in range(8):
            s.set_mixed_variants(v, (v + 2) % 4, tv)
            P = to_dev(np.zeros_like(c["rhs"]))
            s.solve_mixed(f, P, 1)
            assert oracle_rel_resid(to_host(P), c, bands) <= 1e-10, (v, tv)
    say("mixed autotune")
    table = s.autotune(reps=2)
    assert table["mixed_picked"] == s.mixed_variants
    assert all(t > 0 for t in table["mixed_x_ms"] + table["mixed_z_ms"] + table["mixed_tri_ms"])
    P = to_dev(np.zeros_like(c["rhs"]))
    s.solve_mixed(f, P, 1)
    assert oracle_rel_resid(to_host(P), c, bands) <= 1e-10
    s.close()


@pytest.mark.parametrize("ny", [600, 1024])
def test_every_tridiag_variant_tall_columns(cuda, ny):
    """Every tridiagonal workgroup shape of both forms on columns taller than
    16 chunks of 32 rows hold (the FP32 16-chunk slots run another shape
    there), up to the partitioned solver's limit of 1024 rows."""
    say("tall variants", ny)
    c = case_grid(32, 32, ny, ny)
    s, bands = make_solver(c)
    f = to_dev(c["rhs"])
    for tv in range(8):
        s.set_variants(0, 0, tv)
        P = to_dev(np.zeros_like(c["rhs"]))
        s.solve(f, P)
        assert oracle_rel_resid(to_host(P), c, bands) <= 1e-10, ("fp64", tv)
        s.set_mixed_variants(0, 0, tv)
        P = to_dev(np.zeros_like(c["rhs"]))
        s.solve_mixed(f, P, 1)
        assert oracle_rel_resid(to_host(P), c, bands) <= 1e-10, ("mixed", tv)
    s.close()


@pytest.mark.parametrize("nx,nz,ny", [(48, 20, 10), (7, 5, 3), (1, 1, 2), (64, 32, 64)])
def test_device_residual_matches_oracle(cuda, nx, nz, ny):
    say("residual", nx, nz, ny)
    c = case_grid(nx, nz, ny, nx + nz + ny)
    s, bands = make_solver(c, precision=D.PRECISION_FP64)
    p = np.random.default_rng(1).standard_normal(c["rhs"].shape)
    rel = s.residual(to_dev(c["rhs"]), to_dev(p))
    want = oracle_rel_resid(p, c, bands)
    assert abs(rel - want) <= 1e-12 * want
    # the FP64 solve's residual, both ways, below the target
    P = to_dev(c["rhs"])
    s.run(P)
    assert s.residual(to_dev(c["rhs"]), P) <= 1e-10
    assert oracle_rel_resid(to_host(P), c, bands) <= 1e-10
    s.close()


def test_mixed_errors(cuda):
    say("mixed errors")
    c = case_grid(32, 32, 8, 2)
    s, _ = make_solver(c, precision=D.PRECISION_FP64)
    f = to_dev(c["rhs"])
    with pytest.raises(D.DnspError) as e:
        s.solve_mixed(f, f, 1)
    assert e.value.code == D.DNSP_ERR_ARG
    s.close()
    s, _ = make_solver(c)
    with pytest.raises(D.DnspError) as e:
        s.solve_mixed(f, f, D.MAX_REFINEMENTS + 1)
    assert e.value.code == D.DNSP_ERR_ARG
    s.close()
    c = case_grid(7, 32, 8, 2)
    with pytest.raises(D.DnspError) as e:
        make_solver(c)
    assert e.value.code == D.DNSP_ERR_UNSUPPORTED


def test_refine_reports_unmet_tolerance_and_nonfinite(cuda):
    """Failure statuses of the refinement loop: a tolerance no sweep count can
    reach gives DNSP_ERR_NOT_CONVERGED, a NaN in the right-hand side gives
    DNSP_ERR_NONFINITE (the reference stops on NaN, mod_get_bulk.f90:180-186),
    never DNSP_OK; the FP64 solve of the same NaN field is reported by the
    check entry point and by the residual."""
    say("refine failure statuses")
    c = case_grid(64, 32, 16, 5)
    s, _ = make_solver(c)
    f = to_dev(c["rhs"])
    P = to_dev(np.zeros_like(c["rhs"]))
    with pytest.raises(D.DnspError) as e:
        s.solve_refine(f, P, tol=0.0, max_refinements=1)
    assert e.value.code == D.DNSP_ERR_NOT_CONVERGED
    done, rel = s.solve_refine(f, P, tol=1e-10, max_refinements=3)
    assert rel <= 1e-10 and s.check(P) == 0
    bad = c["rhs"].copy()
    bad[3, 5, 7] = np.nan
    fb = to_dev(bad)
    with pytest.raises(D.DnspError) as e:
        s.solve_refine(fb, P, tol=1e-10, max_refinements=2)
    assert e.value.code == D.DNSP_ERR_NONFINITE
    Q = to_dev(np.zeros_like(bad))
    s.solve(fb, Q)
    with pytest.raises(D.DnspError) as e:
        s.check(Q)
    assert e.value.code == D.DNSP_ERR_NONFINITE
    with pytest.raises(D.DnspError) as e:
        s.residual(fb, Q)
    assert e.value.code == D.DNSP_ERR_NONFINITE
    inf = c["rhs"].copy()
    inf[0, 0, 0] = np.inf
    with pytest.raises(D.DnspError) as e:
        s.run_host(inf)
    assert e.value.code == D.DNSP_ERR_NONFINITE
    s.close()


@pytest.mark.slow
def test_config2_mixed_full_size(cuda):
    """768 x 384 x 384: one refinement reaches 1e-10; agreement with the FP64 form."""
    import torch
    say("config2 mixed: build rhs")
    c = W.config2()
    s, bands = make_solver(c)
    f = to_dev(c["rhs"])
    say("config2 mixed: fp64 solve")
    P64 = torch.empty_like(f)
    s.solve(f, P64)
    rel64 = s.residual(f, P64)
    assert rel64 <= 1e-10
    say("config2 mixed: refine")
    P = torch.empty_like(f)
    done, rel = s.solve_refine(f, P, tol=1e-10, max_refinements=4)
    assert done == 1 and rel <= 1e-10
    err = rel_mean_removed(to_host(P), to_host(P64))
    say("config2 mixed: rel residual", rel, "fp64", rel64, "err vs fp64", err)
    assert err <= 1e-9
    s.close()

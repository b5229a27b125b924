"""GPU checks of the batched form (many uneven instances per launch) against
the single-instance solver and the CPU oracle.

Every instance of a batched FP64 solve must equal the single-instance solve
with the same kernel variants bit for bit (the batched launches run the same
tile and mode-group computations), and the oracle within 1e-12 (mean-removed
max-abs, relative).  The mixed-precision batch must equal the single-instance
mixed solve bit for bit and reach a relative residual <= 1e-10 with one
refinement.  A graph capturing the batched solve must reproduce the eager
result on every replay.
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


def rel_mean_removed(got, ref):
    g = got - got.mean()
    r = ref - ref.mean()
    return np.abs(g - r).max() / max(np.abs(r).max(), 1e-300)


def small_cases(seed, n=7):
    """Uneven small instances over the compile-time FFT lengths."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx = int(rng.choice([32, 64, 96, 128, 192]))
        nz = int(rng.choice([32, 64, 96, 128]))
        ny = int(rng.choice([2, 5, 17, 40, 64, 96, 130]))
        yu = W.tanh_faces(ny, float(rng.uniform(1.2, 2.2)))
        bands = D.mk_P_abc(yu, False) + D.mk_P_abc(yu, True)
        rhs = W.rhs_uniform((ny, nz, nx), 100 * seed + i, yu)
        out.append(dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / W.LX, inv_dz=nz / W.LZ, yu=yu, bands=bands, rhs=rhs))
    return out


def bind(batch, cases):
    import torch
    f = [to_dev(c["rhs"]) for c in cases]
    P = [torch.zeros_like(t) for t in f]
    batch.bind(f, P)
    return f, P


@pytest.mark.parametrize("tvar", [0, 1, 2, 3])
def test_batch_fp64_equals_single_and_oracle(cuda, tvar):
    say("batch fp64", tvar)
    cases = small_cases(1)
    b = D.PoissonBatch(cases)
    f, P = bind(b, cases)
    for xv, zv in ((0, 0), (2, 1), (3, 3)):
        b.set_variants(xv, zv, tvar)
        b.run()
        for c, fi, Pi in zip(cases, f, P):
            s = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], c["bands"])
            # same kernels, tile decomposition and 8-mode groups of 32 chunks
            # (one- or two-pass) as the batch (its register caps change no
            # arithmetic)
            s.set_strategy(D.FFT_FIXED, D.TRIDIAG_PARTITION)
            s.set_variants(xv, zv, 1 if tvar < 2 else 5)
            Q = to_dev(np.zeros_like(c["rhs"]))
            s.solve(fi, Q)
            got = to_host(Pi)
            assert np.array_equal(got, to_host(Q)), (c["nx"], c["nz"], c["ny"], xv, zv)
            s.close()
    for c, Pi in zip(cases, P):
        ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], c["bands"])
        assert rel_mean_removed(to_host(Pi), ref) <= 1e-12, (c["nx"], c["nz"], c["ny"])
    b.close()


@pytest.mark.parametrize("tvar", [4, 5])
def test_batch_16_chunk_tridiagonal_variants(cuda, tvar):
    """16 chunks of twice the rows: no single-solve kernel has this shape, so
    the FP64 results are checked against the oracle and against variant 0
    (another partition of the rows, another rounding), both to 1e-12; the
    mixed form with the same shape reaches its 1e-10 residual."""
    say("batch 16 chunks", tvar)
    cases = small_cases(1)
    b = D.PoissonBatch(cases)
    f, P = bind(b, cases)
    b.set_variants(2, 1, 0)
    b.run()
    ref0 = [to_host(Pi) for Pi in P]
    b.set_variants(2, 1, tvar)
    b.run()
    for c, Pi, r0 in zip(cases, P, ref0):
        got = to_host(Pi)
        assert rel_mean_removed(got, O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], c["bands"])) <= 1e-12, c["ny"]
        assert rel_mean_removed(got, r0) <= 1e-12, c["ny"]
    b.close()
    cases = small_cases(2)
    b = D.PoissonBatch(cases, precision=D.PRECISION_MIXED)
    f, P = bind(b, cases)
    b.set_variants(0, 0, 1, 2, 1, tvar)
    b.run_mixed(1)
    for c, Pi in zip(cases, P):
        got = to_host(Pi)
        r = O.apply_operator(got, c["inv_dx"], c["inv_dz"], c["bands"]) - c["rhs"]
        rel = np.linalg.norm(r) / np.linalg.norm(c["rhs"])
        assert rel <= 1e-10, (c["nx"], c["nz"], c["ny"], rel)
    b.close()


def test_batch_persistent_equals_per_size_launches(cuda):
    """One cooperative launch over all instances and stages gives the results
    of the per-size launches with the same variants (x 2, z 2, two-pass
    tridiagonal), bit for bit, and the oracle's to 1e-12."""
    say("batch persistent")
    cases = small_cases(1)
    b = D.PoissonBatch(cases)
    f, P = bind(b, cases)
    b.set_variants(2, 2, 2)
    b.run()
    ref = [to_host(Pi) for Pi in P]
    for Pi in P:
        Pi.zero_()
    for _ in range(2):  # the second launch reuses the group table
        b.run_persistent()
        for c, Pi, r in zip(cases, P, ref):
            assert np.array_equal(to_host(Pi), r), (c["nx"], c["nz"], c["ny"])
    for c, Pi in zip(cases, P):
        assert rel_mean_removed(to_host(Pi), O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], c["bands"])) <= 1e-12
    b.close()


def test_batch_mixed_equals_single_and_reaches_target(cuda):
    say("batch mixed")
    cases = small_cases(2)
    b = D.PoissonBatch(cases, precision=D.PRECISION_MIXED)
    f, P = bind(b, cases)
    b.set_variants(0, 0, 1, 2, 1, 1)
    for refinements in (0, 1):
        b.run_mixed(refinements)
        for c, fi, Pi in zip(cases, f, P):
            s = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], c["bands"],
                                precision=D.PRECISION_MIXED)
            s.set_mixed_variants(2, 1, 1)
            Q = to_dev(np.zeros_like(c["rhs"]))
            s.solve_mixed(fi, Q, refinements)
            got = to_host(Pi)
            assert np.array_equal(got, to_host(Q)), (c["nx"], c["nz"], c["ny"], refinements)
            r = O.apply_operator(got, c["inv_dx"], c["inv_dz"], c["bands"]) - c["rhs"]
            rel = np.linalg.norm(r) / np.linalg.norm(c["rhs"])
            assert rel <= (1e-10 if refinements else 1e-3), (c["nx"], c["nz"], c["ny"], refinements, rel)
            s.close()
    b.close()


def test_batch_autotune_and_graph_replay(cuda):
    import torch
    say("batch tune + graph")
    cases = small_cases(3, n=5)
    b = D.PoissonBatch(cases, precision=D.PRECISION_MIXED)
    f, P = bind(b, cases)
    table = b.autotune(reps=2)
    assert all(t > 0 for t in table["x_ms"] + table["z_ms"] + table["tri_ms"] + table["mixed_x_ms"])
    for form in ("fp64", "mixed"):
        run = b.run if form == "fp64" else (lambda stream=None: b.run_mixed(1, stream))
        run()
        eager = [to_host(p) for p in P]
        for p in P:
            p.zero_()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                run(stream=s)
        torch.cuda.synchronize()
        for rep in range(3):
            for p in P:
                p.fill_(float(rep))  # eager work between replays
            g.replay()
            torch.cuda.synchronize()
            for e, p in zip(eager, P):
                assert np.array_equal(e, to_host(p)), (form, rep)
        del g
    b.close()


def test_batch_errors(cuda):
    say("batch errors")
    cases = small_cases(4, n=2)
    with pytest.raises(D.DnspError):
        D.PoissonBatch([dict(cases[0], nx=7)])
    b = D.PoissonBatch(cases)
    with pytest.raises(D.DnspError):
        b.run()  # nothing bound
    f = [to_dev(c["rhs"]) for c in cases]
    with pytest.raises(D.DnspError):
        b.bind(f, f)  # rhs and P must differ
    with pytest.raises(D.DnspError):
        b.run_mixed(1)  # FP64 batch
    with pytest.raises(D.DnspError):
        b.run_persistent()  # nothing bound
    b.close()
    # the persistent launch holds at most 12 rows per chunk (ny <= 384)
    yu = W.tanh_faces(400, 1.5)
    tall = dict(nx=32, nz=32, ny=400, inv_dx=32 / W.LX, inv_dz=32 / W.LZ,
                bands=D.mk_P_abc(yu, False) + D.mk_P_abc(yu, True), rhs=W.rhs_uniform((400, 32, 32), 9, yu))
    b = D.PoissonBatch([tall])
    bind(b, [tall])
    with pytest.raises(D.DnspError, match="rows per chunk"):
        b.run_persistent()
    b.run()  # the per-size launches take it
    b.close()


@pytest.mark.slow
def test_config3_batch(cuda):
    """The 48-instance batch: every instance's FP64 solution against the
    single-instance solver, the mixed form's residual on the device."""
    import torch
    say("config3 build")
    cases = W.config3()
    for c in cases:
        c["bands"] = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
    b = D.PoissonBatch(cases, precision=D.PRECISION_MIXED)
    f = [to_dev(c["rhs"]) for c in cases]
    P = [torch.zeros_like(t) for t in f]
    b.bind(f, P)
    say("config3 autotune")
    b.autotune(reps=1)
    v = b.variants
    b.run()
    for c, fi, Pi in zip(cases, f, P):
        s = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], c["bands"],
                            precision=D.PRECISION_MIXED)
        s.set_strategy(D.FFT_FIXED, D.TRIDIAG_PARTITION)
        # 8-mode groups of 32 chunks, one- or two-pass; the 16-chunk batch
        # variants (4, 5) have no single-solve kernel: there the same solution
        # to 1e-12, another partition of the rows rounding differently
        s.set_variants(v[0], v[1], 1 if v[2] in (0, 1, 4) else 5)
        Q = torch.zeros_like(fi)
        s.solve(fi, Q)
        if v[2] < 4:
            assert torch.equal(Q, Pi)
        else:
            assert rel_mean_removed(to_host(Pi), to_host(Q)) <= 1e-12, (c["nx"], c["nz"], c["ny"])
        assert s.residual(fi, Pi) <= 1e-10
        s.close()
    say("config3 persistent")
    # one persistent launch equals the per-size launches with its variants
    b.set_variants(2, 2, 2, *v[3:])
    b.run()
    ref = [Pi.clone() for Pi in P]
    for Pi in P:
        Pi.zero_()
    b.run_persistent()
    for c, Pi, r in zip(cases, P, ref):
        assert torch.equal(Pi, r), (c["nx"], c["nz"], c["ny"])
    say("config3 mixed")
    b.run_mixed(1)
    for c, fi, Pi in zip(cases, f, P):
        s = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], c["bands"])
        assert s.residual(fi, Pi) <= 1e-10, (c["nx"], c["nz"], c["ny"])
        s.close()
    b.close()

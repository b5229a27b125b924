"""GPU parity of the FP64 Poisson solve (libdnsproj.so through its C ABI)
against the CPU oracle on the same seeded inputs.

Tolerances:
  * tridiagonal stage: the serial kernel bit-exact (same operation order, no
    contraction); the partitioned kernel within 1e-11 of each mode's max-abs;
  * transform stages: 1e-13 relative to the max-abs of the stage output;
  * whole solve: mean-removed max-abs error <= 1e-12 relative to the oracle's
    mean-removed max-abs (the FP64 form's parity bar) and <= 1e-10 absolute on the
    reference-style random trials (main_poi_test.f90:259-262);
  * full size (768 x 384 x 384): discrete-operator residual <= 1e-10
    relative, forward/backward round trip, and agreement with the oracle.
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


def make_solver(c):
    bands = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
    return D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], bands), bands


def rel_mean_removed(got, ref):
    g = got - got.mean()
    r = ref - ref.mean()
    return np.abs(g - r).max() / max(np.abs(r).max(), 1e-300)


def case_grid(nx, nz, ny, seed):
    rng = np.random.default_rng(seed)
    yu = W.tanh_faces(ny, float(rng.uniform(1.2, 2.2)))
    return dict(nx=nx, nz=nz, ny=ny, inv_dx=nx / W.LX, inv_dz=nz / W.LZ, yu=yu,
                rhs=W.rhs_uniform((ny, nz, nx), seed, yu))


STAGE_CASES = [(64, 32, 64), (5, 7, 3), (6, 10, 9), (1, 1, 2), (2, 3, 4), (15, 14, 5), (768, 6, 3),
               (12, 384, 3), (96, 96, 17), (35, 49, 6), (32, 768, 5), (768, 32, 7), (512, 96, 4),
               (192, 256, 3), (128, 64, 70), (384, 128, 40), (7, 64, 5), (1, 96, 3), (3, 384, 4),
               (45, 128, 6), (2, 32, 3), (13, 11, 5), (624, 312, 3), (143, 26, 4), (44, 169, 3)]


@pytest.mark.parametrize("nx,nz,ny", STAGE_CASES)
def test_stages_match_oracle(cuda, nx, nz, ny):
    say("stages", nx, nz, ny)
    c = case_grid(nx, nz, ny, nx * 31 + nz * 7 + ny)
    s, bands = make_solver(c)
    f = c["rhs"]
    # forward
    t = to_dev(f)
    s.forward(t)
    Sg = to_host(t)
    Sr = O.forward(f)
    assert np.abs(Sg - Sr).max() <= 1e-13 * np.abs(Sr).max()
    # tridiagonal on identical input: the serial kernel is bit-exact
    Tr = O.tridiag(Sr, c["inv_dx"], c["inv_dz"], bands)
    s.set_strategy(D.FFT_AUTO, D.TRIDIAG_SERIAL)
    t = to_dev(Sr)
    s.tridiag(t)
    assert np.array_equal(to_host(t), Tr)
    # the partitioned kernel (registers + PCR on chunk tips) agrees per mode
    if ny <= 1024:
        s.set_strategy(D.FFT_AUTO, D.TRIDIAG_PARTITION)
        t = to_dev(Sr)
        s.tridiag(t)
        Tp = to_host(t)
        scale = np.abs(Tr).max(axis=0, keepdims=True)
        assert np.all(np.abs(Tp - Tr) <= 1e-11 * scale + 1e-300)
    s.set_strategy(D.FFT_AUTO, D.TRIDIAG_AUTO)
    # backward
    t = to_dev(Tr)
    s.backward(t)
    Pg = to_host(t)
    Pr = O.backward(Tr)
    assert np.abs(Pg - Pr).max() <= 1e-13 * np.abs(Pr).max()
    s.close()


def test_config1_solution_parity(cuda):
    say("config1 solve")
    c = W.config1()
    s, bands = make_solver(c)
    t = to_dev(c["rhs"])
    D.run_spec_poisson_multigpu(s, t)
    got = to_host(t)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    assert rel_mean_removed(got, ref) <= 1e-12
    r = O.apply_operator(got, c["inv_dx"], c["inv_dz"], bands) - c["rhs"]
    assert np.abs(r).max() <= 1e-11 * np.abs(c["rhs"]).max()
    mp = s.min_pivot
    assert abs(mp - O.min_pivot(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], bands)) <= 1e-14 * mp
    s.close()


def test_reference_style_random_trials(cuda):
    for n, c in enumerate(W.reference_style_trials(10, seed=0)):
        say("trial", n, c["nx"], c["nz"], c["ny"])
        bands = O.bands_from_yu(c["yu"])
        s = D.init_spec_poisson_multigpu(*bands, c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"])
        t = to_dev(c["rhs"])
        D.run_spec_poisson_multigpu(s, t)
        got = to_host(t)
        ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
        err = np.abs((got - got.mean()) - (ref - ref.mean())).max()
        assert err <= 1e-10
        assert rel_mean_removed(got, ref) <= 1e-12
        s.close()


@pytest.mark.parametrize("nx,nz,ny", [(7, 25, 4), (45, 21, 8), (128, 64, 33), (256, 192, 12), (384, 128, 9),
                                      (512, 256, 5), (192, 96, 40)])
def test_more_sizes_solution_parity(cuda, nx, nz, ny):
    say("sizes", nx, nz, ny)
    c = case_grid(nx, nz, ny, nx + nz + ny)
    s, bands = make_solver(c)
    t = to_dev(c["rhs"])
    s.run(t)
    got = to_host(t)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    assert rel_mean_removed(got, ref) <= 1e-12
    s.close()


STRATEGY_CASES = [(64, 32, 64), (768, 96, 40), (96, 384, 33), (256, 128, 100), (48, 40, 130), (384, 192, 384)]


@pytest.mark.parametrize("nx,nz,ny", STRATEGY_CASES)
def test_every_strategy_matches_oracle(cuda, nx, nz, ny):
    say("strategies", nx, nz, ny)
    c = case_grid(nx, nz, ny, 5 * nx + ny)
    s, bands = make_solver(c)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    outs = {}
    for fft in (D.FFT_GENERIC, D.FFT_AUTO):
        for tri in (D.TRIDIAG_SERIAL, D.TRIDIAG_PARTITION):
            s.set_strategy(fft, tri)
            t = to_dev(c["rhs"])
            s.run(t)
            outs[(fft, tri)] = to_host(t)
            assert rel_mean_removed(outs[(fft, tri)], ref) <= 1e-12, (fft, tri, s.strategy)
    s.close()


@pytest.mark.parametrize("nx,nz,ny", [(768, 384, 24), (96, 768, 40), (512, 64, 130), (32, 192, 7)])
def test_every_kernel_variant_matches_oracle(cuda, nx, nz, ny):
    say("variants", nx, nz, ny)
    c = case_grid(nx, nz, ny, nx + 3 * ny)
    s, bands = make_solver(c)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    for v in range(4):
        for tv in range(8):
            s.set_variants(v, (v + 1) % 4, tv)
            t = to_dev(c["rhs"])
            s.run(t)
            assert rel_mean_removed(to_host(t), ref) <= 1e-12, (v, tv)
    say("autotune", nx, nz, ny)
    table = s.autotune(reps=2)
    assert table["picked"] == s.variants
    assert all(v > 0 for v in table["x_ms"] + table["z_ms"])
    t = to_dev(c["rhs"])
    s.run(t)
    assert rel_mean_removed(to_host(t), ref) <= 1e-12
    s.close()


def test_strategy_reporting(cuda):
    say("strategy report")
    c = W.config2_grid()
    bands = D.mk_P_abc(c["yu"], False) + D.mk_P_abc(c["yu"], True)
    s = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], bands)
    assert s.strategy == {"fft_ct_x": 1, "fft_ct_z": 1, "part_rows": 12}
    s.set_strategy(D.FFT_GENERIC, D.TRIDIAG_SERIAL)
    assert s.strategy == {"fft_ct_x": 0, "fft_ct_z": 0, "part_rows": 0}
    s.close()
    c = case_grid(7, 25, 4, 1)
    s, _ = make_solver(c)
    with pytest.raises(D.DnspError):
        s.set_strategy(D.FFT_FIXED, D.TRIDIAG_AUTO)
    assert s.strategy["fft_ct_x"] == 0
    s.close()


def test_run_host_matches_device_run(cuda):
    say("run_host")
    c = case_grid(48, 20, 10, 3)
    s, bands = make_solver(c)
    h = c["rhs"].copy()
    s.run_host(h)
    t = to_dev(c["rhs"])
    s.run(t)
    assert np.array_equal(h, to_host(t))
    s.close()


def test_unsupported_length_is_an_error(cuda):
    say("unsupported")
    c = case_grid(17, 8, 4, 0)  # prime factor 17 > 13
    with pytest.raises(D.DnspError) as e:
        make_solver(c)
    assert e.value.code == D.DNSP_ERR_UNSUPPORTED


def test_singular_bands_are_an_error(cuda):
    say("singular")
    ny = 4
    yu = W.tanh_faces(ny, 1.5)
    a, b, cc = D.mk_P_abc(yu, False)
    z = np.zeros(ny)
    # all-zero bands for the (0,0) mode: its first pivot b0 + 0 is exactly zero
    with pytest.raises(D.DnspError) as e:
        D.PoissonSolver(4, 4, ny, 1.0, 1.0, (a, b, cc, z, z, z))
    assert e.value.code == D.DNSP_ERR_SINGULAR


def test_repeated_runs_are_deterministic(cuda):
    say("determinism")
    c = case_grid(96, 48, 16, 9)
    s, _ = make_solver(c)
    outs = []
    for _ in range(3):
        t = to_dev(c["rhs"])
        s.run(t)
        outs.append(to_host(t))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    s.close()


def test_eight_byte_aligned_arrays(cuda):
    """Arrays offset by one double from a 16-byte boundary (a tensor view, a
    Fortran section): the FP64 solve takes the runtime-sized x kernels and
    still matches the oracle; the mixed form rejects them (DNSP_ERR_ARG)."""
    import torch
    say("8-byte aligned arrays")
    c = case_grid(768, 32, 12, 4)
    s, bands = make_solver(c)
    n = c["rhs"].size
    buf = torch.zeros(n + 1, dtype=torch.float64, device=cuda)
    P = buf[1:].view(c["rhs"].shape)
    assert P.data_ptr() % 16 == 8
    P.copy_(to_dev(c["rhs"]))
    s.run(P)
    ref = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    assert rel_mean_removed(to_host(P), ref) <= 1e-12
    s.close()
    m = D.PoissonSolver(c["nx"], c["nz"], c["ny"], c["inv_dx"], c["inv_dz"], bands, precision=D.PRECISION_MIXED)
    with pytest.raises(D.DnspError) as e:
        m.solve_mixed(P, P, 1)
    assert e.value.code == D.DNSP_ERR_ARG
    m.close()


@pytest.mark.slow
def test_config2_full_size(cuda):
    """768 x 384 x 384: residual, round trip and oracle agreement."""
    import torch
    say("config2 build rhs")
    c = W.config2()
    s, bands = make_solver(c)
    f = c["rhs"]
    say("config2 round trip")
    t = to_dev(f)
    s.forward(t)
    s.backward(t)
    rt = to_host(t)
    assert np.abs(rt - f).max() <= 1e-13 * np.abs(f).max()
    say("config2 solve")
    t = to_dev(f)
    s.run(t)
    got = to_host(t)
    del t
    torch.cuda.empty_cache()
    say("config2 residual (cpu)")
    r = O.apply_operator(got, c["inv_dx"], c["inv_dz"], bands) - f
    assert np.abs(r).max() <= 1e-10 * np.abs(f).max()
    say("config2 oracle solve (cpu)")
    ref = O.solve(f, c["inv_dx"], c["inv_dz"], bands)
    assert rel_mean_removed(got, ref) <= 1e-12
    s.close()

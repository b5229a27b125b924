"""GPU parity of the projection stencils and IBM forcing against the oracle
(oracle/projection.py) on the same seeded inputs.  The kernels keep the
reference's evaluation order with contraction off: results must be
bit-identical."""
import sys

import numpy as np
import pytest

import projection as PJ
import dnsproj.projection as DP
from dnsproj import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.copy())
    return t.to("cuda:0")


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


SIZES = [(7, 5, 6), (16, 12, 10), (64, 32, 64), (33, 17, 9), (1, 1, 2)]


def fields(nx, nz, ny, seed, n=8):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((ny + 2, nz + 2, nx + 2)) for _ in range(n)], rng


@pytest.mark.parametrize("nx,nz,ny", SIZES)
def test_halo_helpers_bitexact(cuda, nx, nz, ny):
    say("halo", nx, nz, ny)
    (A, *_), _ = fields(nx, nz, ny, 1)
    t = dev(A)
    DP.halo_periodic_xz(t)
    assert np.array_equal(host(t), PJ.halo_periodic_xz(A.copy()))
    for k0, k1, v in [(0, 0, 1.5), (0, 1, -2.0), (-1, -1, 3.0), (2, 1, 0.25)]:
        t = dev(A)
        DP.set_val_slice(t, k0, k1, v)
        if max(k0, k1) < ny + 2:
            assert np.array_equal(host(t), PJ.set_val_slice(A.copy(), k0, k1, v))
    t = dev(A)
    DP.neumann_ghosts_y(t)
    assert np.array_equal(host(t), PJ.neumann_ghosts_y(A.copy()))


@pytest.mark.parametrize("nx,nz,ny", SIZES)
def test_build_p_rhs_bitexact(cuda, nx, nz, ny):
    say("rhs", nx, nz, ny)
    (Uh, Wh, Vh, Pn, Po, dn, rho, _), rng = fields(nx, nz, ny, 2)
    rho = 0.1 + np.abs(rho)
    g = PJ.geom_coeffs(W.tanh_faces(ny, 1.6))
    P = dev(np.zeros((ny, nz, nx)))
    DP.build_p_rhs(P, dev(Uh), dev(Wh), dev(Vh), dev(g["inv_dy_p"]), 40.0, 3.5, 2.25)
    assert np.array_equal(host(P), PJ.build_p_rhs_incomp(Uh, Wh, Vh, g["inv_dy_p"], 40.0, 3.5, 2.25))
    rmin = dev(np.zeros(1))
    DP.rho_min(dev(rho), rmin)
    rm = host(rmin)[0]
    assert rm == PJ.rho_min(rho)
    DP.build_p_rhs_varprop(P, dev(Uh), dev(Wh), dev(Vh), dev(Pn), dev(Po), dev(rho), dev(dn), dev(g["inv_dy_p"]),
                           dev(g["inv_dy_p_pad"]), rmin, 40.0, 3.5, 2.25, 3.5 ** 2, 2.25 ** 2)
    ref = PJ.build_p_rhs_varprop(Uh, Wh, Vh, Pn, Po, rho, dn, g["inv_dy_p"], g["inv_dy_p_pad"], 40.0, 3.5, 2.25,
                                 3.5 ** 2, 2.25 ** 2, rm)
    assert np.array_equal(host(P), ref)


@pytest.mark.parametrize("nx,nz,ny", SIZES)
@pytest.mark.parametrize("incomp", [True, False])
def test_update_vars_bitexact(cuda, nx, nz, ny, incomp):
    say("update_vars", nx, nz, ny, incomp)
    (U, V, Wv, Uh, Vh, Wh, Pn, Po), rng = fields(nx, nz, ny, 3 + incomp)
    rho = 0.1 + np.abs(rng.standard_normal(U.shape))
    p = rng.standard_normal((ny, nz, nx))
    g = PJ.geom_coeffs(W.tanh_faces(ny, 1.2))
    walls = dict(zip(PJ.WALL_KEYS, (0.5, -0.25, 0.125, 1.0, -1.5, 2.0)))
    rm = float(rho[1:-1, 1:-1, 1:-1].min())
    ref = PJ.update_vars(p, U, V, Wv, Uh, Vh, Wh, g["dPdy_ddy_hi"], g["dPdy_ddy_lo"], 0.003, 3.0, 2.0, 1, walls,
                         use_incompressible=incomp, P_now=Pn, P_old=Po, rho=rho, rho_min_v=rm)
    tU, tV, tW, tPn, tPo = dev(U), dev(V), dev(Wv), dev(Pn), dev(Po)
    tPer = dev(np.zeros_like(U))
    DP.update_vars(dev(p), tPer, tPn, tPo, tU, tV, tW, dev(Uh), dev(Vh), dev(Wh), dev(rho), dev(g["dPdy_ddy_hi"]),
                   dev(g["dPdy_ddy_lo"]), dev(np.array([rm])), dev(np.array([walls[k] for k in PJ.WALL_KEYS])), 1,
                   0.003, 3.0, 2.0, use_incompressible=incomp)
    assert np.array_equal(host(tU), ref[0])
    assert np.array_equal(host(tV), ref[1])
    assert np.array_equal(host(tW), ref[2])
    assert np.array_equal(host(tPer), ref[3])
    if not incomp:
        assert np.array_equal(host(tPn), ref[4])
        assert np.array_equal(host(tPo), ref[5])


def mk_rand_ibm(shape, n, rng, disjoint=True):
    used, used_g, out = set(), set(), []
    tries = 0
    while len(out) < n and tries < 100000:
        tries += 1
        t = tuple(int(rng.integers(0, s)) for s in shape)
        d = (0, 0, 0)
        while d == (0, 0, 0):
            d = tuple(int(v) for v in rng.integers(-1, 2, size=3))
        s1, s2 = (tuple((t[q] + m * d[q]) % shape[q] for q in range(3)) for m in (1, 2))
        if t in used_g:
            continue
        if disjoint and (t in used or s1 in used_g or s2 in used_g or t in (s1, s2)):
            continue
        used.update((t, s1, s2))
        used_g.add(t)
        out.append((t, rng.uniform(-2, 2), s1, rng.uniform(-2, 2), s2, rng.uniform(-2, 2)))
    return out


@pytest.mark.parametrize("shape,n", [((7, 5, 6), 30), ((64, 32, 16), 3000), ((3, 2, 2), 4), ((1, 4, 3), 5)])
@pytest.mark.parametrize("nhalo,nskip", [(0, 0), (1, 0), (1, 1), (0, 1)])
@pytest.mark.parametrize("disjoint", [True, False])
def test_ibm_exec_bitexact(cuda, tmp_path, shape, n, nhalo, nskip, disjoint):
    say("ibm", shape, n, nhalo, nskip, disjoint)
    nx, nz, ny = shape
    rng = np.random.default_rng(n + 10 * nhalo + nskip + (7 if disjoint else 0))
    n = min(n, nx * nz * ny // (4 if disjoint else 1))
    recs = mk_rand_ibm(shape, n, rng, disjoint)
    assert len(recs) > 0
    path = tmp_path / "c.dat"
    PJ.ibm_write_coeffs(path, recs)
    ijk, c = PJ.ibm_read_coeffs(path, nx, nz, ny)
    side = nhalo + nskip
    pad = rng.standard_normal((ny + 2 * side, nz + 2 * side, nx + 2 * side))
    ref = PJ.ibm_exec(pad, ijk, c, nx, nz, ny, nhalo, nskip)
    f = DP.IbmForcing(nx, nz, ny, nhalo, nskip, ijk.reshape(-1, 9), c)
    t = dev(pad)
    f.exec(t)
    assert np.array_equal(host(t), ref)
    info = f.info
    assert info["n_forcings"] == len(recs)
    if disjoint:
        assert info["disjoint"]
    f.close()


def test_ibm_rejects_repeated_targets(cuda):
    say("ibm repeated")
    ijk = np.array([[1, 1, 1, 2, 1, 1, 3, 1, 1], [1, 1, 1, 0, 1, 1, 0, 0, 1]], dtype=np.int32)
    with pytest.raises(DP.DnspError):
        DP.IbmForcing(4, 4, 4, 1, 0, ijk, np.ones((2, 3)))


@pytest.mark.parametrize("kmax_off", [0, 1])
def test_ibm_part2_bitexact(cuda, kmax_off):
    say("ibm part2", kmax_off)
    import torch
    rng = np.random.default_rng(11)
    nx, nz, ny = 12, 7, 9
    A = rng.standard_normal((ny + 2, nz + 2, nx + 2))
    bot = rng.integers(-1, 4, size=(nz, nx)).astype(np.int32)
    top = rng.integers(4, 10, size=(nz, nx)).astype(np.int32)
    kmax = ny - 1 + kmax_off
    ref = PJ.ibm_part2(A, bot, top, kmax, 0.75, -0.5)
    t = dev(A)
    DP.ibm_part2(t, torch.from_numpy(bot.ravel()).to("cuda:0"), torch.from_numpy(top.ravel()).to("cuda:0"), kmax,
                 0.75, -0.5)
    assert np.array_equal(host(t), ref)

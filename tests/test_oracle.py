"""Pins the CPU oracle (oracle/dnsp_oracle.c) against checks that do not
restate it: brute-force DFTs, the symbol of the periodic second difference,
closed-form band invariants, dense linear solves of the discrete operator
built independently with Kronecker products, and the reference's own
manufactured-solution and randomized-trial tests, restated.

The reference holds no committed golden vectors for this path; its checks
(tests/test_poisson_solver_and_ibm_rand/) generate data at test time.
"""
import math

import numpy as np
import pytest

import oracle as O
from dnsproj import workloads as W


def dft_matrix(n, sign=-1):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * (np.outer(j, j) % n) / n)


def halfcomplex_from_complex(X):
    n = X.size
    out = np.zeros(n)
    out[: n // 2 + 1] = X.real[: n // 2 + 1]
    for k in range(1, (n + 1) // 2):
        out[n - k] = X.imag[k]
    return out


SIZES = list(range(1, 41)) + [48, 64, 96, 128, 192, 256, 384, 768]


@pytest.mark.parametrize("n", SIZES)
def test_r2hc_matches_brute_force_dft(n):
    x = np.random.default_rng(n).standard_normal(n)
    ref = halfcomplex_from_complex(dft_matrix(n) @ x)
    got = O.r2hc(x)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", SIZES)
def test_hc2r_inverts_r2hc(n):
    x = np.random.default_rng(100 + n).standard_normal(n)
    y = O.hc2r(O.r2hc(x))
    assert np.abs(y - n * x).max() <= 1e-13 * n


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 49, 60, 96, 768])
def test_cfft_matches_numpy(n):
    rng = np.random.default_rng(7 + n)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.cfft(c) - np.fft.fft(c)).max() <= 1e-13 * n
    assert np.abs(O.cfft(c, inverse=True) - n * np.fft.ifft(c)).max() <= 1e-13 * n


@pytest.mark.parametrize("n", list(range(1, 17)) + [31, 64, 96, 768])
def test_eigenvalues_are_symbol_of_periodic_second_difference(n):
    """R2HC(D2 x) == lambda * R2HC(x): pins the halfcomplex -> wavenumber map
    (poisson_solver_multigpu.f90:139-142) and the eigenvalue formula (:147)."""
    inv_d = 3.7
    x = np.random.default_rng(n).standard_normal(n)
    d2x = inv_d ** 2 * (np.roll(x, -1) - 2 * x + np.roll(x, 1))
    lam, is00 = O.eigen(n, inv_d)
    lhs = O.r2hc(d2x)
    rhs = lam * O.r2hc(x)
    assert np.abs(lhs - rhs).max() <= 1e-12 * inv_d ** 2 * max(1.0, np.abs(x).sum())
    assert is00[0] == 1 and is00[1:].sum() == 0


def random_faces(ny, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 0.2, size=ny))]) + rng.uniform(-0.2, 0.2)


@pytest.mark.parametrize("ny", [2, 3, 5, 17, 64])
def test_bands_interior_rows_exact_on_quadratics(ny):
    yu = random_faces(ny, ny)
    ype = O.yp_ext_from_yu(yu)
    a, b, c = O.mk_P_abc(ype, False)
    yp = ype[1:-1]
    for k in range(1, ny - 1):
        raw_b = -(a[k] + c[k])
        for f, d2 in ((lambda y: 1.0 + 0 * y, 0.0), (lambda y: y, 0.0), (lambda y: y * y, 2.0)):
            v = a[k] * f(yp[k - 1]) + raw_b * f(yp[k]) + c[k] * f(yp[k + 1])
            assert abs(v - d2) <= 1e-9 * (abs(a[k]) + abs(c[k]))


@pytest.mark.parametrize("ny", [2, 3, 6, 33])
def test_bands_neumann_and_dirichlet_rows(ny):
    yu = random_faces(ny, 50 + ny)
    ype = O.yp_ext_from_yu(yu)
    a, b, c = O.mk_P_abc(ype, False)
    a0, b0, c0 = O.mk_P_abc(ype, True)
    L = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    L0 = np.diag(b0) + np.diag(a0[1:], -1) + np.diag(c0[:-1], 1)
    scale = np.abs(L).max()
    assert a[0] == 0 and c[-1] == 0 and a0[0] == 0 and c0[-1] == 0
    # Neumann at both walls: constants are in the null space
    assert np.abs(L @ np.ones(ny)).max() <= 1e-12 * scale
    # (W L) symmetric with w = Neumann weights -> w^T L = 0
    w = W.neumann_weights(yu)
    WL = w[:, None] * L
    assert np.abs(WL - WL.T).max() <= 1e-12 * np.abs(WL).max()
    assert np.abs(w @ L).max() <= 1e-12 * np.abs(WL).max()
    # Dirichlet first row: linear function vanishing at the bottom wall is exact
    yp = ype[1:-1]
    g = yp - yu[0]
    r = L0 @ g
    assert np.abs(r[: ny - 1]).max() <= 1e-10 * scale * (yu[-1] - yu[0])
    # the two sets differ only in b[0]
    assert np.array_equal(a, a0) and np.array_equal(c, c0) and np.array_equal(b[1:], b0[1:])
    # Dirichlet set is nonsingular
    assert abs(np.linalg.det(L0)) > 0


def test_bands_uniform_grid_is_classic_stencil():
    ny, h = 12, 0.25
    yu = h * np.arange(ny + 1)
    a, b, c = O.mk_P_abc(O.yp_ext_from_yu(yu), False)
    assert np.allclose(a[1:], 1 / h ** 2, rtol=1e-14) and np.allclose(c[:-1], 1 / h ** 2, rtol=1e-14)
    assert np.allclose(b[1:-1], -2 / h ** 2, rtol=1e-14)
    assert np.isclose(b[0], -1 / h ** 2, rtol=1e-14) and np.isclose(b[-1], -1 / h ** 2, rtol=1e-14)


def kron_operator(nx, nz, ny, inv_dx, inv_dz, bands):
    """Dense discrete operator built from first principles (independent of the oracle)."""
    a, b, c, a0, b0, c0 = bands

    def per(n, inv_d):
        D = -2 * np.eye(n) + np.roll(np.eye(n), 1, axis=1) + np.roll(np.eye(n), -1, axis=1)
        return inv_d ** 2 * D

    Ly = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    Ly0 = np.diag(b0) + np.diag(a0[1:], -1) + np.diag(c0[:-1], 1)
    Ix, Iz, Iy = np.eye(nx), np.eye(nz), np.eye(ny)
    mean_xz = np.full((nz * nx, nz * nx), 1.0 / (nz * nx))
    # C-order flattening of (ny, nz, nx): index = (k*nz + j)*nx + i
    L = (np.kron(Ly, np.kron(Iz, Ix)) + np.kron(Iy, np.kron(per(nz, inv_dz), Ix)) +
         np.kron(Iy, np.kron(Iz, per(nx, inv_dx))) + np.kron(Ly0 - Ly, mean_xz))
    return L


TINY = [(4, 3, 5), (5, 4, 6), (1, 1, 2), (2, 7, 3), (6, 6, 6), (8, 1, 4), (3, 9, 2), (10, 10, 10)]


@pytest.mark.parametrize("nx,nz,ny", TINY)
def test_operator_matches_kronecker_form(nx, nz, ny):
    rng = np.random.default_rng(nx * 100 + nz * 10 + ny)
    yu = random_faces(ny, ny + 3)
    inv_dx, inv_dz = 1 / rng.uniform(0.1, 0.2), 1 / rng.uniform(0.1, 0.2)
    bands = O.bands_from_yu(yu)
    p = rng.standard_normal((ny, nz, nx))
    L = kron_operator(nx, nz, ny, inv_dx, inv_dz, bands)
    got = O.apply_operator(p, inv_dx, inv_dz, bands)
    assert np.abs(got.ravel() - L @ p.ravel()).max() <= 1e-12 * np.abs(L).max() * np.abs(p).max()


@pytest.mark.parametrize("nx,nz,ny", TINY)
def test_solve_matches_dense_solve(nx, nz, ny):
    rng = np.random.default_rng(nx * 7 + nz * 3 + ny)
    yu = random_faces(ny, ny + 11)
    inv_dx, inv_dz = 1 / rng.uniform(0.1, 0.2), 1 / rng.uniform(0.1, 0.2)
    bands = O.bands_from_yu(yu)
    f = 5 * (rng.random((ny, nz, nx)) - 0.5)
    L = kron_operator(nx, nz, ny, inv_dx, inv_dz, bands)
    ref = np.linalg.solve(L, f.ravel()).reshape(f.shape)
    got = O.solve(f, inv_dx, inv_dz, bands)
    assert np.abs(got - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("nx,nz,ny", [(5, 4, 6), (6, 6, 6), (2, 3, 7)])
def test_stages_compose_and_tridiag_matches_dense(nx, nz, ny):
    rng = np.random.default_rng(3)
    yu = random_faces(ny, 9)
    inv_dx, inv_dz = 8.0, 6.0
    bands = O.bands_from_yu(yu)
    f = rng.standard_normal((ny, nz, nx))
    S = O.forward(f)
    # forward = R2HC along x then along z of every plane
    ref = np.empty_like(f)
    for k in range(ny):
        t = np.array([O.r2hc(f[k, j]) for j in range(nz)])
        ref[k] = np.array([O.r2hc(t[:, i]) for i in range(nx)]).T
    assert np.abs(S - ref).max() <= 1e-13 * np.abs(ref).max()
    T = O.tridiag(S, inv_dx, inv_dz, bands)
    lx, i0x = O.eigen(nx, inv_dx)
    lz, i0z = O.eigen(nz, inv_dz)
    a, b, c, a0, b0, c0 = bands
    for j in range(nz):
        for i in range(nx):
            is00 = i0x[i] and i0z[j]
            A, B, C = (a0, b0, c0) if is00 else (a, b, c)
            M = np.diag(B + lx[i] + lz[j]) + np.diag(A[1:], -1) + np.diag(C[:-1], 1)
            x = np.linalg.solve(M, S[:, j, i])
            assert np.abs(T[:, j, i] - x).max() <= 1e-12 * max(1.0, np.abs(x).max())
    P = O.backward(T)
    assert np.abs(P - O.solve(f, inv_dx, inv_dz, bands)).max() == 0.0


def test_reference_style_random_trials():
    """mk_test_poisson_solver.py:117-130 / main_poi_test.f90:224-262, single rank:
    mean-removed solution within 1e-10 of an exact dense solve."""
    for t in W.reference_style_trials(10, seed=0):
        nx, nz, ny = t["nx"], t["nz"], t["ny"]
        bands = O.bands_from_yu(t["yu"])
        L = kron_operator(nx, nz, ny, t["inv_dx"], t["inv_dz"], bands)
        ref = np.linalg.solve(L, t["rhs"].ravel()).reshape(t["rhs"].shape)
        got = O.solve(t["rhs"], t["inv_dx"], t["inv_dz"], bands)
        err = np.abs((got - got.mean()) - (ref - ref.mean())).max()
        assert err <= 1e-10


def test_complex_rfft2_formulation_agrees():
    """py_poi_solver.py:96-131 vs :161-207 (complex rfft2 vs real-real): restated
    with numpy's complex FFT and a dense per-mode solve; must agree at 1e-10."""
    nx, nz, ny = 12, 10, 9
    yu = random_faces(ny, 77)
    inv_dx, inv_dz = 7.0, 5.5
    bands = O.bands_from_yu(yu)
    a, b, c, a0, b0, c0 = bands
    f = np.random.default_rng(5).standard_normal((ny, nz, nx))
    F = np.fft.rfft2(f, axes=(1, 2))  # (ny, nz, nx//2+1)
    kx = np.fft.rfftfreq(nx, d=1 / nx)
    kz = np.fft.fftfreq(nz, d=1 / nz)
    lx = -4 * inv_dx ** 2 * np.sin(kx * np.pi / nx) ** 2
    lz = -4 * inv_dz ** 2 * np.sin(kz * np.pi / nz) ** 2
    G = np.empty_like(F)
    for j in range(nz):
        for i in range(nx // 2 + 1):
            A, B, C = (a0, b0, c0) if (i == 0 and j == 0) else (a, b, c)
            M = np.diag(B + lx[i] + lz[j]) + np.diag(A[1:], -1) + np.diag(C[:-1], 1)
            G[:, j, i] = np.linalg.solve(M, F[:, j, i])
    ref = np.fft.irfft2(G, s=(nz, nx), axes=(1, 2))
    got = O.solve(f, inv_dx, inv_dz, bands)
    assert np.abs(got - ref).max() <= 1e-10


def _manufactured():
    import sympy as sp
    x, y, z = sp.symbols("x y z", real=True)
    Lx, Ly, Lz = 5.63, 2, 5.63 / 2
    f1 = lambda s: 2 * sp.cos(s + sp.cos(s))
    f2 = lambda s: 3 * sp.cos(s + sp.sin(s))
    f3 = lambda s: 5 * sp.sin(s - sp.sin(s) * sp.cos(s))
    p = f1(2 * sp.pi * x / Lx) * f2(2 * sp.pi * y / Ly) * f3(2 * sp.pi * z / Lz)
    lap = sp.diff(p, x, 2) + sp.diff(p, y, 2) + sp.diff(p, z, 2)
    return (sp.lambdify((x, y, z), p, "numpy"), sp.lambdify((x, y, z), lap, "numpy"), Lx, Ly, Lz)


def _grid_y(L, n):
    """Stretched faces of py_poi_solver.py:276-295 (restated)."""
    y, dy = [0.0], 1.0
    all_dy = iter(np.linspace(1, 10, n // 2).tolist())
    for _ in range(n // 2):
        y.append(y[-1] + dy)
        dy = next(all_dy)
    y = np.array(y) / max(y) * 0.5
    y[-1] = 0.5
    other = np.cumsum(np.array(np.diff(y).tolist()[::-1]), 0)
    other[-1] = 0.5
    y = np.array(y.tolist() + (0.5 + other).tolist())
    y[-1] = 1
    y = y[:-1] + 0.5 * np.diff(y)
    return np.array([0.0] + y.tolist() + [1.0]) * L


def test_manufactured_solution_second_order():
    """py_poi_solver.py:236-327: error ratio ~4 per grid doubling."""
    p_f, lap_f, Lx, Ly, Lz = _manufactured()
    mid = lambda u: u[:-1] + 0.5 * np.diff(u)
    errs = []
    for N in [8, 16, 32, 64]:
        yu = _grid_y(Ly, N + 1)
        xp = mid(np.linspace(0, Lx, N + 1))
        zp = mid(np.linspace(0, Lz, N + 1))
        yp = mid(yu)
        Y, Z, X = np.meshgrid(yp, zp, xp, indexing="ij")
        rhs = lap_f(X, Y, Z)
        ref = p_f(X, Y, Z)
        ref -= ref.mean()
        got = O.solve(rhs, 1 / (Lx / N), 1 / (Lz / N), O.bands_from_yu(yu))
        got -= got.mean()
        errs.append(np.abs(got - ref).max())
    ratios = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    assert ratios[-1] > 3.5 and ratios[-2] > 3.0, (errs, ratios)


def test_config1_residual():
    c = W.config1()
    bands = O.bands_from_yu(c["yu"])
    p = O.solve(c["rhs"], c["inv_dx"], c["inv_dz"], bands)
    r = O.apply_operator(p, c["inv_dx"], c["inv_dz"], bands) - c["rhs"]
    assert np.abs(r).max() <= 1e-11 * np.abs(c["rhs"]).max()


def test_min_pivot_brute_force():
    nx, nz, ny = 6, 5, 7
    yu = random_faces(ny, 4)
    bands = O.bands_from_yu(yu)
    lx, i0x = O.eigen(nx, 9.0)
    lz, i0z = O.eigen(nz, 4.0)
    a, b, c, a0, b0, c0 = bands
    mn = np.inf
    for i in range(nx):
        for j in range(nz):
            A, B, C = (a0, b0, c0) if (i0x[i] and i0z[j]) else (a, b, c)
            M = np.diag(B + lx[i] + lz[j]) + np.diag(A[1:], -1) + np.diag(C[:-1], 1)
            # LU pivots of a tridiagonal matrix without pivoting
            piv = [M[0, 0]]
            for k in range(1, ny):
                piv.append(M[k, k] - M[k, k - 1] * M[k - 1, k] / piv[-1])
            mn = min(mn, np.abs(piv).min())
    got = O.min_pivot(nx, nz, ny, 9.0, 4.0, bands)
    assert abs(got - mn) <= 1e-12 * mn


@pytest.mark.parametrize("divs,n", [(1, 7), (3, 10), (4, 4), (5, 3), (8, 386)])
def test_get_lo_n_bounds_1d(divs, n):
    lo, cnt = O.get_lo_n_bounds_1d(divs, n)
    assert cnt.sum() == n and cnt.max() - cnt.min() <= 1
    assert list(cnt) == sorted(cnt, reverse=True)
    assert lo[0] == 0 and all(lo[i + 1] == lo[i] + cnt[i] for i in range(divs - 1))


def test_wall_refined_grid_shape():
    yu = W.wall_refined_faces(384)
    dy = np.diff(yu)
    assert yu.size == 385 and abs(yu[0]) < 1e-15 and abs(yu[-1] - 2.0) < 1e-12
    assert 1.5e-3 < dy[0] < 2.5e-3
    assert np.allclose(dy, dy[::-1], rtol=1e-9, atol=1e-15)
    g = dy[1:193] / dy[:192]
    assert np.all((np.abs(g - 1) < 1e-9) | (np.abs(g - 1.015) < 1e-9))
    # uniform part covers ~5% of the channel height
    nu = int(np.sum(np.abs(dy[:192] - dy[0]) < 1e-12))
    assert abs(nu * dy[0] - 0.1) < 2 * dy[0]

"""Pins the projection-stencil and IBM oracle (oracle/projection.py) with checks
that do not restate it, and checks the library's host helpers against it.

  * div(grad p) built from the divergence stencil (mod_cfd_subroutines.f90:527-537)
    and the gradient-correction stencil (:1124-1152) equals the Neumann band
    operator of the Poisson solve on a uniform y grid;
  * the full projection (RHS -> Poisson solve -> correction) leaves a
    discretely divergence-free velocity for a compatible RHS;
  * constant density reduces the variable-density forms to the incompressible ones;
  * IBM: the closed form A[t] = c1*A[s1] + c2*A[s2] + cw of the reference's
    test generator (mk_test_poisson_solver.py:148-151) on the interior, and
    periodic images in the halo shell.
"""
import os

import numpy as np
import pytest

import oracle as O
import projection as PJ
from dnsproj import workloads as W
import dnsproj.projection as DP


def uniform_faces(ny, h=0.1):
    return h * np.arange(ny + 1)


def rand_halo(rng, nx, nz, ny):
    return rng.standard_normal((ny + 2, nz + 2, nx + 2))


def walls0():
    return {k: 0.0 for k in PJ.WALL_KEYS}


def test_geom_coeffs_library_matches_oracle():
    for ny, yu in [(7, W.tanh_faces(7, 1.5)), (64, W.tanh_faces(64, 1.8)), (384, W.wall_refined_faces(384))]:
        a = PJ.geom_coeffs(yu)
        b = DP.geom_coeffs(yu)
        for k in ("inv_dy_p", "inv_dy_u", "inv_dy_p_pad", "dPdy_ddy_lo", "dPdy_ddy_hi"):
            assert np.array_equal(a[k], b[k]), k
        assert np.allclose(a["inv_dy_p"], 1 / np.diff(yu))
        assert np.array_equal(a["dPdy_ddy_lo"], -a["dPdy_ddy_hi"])
        assert a["dPdy_ddy_hi"][0] == 0.0


def test_halo_and_wall_helpers():
    rng = np.random.default_rng(0)
    A = rand_halo(rng, 5, 4, 3)
    B = PJ.halo_periodic_xz(A.copy())
    assert np.array_equal(B[:, :, 0], B[:, :, 5]) and np.array_equal(B[:, :, -1], B[:, :, 1])
    assert np.array_equal(B[:, 0, 1:-1], B[:, 4, 1:-1]) and np.array_equal(B[:, -1, 1:-1], B[:, 1, 1:-1])
    assert B[2, 0, 0] == A[2, 4, 5]  # corner = double wrap
    C = PJ.set_val_slice(A.copy(), 0, 1, 7.0)
    assert np.all(C[:2] == 7.0) and np.array_equal(C[2:], A[2:])
    C = PJ.set_val_slice(A.copy(), -1, -1, 3.0)
    assert np.all(C[-1] == 3.0) and np.array_equal(C[:-1], A[:-1])
    D = PJ.neumann_ghosts_y(A.copy())
    assert np.array_equal(D[0], A[1]) and np.array_equal(D[-1], A[-2])


def grad_of(p, g, inv_dx, inv_dz):
    """Velocities -grad(p) via update_vars with zero hats and dt = 1."""
    ny, nz, nx = p.shape
    z = np.zeros((ny + 2, nz + 2, nx + 2))
    U, V, W_, _, _, _ = PJ.update_vars(p, z, z, z, z, z, z, g["dPdy_ddy_hi"], g["dPdy_ddy_lo"], 1.0, inv_dx,
                                       inv_dz, 1, walls0())
    return U, V, W_


@pytest.mark.parametrize("nx,nz,ny", [(6, 5, 7), (8, 8, 4), (3, 7, 9)])
def test_divergence_of_gradient_is_the_poisson_operator(nx, nz, ny):
    rng = np.random.default_rng(nx + nz + ny)
    yu = uniform_faces(ny)
    g = PJ.geom_coeffs(yu)
    inv_dx, inv_dz = 7.0, 5.0
    p = rng.standard_normal((ny, nz, nx))
    U, V, W_ = grad_of(p, g, inv_dx, inv_dz)
    div = PJ.build_p_rhs_incomp(U, W_, V, g["inv_dy_p"], 1.0, inv_dx, inv_dz)
    a, b, c = O.mk_P_abc(O.yp_ext_from_yu(yu), False)
    Lp = O.apply_operator(p, inv_dx, inv_dz, (a, b, c, a, b, c))
    assert np.abs(div + Lp).max() <= 1e-12 * np.abs(Lp).max()


def test_projection_leaves_divergence_free_velocity():
    nx, nz, ny = 16, 12, 10
    rng = np.random.default_rng(3)
    yu = uniform_faces(ny, 0.2)
    g = PJ.geom_coeffs(yu)
    inv_dx, inv_dz, dt = 4.0, 3.0, 0.01
    bands = O.bands_from_yu(yu)
    Uh, Wh, Vh = (rand_halo(rng, nx, nz, ny) for _ in range(3))
    for A in (Uh, Wh, Vh):
        PJ.halo_periodic_xz(A)
    PJ.set_val_slice(Vh, 1, 1, 0.0)    # V_hat(:,:,0) = wall (mod_cfd_subroutines.f90:507-514)
    PJ.set_val_slice(Vh, -1, -1, 0.0)  # V_hat(:,:,ny) = wall (:516-523)
    f = PJ.build_p_rhs_incomp(Uh, Wh, Vh, g["inv_dy_p"], 1 / dt, inv_dx, inv_dz)
    assert abs(f.sum()) <= 1e-9 * np.abs(f).sum()  # compatible on a uniform grid
    p = O.solve(f, inv_dx, inv_dz, bands)
    z = np.zeros_like(Uh)
    U, V, W_, _, _, _ = PJ.update_vars(p, z, z, z, Uh, Vh, Wh, g["dPdy_ddy_hi"], g["dPdy_ddy_lo"], dt, inv_dx,
                                       inv_dz, 1, walls0())
    div = PJ.build_p_rhs_incomp(U, W_, V, g["inv_dy_p"], 1.0, inv_dx, inv_dz)
    assert np.abs(div).max() <= 1e-10 * np.abs(f).max() * dt


def test_constant_density_reduces_to_incompressible():
    nx, nz, ny = 7, 6, 5
    rng = np.random.default_rng(4)
    yu = W.tanh_faces(ny, 1.4)
    g = PJ.geom_coeffs(yu)
    rho0 = 0.37
    args = [rand_halo(rng, nx, nz, ny) for _ in range(6)]
    Uh, Wh, Vh, Pn, Po, div_now = args
    rho = np.full_like(Uh, rho0)
    inv_dtmax, inv_dx, inv_dz = 50.0, 3.0, 2.0
    got = PJ.build_p_rhs_varprop(Uh, Wh, Vh, Pn, Po, rho, div_now, g["inv_dy_p"], g["inv_dy_p_pad"], inv_dtmax,
                                 inv_dx, inv_dz, inv_dx ** 2, inv_dz ** 2, rho0)
    ref = PJ.build_p_rhs_incomp(Uh, Wh, Vh, g["inv_dy_p"], inv_dtmax, inv_dx, inv_dz) * rho0 \
        - inv_dtmax * rho0 * div_now[1:-1, 1:-1, 1:-1]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    p = rng.standard_normal((ny, nz, nx))
    out_v = PJ.update_vars(p, *args[:3], Uh, Vh, Wh, g["dPdy_ddy_hi"], g["dPdy_ddy_lo"], 0.01, inv_dx, inv_dz, 1,
                           walls0(), use_incompressible=False, P_now=Pn, P_old=Po, rho=rho, rho_min_v=rho0)
    out_i = PJ.update_vars(p / rho0, *args[:3], Uh, Vh, Wh, g["dPdy_ddy_hi"], g["dPdy_ddy_lo"], 0.01, inv_dx, inv_dz,
                           1, walls0())
    for a, b in zip(out_v[:3], out_i[:3]):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    assert np.array_equal(out_v[5], Pn) and np.array_equal(out_v[4], out_v[3])


def test_rho_min_interior_only():
    rng = np.random.default_rng(5)
    A = rng.uniform(1, 2, (6, 5, 4))
    A[0, 0, 0] = -9.0  # ghost, ignored
    A[2, 2, 2] = 0.5
    assert PJ.rho_min(A) == 0.5


def mk_rand_ibm(shape, n, rng):
    """Restates mk_test_poisson_solver.py:86-103: unique targets; sources t + d, t + 2d
    (d a random nonzero step in {-1,0,1}^3, wrapped) that are not targets."""
    used, used_g, out = set(), set(), []
    while len(out) < n:
        t = tuple(int(rng.integers(0, s)) for s in shape)
        d = (0, 0, 0)
        while d == (0, 0, 0):
            d = tuple(int(v) for v in rng.integers(-1, 2, size=3))
        s1, s2 = (tuple((t[q] + m * d[q]) % shape[q] for q in range(3)) for m in (1, 2))
        if t in used or s1 in used_g or s2 in used_g:
            continue
        used.update((t, s1, s2))
        used_g.add(t)
        out.append((t, rng.uniform(-2, 2), s1, rng.uniform(-2, 2), s2, rng.uniform(-2, 2)))
    return out


@pytest.mark.parametrize("nhalo,nskip", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_ibm_exec_closed_form_and_images(tmp_path, nhalo, nskip):
    rng = np.random.default_rng(10 * nhalo + nskip)
    nx, nz, ny = 7, 5, 6
    recs = mk_rand_ibm((nx, nz, ny), 40, rng)
    path = tmp_path / "ibm_coeffs_A.dat"
    PJ.ibm_write_coeffs(path, recs)
    ijk, c = PJ.ibm_read_coeffs(path, nx, nz, ny)
    side = nhalo + nskip
    A_in = rng.standard_normal((ny, nz, nx))
    pad = np.zeros((ny + 2 * side, nz + 2 * side, nx + 2 * side))
    # padded array filled the reference test's way (main_poi_test.f90:303-318)
    for k in range(nskip, ny + 2 * side - nskip):
        for j in range(nskip, nz + 2 * side - nskip):
            for i in range(nskip, nx + 2 * side - nskip):
                pad[k, j, i] = A_in[(k - side) % ny, (j - side) % nz, (i - side) % nx]
    out = PJ.ibm_exec(pad, ijk, c, nx, nz, ny, nhalo, nskip)
    A_ibm = A_in.copy()
    for t, c1, s1, c2, s2, cw in recs:
        A_ibm[t[2], t[1], t[0]] = c1 * A_in[s1[2], s1[1], s1[0]] + c2 * A_in[s2[2], s2[1], s2[0]] + cw
    inner = out[side:side + ny, side:side + nz, side:side + nx]
    assert np.abs(inner - A_ibm).max() <= 1e-10  # main_poi_test.f90:337-345
    # halo images of the forced targets carry the forced values
    for t, *_ in recs:
        for si in (-1, 0, 1):
            for sj in (-1, 0, 1):
                for sk in (-1, 0, 1):
                    q = (t[0] + si * nx, t[1] + sj * nz, t[2] + sk * ny)
                    if all(-nhalo <= q[d] <= (nx, nz, ny)[d] + nhalo - 1 for d in range(3)):
                        assert out[q[2] + side, q[1] + side, q[0] + side] == inner[t[2], t[1], t[0]]


def test_ibm_file_readers_library_matches_oracle(tmp_path):
    rng = np.random.default_rng(2)
    nx, nz, ny = 9, 4, 8
    recs = mk_rand_ibm((nx, nz, ny), 25, rng)
    # shift some indices out of range: the readers wrap them (mod_utils.f90:197-200)
    recs = [((t[0] - nx, t[1], t[2] + ny), c1, s1, c2, (s2[0] + nx, s2[1] - nz, s2[2]), cw)
            for t, c1, s1, c2, s2, cw in recs]
    path = tmp_path / "ibm_coeffs_U.dat"
    PJ.ibm_write_coeffs(path, recs)
    ijk_o, c_o = PJ.ibm_read_coeffs(path, nx, nz, ny)
    ijk_l, c_l = DP.ibm_read_coeffs(path, nx, nz, ny)
    assert np.array_equal(ijk_o.reshape(-1, 9), ijk_l) and np.array_equal(c_o, c_l)
    yp = W.tanh_faces(ny, 1.3)
    yp = 0.5 * (yp[1:] + yp[:-1])
    assert np.array_equal(PJ.ibm_div_coeffs(ijk_o, yp, nx, nz, 3.0, 2.5),
                          DP.ibm_div_coeffs(ijk_l, yp, nx, nz, 3.0, 2.5))
    # linear extrapolation along the line s1 -> s2 reproduces a linear field at the target
    c = PJ.ibm_div_coeffs(ijk_o, yp, nx, nz, 3.0, 2.5)
    assert np.allclose(c[:, 1] + c[:, 2], 1.0) and np.all(c[:, 0] == 0)


def test_ibm_slab_reader_library_matches_oracle(tmp_path):
    rng = np.random.default_rng(6)
    nx, nz, ny = 6, 5, 10
    bot = rng.integers(-3, 14, size=(nz, nx)).astype(np.int32)
    path = tmp_path / "ibm_inds_U_bot.dat"
    with open(path, "wb") as fw:
        fw.write(np.array([nx, nz], dtype=np.int32).tobytes())
        fw.write(bot.tobytes())
    for is_top in (False, True):
        got = DP.ibm_read_slab(path, nx, nz, ny, 0, is_top)
        b, t = PJ.ibm_clip_slabs(bot, bot, ny)
        assert np.array_equal(got, t if is_top else b)


def test_ibm_part2_columns():
    rng = np.random.default_rng(8)
    nx, nz, ny = 4, 3, 6
    A = rand_halo(rng, nx, nz, ny)
    bot = rng.integers(-1, 3, size=(nz, nx))
    top = rng.integers(3, 7, size=(nz, nx))
    out = PJ.ibm_part2(A, bot, top, ny - 1, 1.5, -2.5)
    for j in range(nz):
        for i in range(nx):
            col = out[1:-1, j + 1, i + 1]
            assert np.all(col[:bot[j, i] + 1] == 1.5)
            assert np.all(col[top[j, i]:] == -2.5)
            assert np.array_equal(col[bot[j, i] + 1:top[j, i]], A[1:-1, j + 1, i + 1][bot[j, i] + 1:top[j, i]])


# ---------------------------------------------------------------------------
# Variable-density projection identity
# ---------------------------------------------------------------------------
def _periodic_neumann(A):
    """Periodic x/z halos and Neumann y ghosts (every padded plane)."""
    PJ.halo_periodic_xz(A)
    return PJ.neumann_ghosts_y(A)


def _varprop_case(nx, nz, ny, seed, h=0.2, rho_lo=0.1):
    rng = np.random.default_rng(seed)
    yu = uniform_faces(ny, h)
    g = PJ.geom_coeffs(yu)
    Uh, Wh, Vh, div_now = (rand_halo(rng, nx, nz, ny) for _ in range(4))
    for A in (Uh, Wh, Vh):
        PJ.halo_periodic_xz(A)
    PJ.set_val_slice(Vh, 1, 1, 0.0)    # V_hat(:,:,0) = wall
    PJ.set_val_slice(Vh, -1, -1, 0.0)  # V_hat(:,:,ny) = wall
    rho = _periodic_neumann(rng.uniform(rho_lo, 1.0, (ny + 2, nz + 2, nx + 2)))
    Pn = _periodic_neumann(rand_halo(rng, nx, nz, ny))
    Po = _periodic_neumann(rand_halo(rng, nx, nz, ny))
    return dict(yu=yu, g=g, Uh=Uh, Wh=Wh, Vh=Vh, div_now=div_now, rho=rho, Pn=Pn, Po=Po)


def _projected_divergence(c, inv_dx, inv_dz, dt, rhs_rho=None, corr_rho=None):
    """build_p_rhs_varprop -> solve (mean removed) -> update_vars, and the
    divergence of the corrected velocity; rhs_rho / corr_rho replace rho in one
    of the two stencils (planted mistakes)."""
    g = c["g"]
    rho = c["rho"]
    rm = PJ.rho_min(rho)
    f = PJ.build_p_rhs_varprop(c["Uh"], c["Wh"], c["Vh"], c["Pn"], c["Po"],
                               rho if rhs_rho is None else rhs_rho, c["div_now"], g["inv_dy_p"],
                               g["inv_dy_p_pad"], 1 / dt, inv_dx, inv_dz, inv_dx ** 2, inv_dz ** 2, rm)
    fbar = f.mean()  # uniform y grid: the compatibility weights are uniform
    p = O.solve(f - fbar, inv_dx, inv_dz, O.bands_from_yu(c["yu"]))
    z = np.zeros_like(c["Uh"])
    U, V, W_, _, _, _ = PJ.update_vars(p, z, z, z, c["Uh"], c["Vh"], c["Wh"], g["dPdy_ddy_hi"],
                                       g["dPdy_ddy_lo"], dt, inv_dx, inv_dz, 1, walls0(),
                                       use_incompressible=False, P_now=c["Pn"], P_old=c["Po"],
                                       rho=rho if corr_rho is None else corr_rho, rho_min_v=rm)
    div = PJ.build_p_rhs_incomp(U, W_, V, g["inv_dy_p"], 1.0, inv_dx, inv_dz)
    return div, fbar, rm


@pytest.mark.parametrize("nx,nz,ny,seed", [(12, 10, 9, 0), (8, 6, 14, 1), (16, 4, 6, 2)])
def test_variable_density_projection_identity(nx, nz, ny, seed):
    """With rho varying by 10x, the variable-density splitting makes the
    projected velocity satisfy div(u) = div_now + dt*mean(f)/rho_min exactly:
    the RHS face factors 1 - rho_min*(1/(2 rho_nb) + 1/(2 rho))
    (mod_cfd_subroutines.f90:552-569) cancel the correction's
    (<1/rho> - 1/rho_min)(2 P_now - P_old) term (:1154-1197) face by face, and
    div(grad p) is the Neumann band operator the solve inverts."""
    inv_dx, inv_dz, dt = 4.0, 3.0, 0.01
    c = _varprop_case(nx, nz, ny, seed)
    assert c["rho"][1:-1, 1:-1, 1:-1].max() / PJ.rho_min(c["rho"]) > 5
    div, fbar, rm = _projected_divergence(c, inv_dx, inv_dz, dt)
    want = c["div_now"][1:-1, 1:-1, 1:-1] + dt * fbar / rm
    scale = np.abs(want).max()
    err = np.abs(div - want).max() / scale
    assert err <= 1e-12, err
    # planted mistakes: rho one cell off in the RHS only, or in the correction
    # only, must break the identity by orders of magnitude
    for axis in (0, 1, 2):
        shifted = _periodic_neumann(np.roll(c["rho"], 1, axis=axis).copy())
        for kw in ({"rhs_rho": shifted}, {"corr_rho": shifted}):
            d2, _, _ = _projected_divergence(c, inv_dx, inv_dz, dt, **kw)
            assert np.abs(d2 - want).max() / scale > 1e-3, (axis, list(kw))


def test_variable_density_identity_needs_both_terms():
    """Dropping the (2 P_now - P_old) history from both stencils (P_now = P_old
    = 0) still satisfies the identity, while a nonzero history that only one
    stencil sees does not: the two terms are each other's exact counterpart."""
    inv_dx, inv_dz, dt = 5.0, 2.0, 0.02
    c = _varprop_case(10, 8, 7, 3)
    c0 = dict(c, Pn=np.zeros_like(c["Pn"]), Po=np.zeros_like(c["Po"]))
    div, fbar, rm = _projected_divergence(c0, inv_dx, inv_dz, dt)
    want = c["div_now"][1:-1, 1:-1, 1:-1] + dt * fbar / rm
    assert np.abs(div - want).max() <= 1e-12 * np.abs(want).max()
    # history in the correction only
    g = c["g"]
    f = PJ.build_p_rhs_varprop(c["Uh"], c["Wh"], c["Vh"], c0["Pn"], c0["Po"], c["rho"], c["div_now"],
                               g["inv_dy_p"], g["inv_dy_p_pad"], 1 / dt, inv_dx, inv_dz, inv_dx ** 2,
                               inv_dz ** 2, rm)
    p = O.solve(f - f.mean(), inv_dx, inv_dz, O.bands_from_yu(c["yu"]))
    z = np.zeros_like(c["Uh"])
    U, V, W_, _, _, _ = PJ.update_vars(p, z, z, z, c["Uh"], c["Vh"], c["Wh"], g["dPdy_ddy_hi"],
                                       g["dPdy_ddy_lo"], dt, inv_dx, inv_dz, 1, walls0(),
                                       use_incompressible=False, P_now=c["Pn"], P_old=c["Po"], rho=c["rho"],
                                       rho_min_v=rm)
    d2 = PJ.build_p_rhs_incomp(U, W_, V, g["inv_dy_p"], 1.0, inv_dx, inv_dz)
    want2 = c["div_now"][1:-1, 1:-1, 1:-1] + dt * f.mean() / rm
    assert np.abs(d2 - want2).max() > 1e-3 * np.abs(want2).max()


# ---------------------------------------------------------------------------
# IBM div coefficients: linear extrapolation along the image line
# ---------------------------------------------------------------------------
def _collinear_records(nx, nz, ny):
    """(target, a, b) triples, collinear in physical space with a between t and
    b, including periodic wraps in x and z: vertical lines on any y grid, and
    lines along x, z and diagonals in (x, z).  Sources lie within half a period
    of the target, where the nearest periodic image is the one on the line."""
    recs = []
    for (i, j) in [(0, 0), (3, 2), (nx - 1, nz - 1)]:
        recs.append(((i, j, 0), (i, j, 1), (i, j, 3)))        # upwards from the bottom
        recs.append(((i, j, ny - 1), (i, j, ny - 2), (i, j, ny - 4)))
    for k in (1, ny // 2):
        recs.append(((nx - 1, 1, k), (0, 1, k), (1, 1, k)))                # x wrap
        recs.append(((2, nz - 1, k), (2, 0, k), (2, 1, k)))                # z wrap
        recs.append(((0, 0, k), (nx - 1, nz - 1, k), (nx - 2, nz - 2, k)))  # diagonal, both wrap
        recs.append(((4, 1, k), (5, 3, k), (6, 5, k)))                    # slope 2 in (x, z)
    return np.array(recs, dtype=np.int64)


def test_ibm_div_coeffs_extrapolate_linear_fields():
    """c1 f(a) + c2 f(b) = f(t) for every linear f when t, a, b are collinear
    with a between t and b (set_ibm_bands.f90:57-77: c1 = 1 + d_ta/d_ab,
    c2 = -d_ta/d_ab).  Swapping c1 and c2, or the distances, fails."""
    nx, nz, ny = 9, 12, 12
    inv_dx, inv_dz = 3.0, 2.5
    yp = W.tanh_faces(ny, 1.6)
    yp = 0.5 * (yp[1:] + yp[:-1])
    ijk = _collinear_records(nx, nz, ny)
    coeffs = {"oracle": PJ.ibm_div_coeffs(ijk, yp, nx, nz, inv_dx, inv_dz),
              "library": DP.ibm_div_coeffs(ijk.reshape(-1, 9), yp, nx, nz, inv_dx, inv_dz)}
    rng = np.random.default_rng(12)
    for _ in range(4):
        al, be, ga, de = rng.standard_normal(4)

        def f(i, j, k):
            return al * i / inv_dx + be * j / inv_dz + ga * yp[k] + de

        for name, c in coeffs.items():
            assert np.all(c[:, 0] == 0)
            for r, rec in enumerate(ijk):
                t, a, b = (tuple(int(v) for v in p) for p in rec)
                # unwrap a and b to the periodic image nearest the target
                ua = [a[0] + nx * round((t[0] - a[0]) / nx), a[1] + nz * round((t[1] - a[1]) / nz), a[2]]
                ub = [b[0] + nx * round((t[0] - b[0]) / nx), b[1] + nz * round((t[1] - b[1]) / nz), b[2]]
                ft, fa, fb = f(*t), f(*ua), f(*ub)
                got = c[r, 1] * fa + c[r, 2] * fb
                assert abs(got - ft) <= 1e-12 * (1 + abs(ft)), (name, r)
                # planted mistake: c1 and c2 swapped
                if abs(fa - fb) > 1e-6:
                    assert abs(c[r, 2] * fa + c[r, 1] * fb - ft) > 1e-6

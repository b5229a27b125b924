"""Pins of the predictor oracle (oracle/predictor.py) by what the mathematics
fixes, independently of its own formulas:

* the finite-difference coefficients are exact on linear functions for every
  row type and on quadratics for interior rows, the y interpolation weights
  reproduce linear functions at the faces, the wall extrapolation is exact on
  linear functions;
* the variable-property momentum and temperature terms reduce to the
  incompressible ones for constant properties (mu = nu, rho = 1, cp = 1,
  div_now = 0, no buoyancy), up to rounding;
* on manufactured smooth fields (periodic in x and z, non-uniform in y) every
  right-hand side converges at second order to the analytic one, including
  the variable-viscosity stress terms: a dropped term, a wrong sign or a
  wrong neighbour breaks the order.
"""
import math

import numpy as np
import pytest

import predictor as PR
from dnsproj import workloads as W


def lin(y):
    return 0.3 + 1.7 * y


def quad(y):
    return 0.3 + 1.7 * y - 2.2 * y * y


@pytest.mark.parametrize("grid", ["config1", "config2"])
def test_fd_coefficients_exact_on_polynomials(grid):
    yu = np.array(W.config1()["yu"] if grid == "config1" else W.config2_grid()["yu"])
    ny = len(yu) - 1
    g = PR.predictor_geom(yu)
    yp = g["yp"]
    ext = np.concatenate([[yu[0]], yp, [yu[-1]]])
    for i in range(ny):
        f = [lin(ext[i]), lin(ext[i + 1]), lin(ext[i + 2])]
        assert abs(g["Ruw_ddy"][i] @ f - 1.7) <= 1e-9 * 1.7 * np.abs(g["Ruw_ddy"][i]).sum() * (1 + abs(ext).max())
        assert abs(g["Ruw_d2dy2"][i] @ f) <= 1e-9 * np.abs(g["Ruw_d2dy2"][i]).sum() * (1 + abs(ext).max())
    for i in range(1, ny - 1):
        f = [quad(ext[i]), quad(ext[i + 1]), quad(ext[i + 2])]
        assert abs(g["Ruw_ddy"][i] @ f - (1.7 - 4.4 * ext[i + 1])) <= 1e-9 * np.abs(g["Ruw_ddy"][i]).sum()
        assert abs(g["Ruw_d2dy2"][i] @ f + 4.4) <= 1e-9 * np.abs(g["Ruw_d2dy2"][i]).sum()
    for i in range(1, ny):
        f = [quad(yu[i - 1]), quad(yu[i]), quad(yu[i + 1])]
        assert abs(g["Rv_ddy"][i] @ f - (1.7 - 4.4 * yu[i])) <= 1e-9 * np.abs(g["Rv_ddy"][i]).sum()
        assert abs(g["Rv_d2dy2"][i] @ f + 4.4) <= 1e-9 * np.abs(g["Rv_d2dy2"][i]).sum()
        ib, it = g["Rv_interp"][i]
        assert abs(ib * lin(yp[i - 1]) + it * lin(yp[i]) - lin(yu[i])) <= 1e-12
    assert np.all(g["Rv_ddy"][[0, ny]] == 0) and np.all(g["Rv_interp"][[0, ny]] == 0)
    e0, e1, em1, em2 = g["extrap"]
    # div_now(-1) = e0 div(0) + e1 div(1) extrapolates linearly to the wall (yp_ext[0] = yu[0])
    assert abs(e0 * lin(yp[0]) + e1 * lin(yp[1]) - lin(yu[0])) <= 1e-12
    assert abs(em1 * lin(yp[-1]) + em2 * lin(yp[-2]) - lin(yu[-1])) <= 1e-12
    assert np.allclose(g["inv_dy_u"], 1 / np.diff(ext))


def _fields(nx, nz, yu, rng, smooth=True):
    """Halo arrays of smooth periodic fields on the staggered positions."""
    ny = len(yu) - 1
    Lx, Lz = 2 * math.pi, math.pi
    x = np.arange(-1, nx + 1) * Lx / nx
    z = np.arange(-1, nz + 1) * Lz / nz
    yp = (yu[:-1] + yu[1:]) / 2
    ypad = np.concatenate([[2 * yu[0] - yp[0]], yp, [2 * yu[-1] - yp[-1]]])
    yf = np.concatenate([[2 * yu[0] - yu[1]], yu])  # V faces: k = -1..ny
    return x, z, ypad, yf, Lx, Lz


def _mk(fun, x, z, y):
    Y, Z, X = np.meshgrid(y, z, x, indexing="ij")
    return fun(X, Z, Y)


def _case(n, incompressible, which, vel=1.0, mus=0.01):
    """Grid n x n/2 x n with a stretched y; returns the max error of the
    chosen right-hand side against its analytic value over interior rows.
    vel scales the velocities and div_now, mus the viscosity: (1, 0.01) is
    convection-dominated, (1e-3, 1) viscous-dominated."""
    nx, nz, ny = n, n // 2, n
    t = np.linspace(0, 1, ny + 1)
    yu = 2 * (t + 0.15 * np.sin(math.pi * t) ** 2 * (t - 0.5))  # smooth stretching on [0, 2]
    g = PR.predictor_geom(yu)
    Lx, Lz = 2 * math.pi, math.pi
    dx, dz = Lx / nx, Lz / nz
    xc = np.arange(-1, nx + 1) * dx + 0.5 * dx  # cell centres (T, P, mu, rho)
    xf = np.arange(-1, nx + 1) * dx             # x faces (U)
    zc = np.arange(-1, nz + 1) * dz + 0.5 * dz
    zf = np.arange(-1, nz + 1) * dz
    yp = (yu[:-1] + yu[1:]) / 2
    ypad = np.concatenate([[2 * yu[0] - yp[0]], yp, [2 * yu[-1] - yp[-1]]])
    yf = np.concatenate([[2 * yu[0] - yu[1]], yu])
    kx, kz = 2 * math.pi / Lx, 2 * math.pi / Lz
    uf = lambda X, Z, Y: vel * np.sin(kx * X) * np.cos(kz * Z) * (1 + 0.3 * Y)
    vf = lambda X, Z, Y: vel * 0.5 * np.cos(kx * X) * np.sin(kz * Z) * np.sin(Y)
    wf = lambda X, Z, Y: vel * np.cos(kx * X + 0.3) * np.sin(kz * Z) * np.cos(0.7 * Y)
    tf = lambda X, Z, Y: 1.5 + 0.2 * np.sin(kx * X) * np.sin(kz * Z + 0.4) * np.cos(Y)
    muf = (lambda X, Z, Y: mus + 0 * X) if incompressible else \
        (lambda X, Z, Y: mus * (1 + 0.3 * np.sin(kx * X + 0.2) * np.cos(kz * Z) * np.exp(-0.3 * Y)))
    rf = (lambda X, Z, Y: 1 + 0 * X) if incompressible else \
        (lambda X, Z, Y: 1 + 0.2 * np.cos(kx * X) * np.sin(kz * Z + 0.1) * (0.5 + 0.2 * Y))
    df = (lambda X, Z, Y: 0 * X) if incompressible else \
        (lambda X, Z, Y: vel * 0.3 * np.sin(kx * X - 0.1) * np.cos(kz * Z) * np.sin(0.5 * Y + 0.2))
    U = _mk(uf, xf, zc, ypad)
    V = _mk(vf, xc, zc, yf)
    Wv = _mk(wf, xc, zf, ypad)
    mu = _mk(muf, xc, zc, ypad)
    rho = _mk(rf, xc, zc, ypad)
    dv = _mk(df, xc, zc, ypad)
    zero = np.zeros_like(U)
    p = dict(dtmax=1.0, inv_dx=1 / dx, inv_dz=1 / dz, nu=mus, Sf_x=0.0, Sf_z=0.0, Gb_x=0.0, Gb_y=0.0, Gb_z=0.0)
    old = [np.zeros((ny, nz, nx)) for _ in range(3)]
    Uh, Vh, Wh = PR.momentum(U, V, Wv, mu, rho, dv, zero, *old, g, p, incompressible, kmin_Vhat=1)
    Ru, Rv, Rw = old  # dt = 1, R_old = 0: hat - value = 1.5 R

    # analytic right-hand sides by finite differences of the analytic fields (h -> 0 at 1e-5)
    h = 1e-5

    def d(fun, axis, X, Z, Y):
        e = [0, 0, 0]
        e[axis] = h
        return (fun(X + e[0], Z + e[1], Y + e[2]) - fun(X - e[0], Z - e[1], Y - e[2])) / (2 * h)

    comps = [uf, wf, vf]  # axis order x, z, y

    def visc(i, X, Z, Y):
        # the continuous operator the variable-property stencils discretise,
        # with the given field div in place of the velocity divergence:
        # mu (lap u_i + 1/3 d_i div) + d_i mu (2 d_i u_i - 2/3 div) + sum_{j != i} d_j mu (d_j u_i + d_i u_j)
        lap = sum(d(lambda X, Z, Y, a=a: d(comps[i], a, X, Z, Y), a, X, Z, Y) for a in range(3))
        out = muf(X, Z, Y) * (lap + d(df, i, X, Z, Y) / 3.0)
        out = out + d(muf, i, X, Z, Y) * (2 * d(comps[i], i, X, Z, Y) - 2.0 / 3.0 * df(X, Z, Y))
        for j in range(3):
            if j != i:
                out = out + d(muf, j, X, Z, Y) * (d(comps[i], j, X, Z, Y) + d(comps[j], i, X, Z, Y))
        return out

    def rhs(i, X, Z, Y):
        conv = uf(X, Z, Y) * d(comps[i], 0, X, Z, Y) + wf(X, Z, Y) * d(comps[i], 1, X, Z, Y) + \
            vf(X, Z, Y) * d(comps[i], 2, X, Z, Y)
        return (-conv + visc(i, X, Z, Y)) / rf(X, Z, Y)

    ks = slice(2, ny - 2)  # rows away from the one-sided wall stencils
    if which == "u":
        Y, Z, X = np.meshgrid(ypad[1:-1][ks], zc[1:-1], xf[1:-1], indexing="ij")
        ref = rhs(0, X, Z, Y)
        got = Ru[ks]
    elif which == "w":
        Y, Z, X = np.meshgrid(ypad[1:-1][ks], zf[1:-1], xc[1:-1], indexing="ij")
        ref = rhs(1, X, Z, Y)
        got = Rw[ks]
    elif which == "v":
        Y, Z, X = np.meshgrid(yf[1:-1][ks], zc[1:-1], xc[1:-1], indexing="ij")
        ref = rhs(2, X, Z, Y)
        got = Rv[ks]
    else:
        T = _mk(tf, xc, zc, ypad)
        Rt = np.zeros((ny, nz, nx))
        cp = np.ones_like(T)
        pt = dict(p, cond=mus, Sq=0.0)
        if incompressible:
            T2 = PR.advance_T(T.copy(), U, V, Wv, rho, cp, None, Rt, g, pt, True)
        else:
            cond = mu.copy()
            divj = np.ones_like(T)
            dv2 = np.zeros_like(T)
            RHS = PR.rhs_energy(T, cond, divj, dv2, g, pt)
            T2 = PR.advance_T(T.copy(), U, V, Wv, rho, cp, RHS, Rt, g, pt, False)
        Y, Z, X = np.meshgrid(ypad[1:-1][ks], zc[1:-1], xc[1:-1], indexing="ij")
        conv = uf(X, Z, Y) * d(tf, 0, X, Z, Y) + wf(X, Z, Y) * d(tf, 1, X, Z, Y) + vf(X, Z, Y) * d(tf, 2, X, Z, Y)

        def flux(a):
            return lambda X, Z, Y: muf(X, Z, Y) * d(tf, a, X, Z, Y)
        cond_term = sum(d(flux(a), a, X, Z, Y) for a in range(3))
        ref = -conv + cond_term / (1.0 if incompressible else rf(X, Z, Y))
        got = Rt[ks]
    return np.abs(got - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("regime", [(1.0, 0.01), (1e-3, 1.0)], ids=["convective", "viscous"])
@pytest.mark.parametrize("incompressible", [True, False])
@pytest.mark.parametrize("which", ["u", "v", "w", "t"])
def test_second_order_convergence(which, incompressible, regime):
    e1 = _case(24, incompressible, which, *regime)
    e2 = _case(48, incompressible, which, *regime)
    assert e1 < 0.15, (which, e1)
    assert e2 < e1 / 3.5, (which, incompressible, e1, e2)


def test_varprop_reduces_to_incompressible():
    rng = np.random.default_rng(5)
    nx, nz = 12, 8
    yu = np.array(W.tanh_faces(10, 1.6))
    ny = len(yu) - 1
    g = PR.predictor_geom(yu)
    shp = (ny + 2, nz + 2, nx + 2)
    U, V, Wv, T = (rng.uniform(-1, 1, shp) for _ in range(4))
    nu = 0.013
    p = dict(dtmax=0.01, inv_dx=3.0, inv_dz=5.0, nu=nu, Sf_x=0.2, Sf_z=-0.1, Gb_x=0.0, Gb_y=0.0, Gb_z=0.0,
             cond=nu, Sq=0.3)
    mu = np.full(shp, nu)
    rho = np.ones(shp)
    zero = np.zeros(shp)
    outs = []
    for inc in (True, False):
        old = [np.zeros((ny, nz, nx)) for _ in range(4)]
        hats = PR.momentum(U, V, Wv, mu, rho, zero, zero, *old[:3], g, p, inc, kmin_Vhat=1)
        if inc:
            T2 = PR.advance_T(T.copy(), U, V, Wv, rho, np.ones(shp), None, old[3], g, p, True)
        else:
            RHS = PR.rhs_energy(T, mu, np.ones(shp), zero.copy(), g, p)
            T2 = PR.advance_T(T.copy(), U, V, Wv, rho, np.ones(shp), RHS, old[3], g, p, False)
        outs.append([h[1:-1, 1:-1, 1:-1] for h in hats] + [T2[1:-1, 1:-1, 1:-1]])
    for a, b in zip(*outs):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max())

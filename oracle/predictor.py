"""CPU restatement of the time-step stages before the projection: the
momentum predictor, the temperature update, the thermophysical property
update and the energy right-hand side (TEST INFRASTRUCTURE ONLY: the parity
checker for the HIP kernels; the product never imports it).

Arrays follow oracle/projection.py: halo arrays (-1:nx, -1:nz, -1:ny) are
numpy (ny+2, nz+2, nx+2) C-order; fields without halo are (ny, nz, nx).
Every expression keeps the reference's evaluation order (left to right,
parentheses as written), so a GPU kernel without FMA contraction reproduces
it bit for bit.

Reference lines restated (src/ unless stated):
  fd_coeffs              pycode/calc_bands.py:84-104 (the three-point quadratic-fit
                         coefficients of fd_coeffs_ref, evaluated at xc = 0)
  predictor_geom         pycode/calc_bands.py:212-220 (extrap_T_wall_*), :272-277 (inv_dy_u),
                         :285-295 (Ruw_*), :309-320 (Rv_*), :345-357 (Rv_interp_uw_bottom_top)
  div_now_walls          mod_cfd_subroutines.f90:120-144
  momentum               mod_cfd_subroutines.f90:147-183 (U_hat), :241-297 (V_hat), :299-354 (W_hat);
                         force_Sf_x_pressure (:184-237) excluded (bulk forcing control)
  advance_T              mod_cfd_subroutines.f90:367-407
  periodicity            mod_cfd_subroutines.f90:860-896 (cfd_periodicity)
  update_props           mod_cfd_subroutines.f90:776-857 with the property laws of the examples'
                         expressions_rho_mu_cond_buoyancy_cp_div_jacobian.f90
  rhs_energy             mod_cfd_subroutines.f90:618-745 (build_rhs_energy; use_utau_work excluded:
                         build_utau_work is a removed legacy routine, :770)
  predictor_step         the order of build_vars, mod_cfd_subroutines.f90:10-468
"""
import math

import numpy as np

from projection import halo_periodic_xz, ibm_exec, ibm_part2, set_val_slice

THIRD = 0.3333333333333333
TWO_THIRDS = 0.6666666666666666


# ---------------------------------------------------------------------------
# geometry (calc_bands.py)
# ---------------------------------------------------------------------------
def fd_coeffs(xs, xn, bc):
    """d/dx and d2/dx2 at xc = 0 of the quadratic through (xs, fs), (0, fc),
    (xn, fn); 'dirich_s' / 'dirich_n' replace the s / n condition by the
    mirrored ghost f(2 xs) = 2 fs - fc (resp. at xn).  Term order of the
    reference's expressions with the xc terms (exact zeros) dropped."""
    if bc == "normal":
        D = xn ** 2 * xs - xn * xs ** 2
        ddx = [xn ** 2 / D, (-xn ** 2 + xs ** 2) / D, (-xs ** 2) / D]
        d2 = [2 * (-xn) / D, 2 * (xn - xs) / D, 2 * (xs) / D]
    elif bc == "dirich_s":
        D = xn ** 2 * xs - 2 * xn * xs ** 2
        ddx = [xn ** 2 / D, (-xn ** 2 + 2 * xs ** 2) / D, (-(2 * xs ** 2)) / D]
        d2 = [2 * (-xn) / D, 2 * (xn - xs) / D, 2 * (xs) / D]
    elif bc == "dirich_n":
        D = -(2 * xn ** 2 * xs) + xn * xs ** 2
        ddx = [(-(2 * xn ** 2)) / D, (2 * xn ** 2 - xs ** 2) / D, xs ** 2 / D]
        d2 = [2 * (xn) / D, 2 * (-xn + xs) / D, 2 * (-xs) / D]
    else:
        raise ValueError(bc)
    return ddx, d2


def predictor_geom(yu):
    yu = [float(v) for v in yu]
    ny = len(yu) - 1
    yp = [(yu[i] + yu[i + 1]) / 2 for i in range(ny)]
    yp_ext = [yu[0]] + yp + [yu[-1]]
    ex = lambda xc, xa, xb: [float((xc - xa) / (xb - xa) * (-1) + 1), float((xc - xa) / (xb - xa))]
    e0, e1 = ex(*yp_ext[:3])
    em1, em2 = ex(*(yp_ext[-3:][::-1]))
    Ruw_ddy, Ruw_d2 = [], []
    for i in range(ny):
        bc = "dirich_s" if i == 0 else ("dirich_n" if i == ny - 1 else "normal")
        a, b = fd_coeffs(yp_ext[i] - yp_ext[i + 1], yp_ext[i + 2] - yp_ext[i + 1], bc)
        Ruw_ddy.append(a)
        Ruw_d2.append(b)
    Rv_ddy, Rv_d2, interp = [], [], []
    for i in range(ny + 1):
        if i in (0, ny):
            Rv_ddy.append([0.0] * 3)
            Rv_d2.append([0.0] * 3)
            interp.append([0.0, 0.0])
        else:
            a, b = fd_coeffs(yu[i - 1] - yu[i], yu[i + 1] - yu[i], "normal")
            Rv_ddy.append(a)
            Rv_d2.append(b)
            dy_top = yp[i] - yu[i]
            dy_bottom = yu[i] - yp[i - 1]
            interp.append([1 - (dy_bottom / (dy_top + dy_bottom)), 1 - (dy_top / (dy_top + dy_bottom))])
    inv_dy_u = [1. / (yp_ext[i + 1] - yp_ext[i]) for i in range(ny + 1)]
    return dict(Ruw_ddy=np.array(Ruw_ddy), Ruw_d2dy2=np.array(Ruw_d2), Rv_ddy=np.array(Rv_ddy),
                Rv_d2dy2=np.array(Rv_d2), Rv_interp=np.array(interp), inv_dy_u=np.array(inv_dy_u),
                extrap=np.array([e0, e1, em1, em2]), yp=np.array(yp))


# ---------------------------------------------------------------------------
# shifted views of halo arrays over a k range of interior points
# ---------------------------------------------------------------------------
class F:
    """A halo array seen from the interior points of planes k0..k1-1:
    F(di, dj, dk) is the (k1-k0, nz, nx) view of A(i+di, j+dj, k+dk)."""

    def __init__(self, A, k0, k1):
        self.A, self.k0, self.k1 = A, k0, k1

    def __call__(self, di, dj, dk):
        ny2, nz2, nx2 = self.A.shape
        return self.A[1 + self.k0 + dk:1 + self.k1 + dk, 1 + dj:nz2 - 1 + dj, 1 + di:nx2 - 1 + di]


def _col(v, k0, k1):
    return np.asarray(v)[k0:k1][:, None, None]


# ---------------------------------------------------------------------------
# momentum predictor
# ---------------------------------------------------------------------------
def momentum(U, V, W, mu, rho, div_now, buoy, Ru_old, Rv_old, Rw_old, g, p, incompressible, kmin_Vhat=1):
    """U_hat, V_hat, W_hat (halo arrays, interior written) and the updated
    R*_old (interior arrays, modified in place)."""
    ny2, nz2, nx2 = U.shape
    ny, nz, nx = ny2 - 2, nz2 - 2, nx2 - 2
    dt = p["dtmax"]
    idx, idz = p["inv_dx"], p["inv_dz"]
    idx2, idz2 = idx * idx, idz * idz
    idx__2, idz__2 = 0.5 * idx, 0.5 * idz
    nu, Sf_x, Sf_z = p["nu"], p["Sf_x"], p["Sf_z"]
    Gb_x, Gb_y, Gb_z = p["Gb_x"], p["Gb_y"], p["Gb_z"]
    U_hat, V_hat, W_hat = np.zeros_like(U), np.zeros_like(V), np.zeros_like(W)

    # U_hat and W_hat: planes 0..ny-1
    k0, k1 = 0, ny
    u, v, w = F(U, k0, k1), F(V, k0, k1), F(W, k0, k1)
    Ra, Rb, Rc = (_col(g["Ruw_ddy"][:, c], k0, k1) for c in range(3))
    R2a, R2b, R2c = (_col(g["Ruw_d2dy2"][:, c], k0, k1) for c in range(3))
    # shared pieces, each in the reference's left-to-right order
    dudx = (u(1, 0, 0) - u(-1, 0, 0)) * idx__2
    dudz = (u(0, 1, 0) - u(0, -1, 0)) * idz__2
    dudy = (u(0, 0, -1) * Ra + u(0, 0, 0) * Rb) + u(0, 0, 1) * Rc
    d2ux = ((u(1, 0, 0) - 2 * u(0, 0, 0)) + u(-1, 0, 0)) * idx2
    d2uz = ((u(0, 1, 0) - 2 * u(0, 0, 0)) + u(0, -1, 0)) * idz2
    d2uy_a, d2uy_b, d2uy_c = u(0, 0, -1) * R2a, u(0, 0, 0) * R2b, u(0, 0, 1) * R2c
    dwdx = (w(1, 0, 0) - w(-1, 0, 0)) * idx__2
    dwdz = (w(0, 1, 0) - w(0, -1, 0)) * idz__2
    dwdy = (w(0, 0, -1) * Ra + w(0, 0, 0) * Rb) + w(0, 0, 1) * Rc
    d2wx = ((w(1, 0, 0) - 2 * w(0, 0, 0)) + w(-1, 0, 0)) * idx2
    d2wz = ((w(0, 1, 0) - 2 * w(0, 0, 0)) + w(0, -1, 0)) * idz2
    d2wy_a, d2wy_b, d2wy_c = w(0, 0, -1) * R2a, w(0, 0, 0) * R2b, w(0, 0, 1) * R2c
    # V at the U point and at the W point (4-point averages)
    v_u = 0.25 * (((v(0, 0, 0) + v(-1, 0, 0)) + v(0, 0, 1)) + v(-1, 0, 1))
    w_u = 0.25 * (((w(0, 0, 0) + w(-1, 0, 0)) + w(0, 1, 0)) + w(-1, 1, 0))
    u_w = 0.25 * (((u(0, 0, 0) + u(0, -1, 0)) + u(1, 0, 0)) + u(1, -1, 0))
    v_w = 0.25 * (((v(0, 0, 0) + v(0, -1, 0)) + v(0, 0, 1)) + v(0, -1, 1))
    if incompressible:
        Ru = (((-(u(0, 0, 0) * dudx) - v_u * dudy) - w_u * dudz)
              + nu * ((((d2ux + d2uy_a) + d2uy_b) + d2uy_c) + d2uz)) + Sf_x
        Rw = (((-(u_w * dwdx) - v_w * dwdy) - w(0, 0, 0) * dwdz)
              + nu * ((((d2wx + d2wy_a) + d2wy_b) + d2wy_c) + d2wz)) + Sf_z
    else:
        m, r, dv, b = F(mu, k0, k1), F(rho, k0, k1), F(div_now, k0, k1), F(buoy, k0, k1)
        # x momentum at the U point (between i-1 and i)
        conv = (u(0, 0, 0) * dudx + w_u * dudz) + v_u * dudy
        D = (((d2ux + d2uz) + ((d2uy_a + d2uy_b) + d2uy_c)) + THIRD * ((dv(0, 0, 0) - dv(-1, 0, 0)) * idx))
        E = 2 * dudx - TWO_THIRDS * (0.5 * (dv(-1, 0, 0) + dv(0, 0, 0)))
        Fx = (m(0, 0, 0) - m(-1, 0, 0)) * idx
        G = dudy + ((0.5 * (v(0, 0, 0) + v(0, 0, 1))) - (0.5 * (v(-1, 0, 0) + v(-1, 0, 1)))) * idx
        H = (((0.5 * (m(-1, 0, -1) + m(0, 0, -1))) * Ra + (0.5 * (m(-1, 0, 0) + m(0, 0, 0))) * Rb)
             + (0.5 * (m(-1, 0, 1) + m(0, 0, 1))) * Rc)
        I = dudz + ((0.5 * (w(0, 0, 0) + w(0, 1, 0))) - (0.5 * (w(-1, 0, 0) + w(-1, 1, 0)))) * idx
        J = ((0.5 * (m(-1, 1, 0) + m(0, 1, 0))) - (0.5 * (m(-1, -1, 0) + m(0, -1, 0)))) * idz__2
        visc = (((0.5 * (m(-1, 0, 0) + m(0, 0, 0))) * D + E * Fx) + G * H) + I * J
        Ru = (((-conv + visc) + Sf_x) + Gb_x * (0.5 * (b(0, 0, 0) + b(-1, 0, 0)))) / (0.5 * (r(-1, 0, 0) + r(0, 0, 0)))
        # z momentum at the W point (between j-1 and j)
        conv = (u_w * dwdx + w(0, 0, 0) * dwdz) + v_w * dwdy
        D = (((d2wx + d2wz) + ((d2wy_a + d2wy_b) + d2wy_c)) + THIRD * ((dv(0, 0, 0) - dv(0, -1, 0)) * idz))
        E = dwdx + ((0.5 * (u(0, 0, 0) + u(1, 0, 0))) - (0.5 * (u(0, -1, 0) + u(1, -1, 0)))) * idz
        Fx = ((0.5 * (m(1, -1, 0) + m(1, 0, 0))) - (0.5 * (m(-1, -1, 0) + m(-1, 0, 0)))) * idx__2
        G = dwdy + ((0.5 * (v(0, 0, 0) + v(0, 0, 1))) - (0.5 * (v(0, -1, 0) + v(0, -1, 1)))) * idz
        H = (((0.5 * (m(0, -1, -1) + m(0, 0, -1))) * Ra + (0.5 * (m(0, -1, 0) + m(0, 0, 0))) * Rb)
             + (0.5 * (m(0, -1, 1) + m(0, 0, 1))) * Rc)
        I = 2 * dwdz - TWO_THIRDS * (0.5 * (dv(0, -1, 0) + dv(0, 0, 0)))
        J = (m(0, 0, 0) - m(0, -1, 0)) * idz
        visc = (((0.5 * (m(0, -1, 0) + m(0, 0, 0))) * D + E * Fx) + G * H) + I * J
        Rw = (((-conv + visc) + Sf_z) + Gb_z * (0.5 * (b(0, 0, 0) + b(0, -1, 0)))) / (0.5 * (r(0, -1, 0) + r(0, 0, 0)))
    Uc, Wc = F(U, k0, k1)(0, 0, 0), F(W, k0, k1)(0, 0, 0)
    U_hat[1 + k0:1 + k1, 1:-1, 1:-1] = Uc + dt * (1.5 * Ru - 0.5 * Ru_old[k0:k1])
    W_hat[1 + k0:1 + k1, 1:-1, 1:-1] = Wc + dt * (1.5 * Rw - 0.5 * Rw_old[k0:k1])
    Ru_old[k0:k1] = Ru
    Rw_old[k0:k1] = Rw

    # V_hat: planes kmin_Vhat..ny-1 (V lives on the y faces)
    k0, k1 = kmin_Vhat, ny
    if k1 > k0:
        u, v, w = F(U, k0, k1), F(V, k0, k1), F(W, k0, k1)
        Va, Vb, Vc = (_col(g["Rv_ddy"][:, c], k0, k1) for c in range(3))
        V2a, V2b, V2c = (_col(g["Rv_d2dy2"][:, c], k0, k1) for c in range(3))
        ib, it = _col(g["Rv_interp"][:, 0], k0, k1), _col(g["Rv_interp"][:, 1], k0, k1)
        idyu = _col(g["inv_dy_u"], k0, k1)
        dvdx = (v(1, 0, 0) - v(-1, 0, 0)) * idx__2
        dvdz = (v(0, 1, 0) - v(0, -1, 0)) * idz__2
        dvdy = (v(0, 0, -1) * Va + v(0, 0, 0) * Vb) + v(0, 0, 1) * Vc
        d2vx = ((v(1, 0, 0) - 2 * v(0, 0, 0)) + v(-1, 0, 0)) * idx2
        d2vz = ((v(0, 1, 0) - 2 * v(0, 0, 0)) + v(0, -1, 0)) * idz2
        d2vy_a, d2vy_b, d2vy_c = v(0, 0, -1) * V2a, v(0, 0, 0) * V2b, v(0, 0, 1) * V2c
        # U and W at the V point: x (z) pair averages, interpolated in y
        u_v = (0.5 * (u(0, 0, -1) + u(1, 0, -1))) * ib + (0.5 * (u(0, 0, 0) + u(1, 0, 0))) * it
        w_v = (0.5 * (w(0, 0, -1) + w(0, 1, -1))) * ib + (0.5 * (w(0, 0, 0) + w(0, 1, 0))) * it
        if incompressible:
            Rv = ((-(u_v * dvdx) - v(0, 0, 0) * dvdy) - w_v * dvdz) + nu * ((((d2vx + d2vy_a) + d2vy_b) + d2vy_c) + d2vz)
        else:
            m, r, dv, b = F(mu, k0, k1), F(rho, k0, k1), F(div_now, k0, k1), F(buoy, k0, k1)
            conv = (u_v * dvdx + w_v * dvdz) + v(0, 0, 0) * dvdy
            D = (((d2vx + d2vz) + ((d2vy_a + d2vy_b) + d2vy_c)) + THIRD * ((dv(0, 0, 0) - dv(0, 0, -1)) * idyu))
            E = dvdx + ((0.5 * (u(0, 0, 0) + u(1, 0, 0))) - (0.5 * (u(0, 0, -1) + u(1, 0, -1)))) * idyu
            Fx = ((0.5 * (m(1, 0, -1) + m(1, 0, 0))) - (0.5 * (m(-1, 0, -1) + m(-1, 0, 0)))) * idx__2
            G = 2 * dvdy - TWO_THIRDS * (0.5 * (dv(0, 0, -1) + dv(0, 0, 0)))
            H = (m(0, 0, 0) - m(0, 0, -1)) * idyu
            I = dvdz + ((0.5 * (w(0, 0, 0) + w(0, 1, 0))) - (0.5 * (w(0, 0, -1) + w(0, 1, -1)))) * idyu
            J = ((0.5 * (m(0, 1, -1) + m(0, 1, 0))) - (0.5 * (m(0, -1, -1) + m(0, -1, 0)))) * idz__2
            visc = (((0.5 * (m(0, 0, -1) + m(0, 0, 0))) * D + E * Fx) + G * H) + I * J
            Rv = ((-conv + visc) + Gb_y * (0.5 * (b(0, 0, 0) + b(0, 0, -1)))) / (0.5 * (r(0, 0, -1) + r(0, 0, 0)))
        Vc_ = F(V, k0, k1)(0, 0, 0)
        V_hat[1 + k0:1 + k1, 1:-1, 1:-1] = Vc_ + dt * (1.5 * Rv - 0.5 * Rv_old[k0:k1])
        Rv_old[k0:k1] = Rv
    return U_hat, V_hat, W_hat


# ---------------------------------------------------------------------------
# temperature
# ---------------------------------------------------------------------------
def advance_T(T, U, V, W, rho, cp, RHS_energy, Rt_old, g, p, incompressible):
    """T (halo array) advanced in place on the interior; Rt_old updated in place."""
    ny2, nz2, nx2 = T.shape
    ny = ny2 - 2
    idx, idz = p["inv_dx"], p["inv_dz"]
    idx2, idz2 = idx * idx, idz * idz
    idx__2, idz__2 = 0.5 * idx, 0.5 * idz
    t, u, v, w = F(T, 0, ny), F(U, 0, ny), F(V, 0, ny), F(W, 0, ny)
    Ra, Rb, Rc = (_col(g["Ruw_ddy"][:, c], 0, ny) for c in range(3))
    R2a, R2b, R2c = (_col(g["Ruw_d2dy2"][:, c], 0, ny) for c in range(3))
    dtdx = (t(1, 0, 0) - t(-1, 0, 0)) * idx__2
    dtdz = (t(0, 1, 0) - t(0, -1, 0)) * idz__2
    dtdy = (t(0, 0, -1) * Ra + t(0, 0, 0) * Rb) + t(0, 0, 1) * Rc
    u_t = 0.5 * (u(0, 0, 0) + u(1, 0, 0))
    v_t = 0.5 * (v(0, 0, 0) + v(0, 0, 1))
    w_t = 0.5 * (w(0, 0, 0) + w(0, 1, 0))
    if incompressible:
        d2tx = ((t(1, 0, 0) - 2 * t(0, 0, 0)) + t(-1, 0, 0)) * idx2
        d2tz = ((t(0, 1, 0) - 2 * t(0, 0, 0)) + t(0, -1, 0)) * idz2
        lap = (((d2tx + t(0, 0, -1) * R2a) + t(0, 0, 0) * R2b) + t(0, 0, 1) * R2c) + d2tz
        Rt = (((-(u_t * dtdx) - v_t * dtdy) - w_t * dtdz) + p["cond"] * lap) + p["Sq"]
    else:
        r, c = F(rho, 0, ny)(0, 0, 0), F(cp, 0, ny)(0, 0, 0)
        Rt = -((u_t * dtdx + w_t * dtdz) + v_t * dtdy) + RHS_energy / (r * c)
    T[1:-1, 1:-1, 1:-1] = T[1:-1, 1:-1, 1:-1] + p["dtmax"] * (1.5 * Rt - 0.5 * Rt_old)
    Rt_old[...] = Rt
    return T


# ---------------------------------------------------------------------------
# periodicity and walls
# ---------------------------------------------------------------------------
def periodicity(U, V, W, T, walls):
    """cfd_periodicity on one rank: x/z halos of all four, then the wall planes."""
    for A in (U, V, W, T):
        halo_periodic_xz(A)
    set_val_slice(U, 0, 0, walls["U_bot"])
    set_val_slice(V, 0, 1, walls["V_bot"])
    set_val_slice(W, 0, 0, walls["W_bot"])
    set_val_slice(T, 0, 0, walls["T_bot"])
    set_val_slice(U, -1, -1, walls["U_top"])
    set_val_slice(V, -1, -1, walls["V_top"])
    set_val_slice(W, -1, -1, walls["W_top"])
    set_val_slice(T, -1, -1, walls["T_top"])


def div_now_walls(div_now, extrap):
    e0, e1, em1, em2 = extrap
    div_now[0] = e0 * div_now[1] + e1 * div_now[2]
    div_now[-1] = em1 * div_now[-2] + em2 * div_now[-3]


# ---------------------------------------------------------------------------
# thermophysical properties
# ---------------------------------------------------------------------------
BASE = ("enth", "rho", "mu", "cond", "cp", "drho_dT")


def law_eval(law, T):
    """law: ("const", c) | ("power", c, p) | ("cubic", segments) with segments
    [(hi, x0, c0, c1, c2, c3), ...] sorted by hi; a point takes the first
    segment with T <= hi (the last one above every hi): the in-order leaves of
    the examples' if (T <= hi) then ... else ... trees."""
    T = np.asarray(T, dtype=np.float64)
    kind = law[0]
    if kind == "const":
        return np.full_like(T, law[1])
    if kind == "power":
        c, pw = law[1], law[2]
        if pw == -1.0:
            return c / T
        if pw == -0.5:
            return c / np.sqrt(T)
        if pw == 0.5:
            return c * np.sqrt(T)
        if pw == 1.0:
            return c * T
        return c * np.power(T, pw)
    if kind == "cubic":
        seg = law[1]
        his = np.array([s[0] for s in seg[:-1]])
        idx = np.searchsorted(his, T, side="left")  # first hi >= T
        out = np.empty_like(T)
        for s, (hi, x0, c0, c1, c2, c3) in enumerate(seg):
            mask = idx == s
            d = T[mask] - x0
            out[mask] = (c0) + (c1) * (d) + (c2) * (d * d) + (c3) * (d * d * d)
        return out
    raise ValueError(kind)


def update_props(T, props, laws, derived, walls, bot=None, top=None):
    """T_ext and the seven properties over the whole padded array.  laws: dict
    over BASE; derived: {"divjac": ("const", c) | ("formula",),
    "buoy": ("none",) | ("linear", b1, b0)}.  bot/top: rough-wall column slabs
    (nx, nz) or None."""
    T_ext = T.copy()
    ny2, nz2, nx2 = T.shape
    ny = ny2 - 2
    if bot is not None:
        for i in range(nx2 - 2):
            for j in range(nz2 - 2):
                kb = int(bot[i, j])
                if kb >= 0:
                    T_ext[1:1 + min(kb, ny - 1) + 1, j + 1, i + 1] = walls["T_bot"]
                kt = int(top[i, j])
                if kt <= ny - 1:
                    T_ext[1 + max(kt, 0):1 + ny, j + 1, i + 1] = walls["T_top"]
        halo_periodic_xz(T_ext)
    vals = {q: law_eval(laws[q], T_ext) for q in BASE}
    props["T_ext"][...] = T_ext
    props["enth"][...] = vals["enth"]
    props["rho"][...] = vals["rho"]
    props["mu"][...] = vals["mu"]
    props["cond"][...] = vals["cond"]
    props["cp"][...] = vals["cp"]
    dj = derived["divjac"]
    if dj[0] == "const":
        props["divjac"][...] = dj[1]
    else:
        props["divjac"][...] = -1.0 / ((vals["rho"] ** 2) * vals["cp"]) * vals["drho_dT"]
    bz = derived["buoy"]
    if bz[0] == "none":
        props["buoy"][...] = 0.0
    else:
        props["buoy"][...] = bz[1] * (vals["rho"] - bz[2])
    return props


# ---------------------------------------------------------------------------
# energy right-hand side and div_now
# ---------------------------------------------------------------------------
def rhs_energy(T, cond, divjac, div_now, g, p, bot=None, top=None):
    """RHS_energy (interior array) and div_now (halo array, in place)."""
    ny2, nz2, nx2 = T.shape
    ny = ny2 - 2
    idx, idz = p["inv_dx"], p["inv_dz"]
    idx2, idz2 = idx * idx, idz * idz
    idx__2, idz__2 = 0.5 * idx, 0.5 * idz
    t, k_ = F(T, 0, ny), F(cond, 0, ny)
    Ra, Rb, Rc = (_col(g["Ruw_ddy"][:, c], 0, ny) for c in range(3))
    R2a, R2b, R2c = (_col(g["Ruw_d2dy2"][:, c], 0, ny) for c in range(3))
    d2tx = ((t(1, 0, 0) - 2 * t(0, 0, 0)) + t(-1, 0, 0)) * idx2
    d2tz = ((t(0, 1, 0) - 2 * t(0, 0, 0)) + t(0, -1, 0)) * idz2
    d2ty = (t(0, 0, -1) * R2a + t(0, 0, 0) * R2b) + t(0, 0, 1) * R2c
    grad = ((((k_(1, 0, 0) - k_(-1, 0, 0)) * idx__2) * ((t(1, 0, 0) - t(-1, 0, 0)) * idx__2)
             + ((k_(0, 1, 0) - k_(0, -1, 0)) * idz__2) * ((t(0, 1, 0) - t(0, -1, 0)) * idz__2))
            + ((k_(0, 0, -1) * Ra + k_(0, 0, 0) * Rb) + k_(0, 0, 1) * Rc)
            * ((t(0, 0, -1) * Ra + t(0, 0, 0) * Rb) + t(0, 0, 1) * Rc))
    RHS = (k_(0, 0, 0) * ((d2tx + d2tz) + d2ty) + grad) + p["Sq"]
    div_now[1:-1, 1:-1, 1:-1] = F(divjac, 0, ny)(0, 0, 0) * RHS
    if bot is not None:
        for i in range(nx2 - 2):
            for j in range(nz2 - 2):
                kb = int(bot[i, j])
                if kb >= 0:
                    div_now[1:1 + min(kb, ny - 1) + 1, j + 1, i + 1] = 0.0
                kt = int(top[i, j])
                if kt <= ny - 1:
                    div_now[1 + max(kt, 0):1 + ny, j + 1, i + 1] = 0.0
    halo_periodic_xz(div_now)
    div_now[0] = 0.0
    div_now[-1] = 0.0
    return RHS


# ---------------------------------------------------------------------------
# the predictor stage of one time step, in build_vars' order (smooth walls)
# ---------------------------------------------------------------------------
def predictor_step(st, g, p, laws=None, derived=None, incompressible=True, kmin_Vhat=1, bot=None, top=None,
                   ibm=None, quady=False):
    """st: dict of arrays (U, V, W, T, mu, rho, div_now, buoy, cp, cond, divjac,
    enth, T_ext halo arrays; Ru_old, Rv_old, Rw_old, Rt_old, RHS_energy interior
    arrays); updated in place; returns (U_hat, V_hat, W_hat).  bot/top: column
    slabs (nx, nz), one pair for every component or dicts per component
    {"U"|"V"|"W"|"T": slab} (the reference's rough_<X>_bot/top_inds:
    ibm_part_2 fills each field's columns from its own pair, update_props and
    build_rhs_energy use T's); ibm: {"U"|"V"|"W"|"T"|"div": (ijk, c)} forcing
    records.  quady (use_rough_surf_quady, mod_cfd_subroutines.f90:99-108 and
    :417-427): with rough walls the opening is ibm_part_2 on U, V, W, T,
    cfd_periodicity, then ibm_part_1; after the predictor ibm_part_2 on the
    hats and T is followed by cfd_periodicity and ibm_part_1 on the hats and
    T, whose last argument .false. leaves the div_now forcing on, then T's
    halos.
    With columns or forcing, build_vars' rough-wall branch (without quady):
    ibm_part_1 opens (mod_cfd_ibm.f90:10-45), ibm_part_2 on the hats and T
    follows the predictor (:47-141), T is forced again in build_rhs_energy
    (mod_cfd_subroutines.f90:640-645)."""
    walls = p["walls"]
    ibm = ibm or {}
    rough = bot is not None or bool(ibm)

    def comp(slab, name):
        return slab.get(name) if isinstance(slab, dict) else slab

    ny2, nz2, nx2 = st["U"].shape
    nx, nz, ny = nx2 - 2, nz2 - 2, ny2 - 2

    def force(name, A):
        if name in ibm:
            A[...] = ibm_exec(A, ibm[name][0], ibm[name][1], nx, nz, ny, 1, 0)

    def columns(fields):
        if bot is None:
            return
        for A, name, kmax in zip(fields, ("U", "V", "W", "T"), (ny - 1, ny, ny - 1, ny - 1)):
            b, t = comp(bot, name), comp(top, name)
            if b is not None:  # (nz, nx) for ibm_part2
                A[...] = ibm_part2(A, np.asarray(b).T, np.asarray(t).T, kmax, walls[name + "_bot"],
                                   walls[name + "_top"])

    def part_1(fields, force_div):  # ibm_part_1 (mod_cfd_ibm.f90:10-45)
        for A, name in zip(fields, ("U", "V", "W", "T")):
            force(name, A)
        for A, name in zip(fields, ("U", "V", "W", "T")):
            set_val_slice(A, 0, 0, walls[name + "_bot"])
        for A, name in zip(fields, ("U", "V", "W", "T")):
            set_val_slice(A, -1, -1, walls[name + "_top"])
        if force_div:
            force("div", st["div_now"])

    if not rough:
        periodicity(st["U"], st["V"], st["W"], st["T"], walls)
    else:
        if quady:
            columns((st["U"], st["V"], st["W"], st["T"]))
            periodicity(st["U"], st["V"], st["W"], st["T"], walls)
        part_1((st["U"], st["V"], st["W"], st["T"]), not incompressible)
    if not incompressible:
        div_now_walls(st["div_now"], g["extrap"])
    U_hat, V_hat, W_hat = momentum(st["U"], st["V"], st["W"], st["mu"], st["rho"], st["div_now"], st["buoy"],
                                   st["Ru_old"], st["Rv_old"], st["Rw_old"], g, p, incompressible, kmin_Vhat)
    advance_T(st["T"], st["U"], st["V"], st["W"], st["rho"], st["cp"], st["RHS_energy"], st["Rt_old"], g, p,
              incompressible)
    if rough:
        columns((U_hat, V_hat, W_hat, st["T"]))
    periodicity(U_hat, V_hat, W_hat, st["T"], walls)
    if rough and quady:
        part_1((U_hat, V_hat, W_hat, st["T"]), True)
        halo_periodic_xz(st["T"])
    bot, top = (None, None) if bot is None else (comp(bot, "T"), comp(top, "T"))
    if not incompressible:
        update_props(st["T"], st, laws, derived, walls, bot, top)
        if "T" in ibm:
            force("T", st["T"])
            set_val_slice(st["T"], 0, 0, walls["T_bot"])
            set_val_slice(st["T"], -1, -1, walls["T_top"])
        st["RHS_energy"][...] = rhs_energy(st["T"], st["cond"], st["divjac"], st["div_now"], g, p, bot, top)
    return U_hat, V_hat, W_hat

"""CPU restatement of the projection stencils and IBM forcing around the
Poisson solve (TEST INFRASTRUCTURE ONLY: the parity checker for the HIP
kernels; the product never imports it).

Arrays:
  * halo arrays, the reference's (-1:nx, -1:nz, -1:ny) fields, are numpy
    (ny+2, nz+2, nx+2) C-order arrays: element (i,j,k) at [k+1, j+1, i+1];
  * P and div-free fields without halo are (ny, nz, nx).
Every arithmetic expression keeps the reference's evaluation order so that a
GPU kernel without FMA contraction reproduces it bit for bit.

Reference lines restated (src/ unless stated):
  geom_coeffs            pycode/calc_bands.py:272-277 (inv_dy_p, inv_dy_p_pad), :336-343 (dPdy_ddy_lh)
  halo_periodic_xz       the one-rank effect of diezdecomp_halos_execute_generic with hl(0:1)
                         periodic in x and z, hl(2) non-periodic in y (inc_set_alloc_arrays.f90:338-356)
  set_val_slice          mod_utils.f90:38-61
  neumann_ghosts_y       mod_cfd_subroutines.f90:1100-1120 (update_vars), :1022-1038 (periodicity_pressure)
  build_p_rhs            mod_cfd_subroutines.f90:527-537 (incompressible), :538-574 (variable density)
  rho_min                mod_cfd_subroutines.f90:470-499
  update_vars            mod_cfd_subroutines.f90:1087-1216 and :1225-1244 (bulk forcing :1217-1221 excluded)
  ibm_read_coeffs        mod_utils.f90:181-212
  ibm_exec               diezdecomp_api_ibm.f90:374-445 with the one-rank index maps of :109-371
  ibm_div_coeffs         ibm_setup/set_ibm_bands.f90:31-81
  ibm_part2              ibm_setup/mod_cfd_ibm.f90:47-141
  ibm_clip_slabs         ibm_setup/set_ibm_bands.f90:151-218
"""
import struct

import numpy as np


# ---------------------------------------------------------------------------
# geometry coefficients (calc_bands.py)
# ---------------------------------------------------------------------------
def geom_coeffs(yu):
    yu = [float(v) for v in yu]
    ny = len(yu) - 1
    yp = [(yu[i] + yu[i + 1]) / 2 for i in range(ny)]
    yp_ext = [yu[0]] + yp + [yu[-1]]
    yp_pad = [2 * yu[0] - yp[0]] + yp + [2 * yu[-1] - yp[-1]]
    inv_dy_p = np.array([1. / (yu[i + 1] - yu[i]) for i in range(ny)])
    inv_dy_u = np.array([1. / (yp_ext[i + 1] - yp_ext[i]) for i in range(ny + 1)])
    inv_dy_p_pad = np.array([1. / (yp_pad[i + 1] - yp_pad[i]) for i in range(ny + 1)])
    lo, hi = [], []
    for i in range(ny + 1):
        if i in (0, ny):
            lo.append(0.0)
            hi.append(0.0)
        else:
            dy = yp[i] - yp[i - 1]
            lo.append(-1 / dy)
            hi.append(1 / dy)
    # read_column_r8 reads the first ny rows (inc_set_alloc_arrays.f90:51-52)
    return dict(inv_dy_p=inv_dy_p, inv_dy_u=inv_dy_u, inv_dy_p_pad=inv_dy_p_pad,
                dPdy_ddy_lo=np.array(lo[:ny]), dPdy_ddy_hi=np.array(hi[:ny]), yp=np.array(yp))


# ---------------------------------------------------------------------------
# halos and wall planes
# ---------------------------------------------------------------------------
def halo_periodic_xz(A):
    """x pass over the whole padded (z, y) range, then z pass over the whole padded (x, y) range."""
    A[:, :, 0] = A[:, :, -2]
    A[:, :, -1] = A[:, :, 1]
    A[:, 0, :] = A[:, -2, :]
    A[:, -1, :] = A[:, 1, :]
    return A


def set_val_slice(A, k0, k1, val):
    """mod_utils.f90:38-61 on a halo array seen with lower bound 0 (so plane 0 is k = -1)."""
    n2 = A.shape[0]
    if k0 < 0:
        k0 += n2
    if k1 < 0:
        k1 += n2
    A[min(k0, k1):max(k0, k1) + 1, :, :] = val
    return A


def neumann_ghosts_y(A):
    A[0, :, :] = A[1, :, :]
    A[-1, :, :] = A[-2, :, :]
    return A


def halo_from_interior(P):
    ny, nz, nx = P.shape
    H = np.zeros((ny + 2, nz + 2, nx + 2))
    H[1:-1, 1:-1, 1:-1] = P
    return H


# ---------------------------------------------------------------------------
# Poisson right-hand side
# ---------------------------------------------------------------------------
def _c(A):
    return A[1:-1, 1:-1, 1:-1]


def _xp(A):
    return A[1:-1, 1:-1, 2:]


def _xm(A):
    return A[1:-1, 1:-1, :-2]


def _zp(A):
    return A[1:-1, 2:, 1:-1]


def _zm(A):
    return A[1:-1, :-2, 1:-1]


def _yp(A):
    return A[2:, 1:-1, 1:-1]


def _ym(A):
    return A[:-2, 1:-1, 1:-1]


def build_p_rhs_incomp(U_hat, W_hat, V_hat, inv_dy_p, inv_dtmax, inv_dx, inv_dz):
    idy = np.asarray(inv_dy_p)[:, None, None]
    return inv_dtmax * (((_xp(U_hat) - _c(U_hat)) * inv_dx + (_zp(W_hat) - _c(W_hat)) * inv_dz) +
                        (_yp(V_hat) - _c(V_hat)) * idy)


def build_p_rhs_varprop(U_hat, W_hat, V_hat, P_now, P_old, rho, div_now, inv_dy_p, inv_dy_p_pad,
                        inv_dtmax, inv_dx, inv_dz, inv_dx2, inv_dz2, rho_min):
    idy = np.asarray(inv_dy_p)[:, None, None]
    ipp = np.asarray(inv_dy_p_pad)
    ipp_hi = ipp[1:][:, None, None]
    ipp_lo = ipp[:-1][:, None, None]
    Ps = 2 * P_now - P_old
    Ps_00 = _c(Ps)
    rc = _c(rho)

    def rfac(nb):
        return 1.0 - rho_min * (0.5 / nb + 0.5 / rc)

    rho_xp, rho_xm = rfac(_xp(rho)), rfac(_xm(rho))
    rho_zp, rho_zm = rfac(_zp(rho)), rfac(_zm(rho))
    rho_yp, rho_ym = rfac(_yp(rho)), rfac(_ym(rho))
    t1 = inv_dtmax * (((_xp(U_hat) - _c(U_hat)) * inv_dx + (_zp(W_hat) - _c(W_hat)) * inv_dz) +
                      (_yp(V_hat) - _c(V_hat)) * idy) * rho_min
    t2 = inv_dtmax * rho_min * _c(div_now)
    t3 = (((rho_xp * (_xp(Ps) - Ps_00) - rho_xm * (Ps_00 - _xm(Ps))) * inv_dx2 +
           (rho_zp * (_zp(Ps) - Ps_00) - rho_zm * (Ps_00 - _zm(Ps))) * inv_dz2) +
          (rho_yp * (_yp(Ps) - Ps_00) * ipp_hi - rho_ym * (Ps_00 - _ym(Ps)) * ipp_lo) * idy)
    return (t1 - t2) + t3


def rho_min(rho):
    return float(_c(rho).min())


# ---------------------------------------------------------------------------
# gradient correction
# ---------------------------------------------------------------------------
WALL_KEYS = ("U_bot", "V_bot", "W_bot", "U_top", "V_top", "W_top")


def make_P_per(P):
    P_per = halo_from_interior(P)
    halo_periodic_xz(P_per)
    return neumann_ghosts_y(P_per)


def update_vars(P, U, V, W, U_hat, V_hat, W_hat, dPdy_ddy_hi, dPdy_ddy_lo, dtmax, inv_dx, inv_dz, kmin_Vhat,
                walls, use_incompressible=True, P_now=None, P_old=None, rho=None, rho_min_v=None):
    """Returns (U, V, W, P_per, P_now, P_old); the inputs are not modified."""
    U, V, W = U.copy(), V.copy(), W.copy()
    P_per = make_P_per(P)
    hi = np.asarray(dPdy_ddy_hi)[:, None, None]
    lo = np.asarray(dPdy_ddy_lo)[:, None, None]
    if use_incompressible:
        U[1:-1, 1:-1, 1:-1] = _c(U_hat) - dtmax * (_c(P_per) - _xm(P_per)) * inv_dx
        W[1:-1, 1:-1, 1:-1] = _c(W_hat) - dtmax * (_c(P_per) - _zm(P_per)) * inv_dz
        Vn = _c(V_hat) - dtmax * (hi * _c(P_per) + lo * _ym(P_per))
    else:
        inv_rho_min = 1.0 / rho_min_v
        Ps = 2 * P_now - P_old
        rc = _c(rho)

        def pf(nb_per, nb_ps, nb_rho):
            inv_rho_now = 0.5 / rc + 0.5 / nb_rho
            P_fwd = _c(P_per) * inv_rho_min + (inv_rho_now - inv_rho_min) * _c(Ps)
            P_bwd = nb_per * inv_rho_min + (inv_rho_now - inv_rho_min) * nb_ps
            return P_fwd, P_bwd

        f, b = pf(_xm(P_per), _xm(Ps), _xm(rho))
        U[1:-1, 1:-1, 1:-1] = _c(U_hat) - dtmax * (f - b) * inv_dx
        f, b = pf(_zm(P_per), _zm(Ps), _zm(rho))
        W[1:-1, 1:-1, 1:-1] = _c(W_hat) - dtmax * (f - b) * inv_dz
        f, b = pf(_ym(P_per), _ym(Ps), _ym(rho))
        Vn = _c(V_hat) - dtmax * (hi * f + lo * b)
        P_old = P_now.copy()
        P_now = P_per.copy()
    V[1 + kmin_Vhat:-1, 1:-1, 1:-1] = Vn[kmin_Vhat:]
    for A in (U, V, W):
        halo_periodic_xz(A)
    set_val_slice(U, 0, 0, walls["U_bot"])
    set_val_slice(V, 0, 1, walls["V_bot"])
    set_val_slice(W, 0, 0, walls["W_bot"])
    set_val_slice(U, -1, -1, walls["U_top"])
    set_val_slice(V, -1, -1, walls["V_top"])
    set_val_slice(W, -1, -1, walls["W_top"])
    return U, V, W, P_per, P_now, P_old


# ---------------------------------------------------------------------------
# IBM
# ---------------------------------------------------------------------------
def ibm_write_coeffs(path, records):
    """records: list of (t_ijk, c1, s1_ijk, c2, s2_ijk, cw) in the 60-byte layout
    of calc_bands.py:150-161 / mk_test_poisson_solver.py:69-74."""
    with open(path, "wb") as fw:
        fw.write(struct.pack("i", len(records)))
        for t, c1, s1, c2, s2, cw in records:
            fw.write(struct.pack("3i", *t))
            fw.write(struct.pack("d", c1))
            fw.write(struct.pack("3i", *s1))
            fw.write(struct.pack("d", c2))
            fw.write(struct.pack("3i", *s2))
            fw.write(struct.pack("d", cw))


def ibm_read_coeffs(path, nx, nz, ny):
    """Returns ijk (n, 3 points, 3) int and c (n, 3) = (cw, c1, c2), indices wrapped mod n."""
    raw = open(path, "rb").read()
    n = struct.unpack_from("i", raw, 0)[0]
    ijk = np.zeros((n, 3, 3), dtype=np.int64)
    c = np.zeros((n, 3))
    off = 4
    for r in range(n):
        t = struct.unpack_from("3i", raw, off); off += 12
        c1 = struct.unpack_from("d", raw, off)[0]; off += 8
        s1 = struct.unpack_from("3i", raw, off); off += 12
        c2 = struct.unpack_from("d", raw, off)[0]; off += 8
        s2 = struct.unpack_from("3i", raw, off); off += 12
        cw = struct.unpack_from("d", raw, off)[0]; off += 8
        for p, v in enumerate((t, s1, s2)):
            ijk[r, p] = (v[0] % nx, v[1] % nz, v[2] % ny)
        c[r] = (cw, c1, c2)
    return ijk, c


def ibm_exec(A, ijk, c, nx, nz, ny, nhalo, nskip):
    """One-rank diezdecomp_ibm_exec on a padded array A (shape (ny, nz, nx) + 2*(nhalo+nskip)).
    Targets are unique (the reference's generator guarantees it), so the
    record loops are written as whole-array operations."""
    A = A.copy()
    side = nhalo + nskip
    ijk = np.asarray(ijk, dtype=np.int64).reshape(-1, 3, 3)
    c = np.asarray(c, dtype=np.float64).reshape(-1, 3)
    t, s1, s2 = ijk[:, 0], ijk[:, 1], ijk[:, 2]

    def at(p):
        return (p[:, 2] + side, p[:, 1] + side, p[:, 0] + side)

    # fluids pass: pack every source, then apply (diezdecomp_api_ibm.f90:381-410, :413-424)
    b1, b2 = A[at(s1)], A[at(s2)]
    A[at(t)] = c[:, 1] * b1 + c[:, 2] * b2 + c[:, 0]
    # forcing pass: targets with a periodic image in the halo shell copy their
    # forced value to every image inside the padded array (:138-149, :252-274)
    n = np.array([nx, nz, ny])
    val = A[at(t)]
    shape = n + 2 * side
    shifts = [np.array(v) for v in ((si, sj, sk) for si in (-1, 0, 1) for sj in (-1, 0, 1) for sk in (-1, 0, 1))]
    halo = np.zeros(len(t), dtype=bool)
    for sh in shifts:
        q = t + sh * n
        inner = np.all((q >= 0) & (q < n), axis=1)
        in_pad = np.all((q >= -nhalo) & (q <= n + nhalo - 1), axis=1)
        halo |= in_pad & ~inner
    for sh in shifts:
        q = t + sh * n
        ok = halo & np.all((q + side >= 0) & (q + side < shape), axis=1)
        A[at(q[ok])] = val[ok]
    return A


def ibm_div_coeffs(ijk, yp, nx, nz, inv_dx, inv_dz):
    """set_ibm_bands.f90:57-77: c = (cw=0, 1 + d_ta/d_ab, -d_ta/d_ab)."""
    dx = 1 / inv_dx
    dz = 1 / inv_dz

    def qp(i0, t, n):
        i = i0
        if abs((i0 - n) - t) < abs(i - t):
            i = i0 - n
        if abs((i0 + n) - t) < abs(i - t):
            i = i0 + n
        return i

    c = np.zeros((len(ijk), 3))
    for r in range(len(ijk)):
        it, jt, kt = (int(v) for v in ijk[r, 0])
        ia = qp(int(ijk[r, 1, 0]), it, nx)
        ja = qp(int(ijk[r, 1, 1]), jt, nz)
        ka = int(ijk[r, 1, 2])
        ib = qp(int(ijk[r, 2, 0]), it, nx)
        jb = qp(int(ijk[r, 2, 1]), jt, nz)
        kb = int(ijk[r, 2, 2])
        d_ab = np.sqrt((dx * (ib - ia)) ** 2 + (dz * (jb - ja)) ** 2 + (yp[kb] - yp[ka]) ** 2)
        d_ta = np.sqrt((dx * (it - ia)) ** 2 + (dz * (jt - ja)) ** 2 + (yp[kt] - yp[ka]) ** 2)
        c[r, 1] = 1 + d_ta / d_ab
        c[r, 2] = -d_ta / d_ab
    return c


def ibm_part2(A, bot, top, kmax, wall_bot, wall_top):
    """Column fill of a halo array: k in [0, bot(i,j)] and [top(i,j), kmax]."""
    A = A.copy()
    nz, nx = bot.shape
    for j in range(nz):
        for i in range(nx):
            for k in range(0, int(bot[j, i]) + 1):
                A[k + 1, j + 1, i + 1] = wall_bot
            for k in range(int(top[j, i]), kmax + 1):
                A[k + 1, j + 1, i + 1] = wall_top
    return A


def ibm_clip_slabs(bot, top, ny, k_first=0):
    """Slab files minus k_first, clipped as set_ibm_bands.f90:196-218."""
    return np.minimum(bot - k_first, ny - 1), np.maximum(top - k_first, 0)

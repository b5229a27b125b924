"""Oracle (test infrastructure only): restatement of the reference's IBM
coefficient engine, src/pycode/ibm_coeffs_engine.f90 (search_mid_h :6-76,
get_contained_alpha :79-89, apply_Fadlun :91-131, make_rough_surface_quady
:133-189, apply_quady :191-210, make_rough_surface_ghost :212-345) and its
driver main_runner_ibm_coeffs_engine.f90 (heights :39-47, passes :56-66,
repetitions and slabs :69-108, sort by (k, j, i) :111-146, files :148-175),
with the height function of the examples' rough_surf_define.f90
(Mod_Height_Function get_h).  Plain Python loops: small grids only.

Where the reference reads y outside 0..ny-1 (y(k-1) at k = 0 and y(k+1) at
k = ny-1 in the quadratic mode, y(k+2)/y(k-2) past the ends, the second
source past the ends in the ghost mode) its result is undefined; here the
first two are taken as -inf/+inf, records whose sources leave the grid are
skipped in the quadratic mode and rejected in the ghost mode.
"""
import math

DIRS = [(0, 0, -1), (0, 0, 1), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (-1, 0, -1), (-1, 0, 1), (1, 0, -1),
        (1, 0, 1), (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1), (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0),
        (-1, -1, -1), (-1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1), (1, -1, 1), (1, 1, -1), (1, 1, 1)]


def height(mnab, Lx, Lz, Ly, x, z, bottom):
    two_pi = 8.0 * math.atan(1.0)
    H = 0.0
    for m, n, a, b in mnab:
        th = two_pi * (m * (x / Lx) + n * (z / Lz))
        H = H + a * math.cos(th) + b * math.sin(th)
    return H if bottom else Ly - H


def search_mid_h(ta, tb, ha, hb, geth):
    a, b = list(ta), list(tb)
    if (a[2] - ha) > (b[2] - hb):
        a, b, ha, hb = list(tb), list(ta), hb, ha
    dx, dz, dy = b[0] - a[0], b[1] - a[1], b[2] - a[2]
    if abs(dx) < 1e-10 and abs(dz) < 1e-10:
        return [0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]), 0.5 * (ha + hb)]
    x0, z0, y0 = a
    assert not ((hb - b[2] - 1e-10) > (ha - a[2]))
    assert not ((b[2] + 1e-10) < hb)
    assert not ((a[2] - 1e-10) > ha)
    lo, hi, iters = 0.0, 1.0, 0
    while (hi - lo) > 1e-10:
        iters += 1
        assert iters < 10000
        mid = 0.5 * (lo + hi)
        new_h = geth(dx * mid + x0, dz * mid + z0) - (dy * mid + y0)
        if (new_h - 1e-10) <= 0:
            hi = mid
        if (new_h + 1e-10) >= 0:
            lo = mid
    mid = 0.5 * (lo + hi)
    xn, zn = dx * mid + x0, dz * mid + z0
    return [xn, zn, geth(xn, zn)]


def contained_alpha(T, A, B, ee, p1):
    nA = nB = nT = 0.0
    for d in range(3):
        nA += (A[d] - p1[d]) * ee[d]
        nB += (B[d] - p1[d]) * ee[d]
        nT += (T[d] - p1[d]) * ee[d]
    return ((nA - 1e-10) <= nT) and (nT <= (nB + 1e-10)), 1 - (nT - nA) / (nB - nA)


def apply_fadlun(t, p1, p2, ht, h2, geth):
    n = search_mid_h(t, p2, ht, h2, geth)
    g = [2 * n[d] - t[d] for d in range(3)]
    nrm = 0.0
    for d in range(3):
        nrm += (p2[d] - p1[d]) * (p2[d] - p1[d])
    ee = [(p2[d] - p1[d]) / nrm for d in range(3)]
    c_n1, a_n = contained_alpha(g, n, p1, ee, p1)
    c_12, a_p1 = contained_alpha(g, p1, p2, ee, p1)
    if c_n1 and c_12:
        assert abs(a_n) < 1e-10 and abs(a_p1 - 1) < 1e-10
    if c_n1:
        return -(1 - a_n), 0.0, 2 - a_n
    assert c_12
    return -a_p1, -(1 - a_p1), 2.0


def apply_quady(y0, ya, yb, hw):
    xw, xa, xb = hw - y0, ya - y0, yb - y0
    den = xa * xa * xb - xa * xa * xw - xa * xb * xb + xa * xw * xw + xb * xb * xw - xb * xw * xw
    cw = (xa * xa * xb - xa * xb * xb) / den
    ca = (xb * xb * xw - xb * xw * xw) / den
    cb = (-xa * xa * xw + xa * xw * xw) / den
    return ca, cb, cw


def ghost_pass(x, z, y, hxz, bottom, bc_w, geth, out):
    nx, nz, ny = len(x), len(z), len(y)
    mag_x = (x[nx - 1] - x[0]) / (nx - 1.0)
    mag_z = (z[nz - 1] - z[0]) / (nz - 1.0)
    fluid = (lambda i, j, k: hxz[i][j] < y[k]) if bottom else (lambda i, j, k: hxz[i][j] > y[k])
    ok = [[[fluid(i, j, k) for k in range(ny)] for j in range(nz)] for i in range(nx)]
    slab = [[-1 if bottom else ny for _ in range(nz)] for _ in range(nx)]
    for dx, dz, dy in DIRS:
        for i in range(nx):
            for j in range(nz):
                for k in range(ny):
                    if not fluid(i, j, k):
                        continue
                    i2, j2, k2 = (i + dx) % nx, (j + dz) % nz, k + dy
                    if not (0 <= k2 < ny) or ok[i2][j2][k2]:
                        continue
                    assert 0 <= k - dy < ny
                    sx = nx if (i + dx) < 0 else (-nx if (i + dx) >= nx else 0)
                    sz = nz if (j + dz) < 0 else (-nz if (j + dz) >= nz else 0)
                    t = (i + dx + sx, j + dz + sz, k + dy)
                    p1 = (i + sx, j + sz, k)
                    p2 = (i - dx + sx, j - dz + sz, k - dy)
                    xt = [x[i] + dx * mag_x, z[j] + dz * mag_z, y[k + dy]]
                    x1 = [x[i], z[j], y[k]]
                    x2 = [x[i] - dx * mag_x, z[j] - dz * mag_z, y[k - dy]]
                    ca, cb, cw = apply_fadlun(xt, x1, x2, hxz[i2][j2], hxz[(i - dx) % nx][(j - dz) % nz], geth)
                    out.append((t, p1, p2, ca, cb, cw * bc_w))
                    slab[t[0]][t[1]] = max(t[2], slab[t[0]][t[1]]) if bottom else min(t[2], slab[t[0]][t[1]])
                    ok[i2][j2][k2] = True
    return slab


def quady_pass(x, z, y, hxz, bottom, bc_w, out):
    nx, nz, ny = len(x), len(z), len(y)
    slab = [[-1 if bottom else ny for _ in range(nz)] for _ in range(nx)]
    for i in range(nx):
        for j in range(nz):
            h = hxz[i][j]
            for k in range(ny):
                if bottom and h < y[k] and (k == 0 or y[k - 1] <= h) and k + 2 < ny:
                    ca, cb, cw = apply_quady(y[k], y[k + 1], y[k + 2], h)
                    out.append(((i, j, k), (i, j, k + 1), (i, j, k + 2), ca, cb, cw * bc_w))
                    slab[i][j] = max(k - 1, slab[i][j])
                if (not bottom) and y[k] < h and (k == ny - 1 or h <= y[k + 1]) and k - 2 >= 0:
                    ca, cb, cw = apply_quady(y[k], y[k - 1], y[k - 2], h)
                    out.append(((i, j, k), (i, j, k - 1), (i, j, k - 2), ca, cb, cw * bc_w))
                    slab[i][j] = min(k + 1, slab[i][j])
    return slab


def generate(mnab, Lx, Lz, Ly, x, z, y, cc_x=1, cc_z=1, quady=False, bc_bot=0.0, bc_top=0.0):
    """Returns (records, slab_bot, slab_top): records (t, p1, p2, ca, cb, cw)
    sorted by (k, j, i); slabs as nested lists [i][j] over the repeated grid."""
    nx, nz = len(x), len(z)
    hb = [[height(mnab, Lx, Lz, Ly, x[i], z[j], True) for j in range(nz)] for i in range(nx)]
    ht = [[height(mnab, Lx, Lz, Ly, x[i], z[j], False) for j in range(nz)] for i in range(nx)]
    recs = []
    if quady:
        sb = quady_pass(x, z, y, hb, True, bc_bot, recs)
        st = quady_pass(x, z, y, ht, False, bc_top, recs)
    else:
        sb = ghost_pass(x, z, y, hb, True, bc_bot, lambda a, b: height(mnab, Lx, Lz, Ly, a, b, True), recs)
        st = ghost_pass(x, z, y, ht, False, bc_top, lambda a, b: height(mnab, Lx, Lz, Ly, a, b, False), recs)
    n0 = len(recs)
    rep = [None] * (n0 * cc_x * cc_z)
    for ii in range(cc_x):
        for jj in range(cc_z):
            for r, (t, p1, p2, ca, cb, cw) in enumerate(recs):
                sh = lambda p: (p[0] + nx * ii, p[1] + nz * jj, p[2])
                rep[r + n0 * (ii * cc_z + jj)] = (sh(t), sh(p1), sh(p2), ca, cb, cw)
    order = sorted(range(len(rep)), key=lambda r: (rep[r][0][2], rep[r][0][1], rep[r][0][0], r))
    recs = [rep[r] for r in order]
    NX, NZ = nx * cc_x, nz * cc_z
    slab_bot = [[sb[I % nx][J % nz] for J in range(NZ)] for I in range(NX)]
    slab_top = [[st[I % nx][J % nz] for J in range(NZ)] for I in range(NX)]
    return recs, slab_bot, slab_top

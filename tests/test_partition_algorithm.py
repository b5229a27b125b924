"""Numerical design check of the partitioned tridiagonal kernel
(tri_group in csrc/poisson_ct.h), emulated in numpy on single columns: P
chunks of MR rows, chunk elimination with fill-in towards both chunk ends,
elimination of every chunk's start row (its own and the next chunk's) from
its end row, the P-row reduced system in the end unknowns solved by parallel
cyclic reduction or by a serial Thomas sweep, back-substitution.  The
two-pass form keeps only the chunk's right-hand side through the reduced
solve: the elimination accumulates x_{s+1} = Pp - Qq x_s - Rr x_e forwards,
and once x_s and x_e are known the chunk interior is solved again as a
tridiagonal system with those two values as boundary data.  Padding rows
(past ny) are a = c = 0, b = -1, as the kernel stages them.

Cyclic reduction can lose digits on nearly singular (low |lambda|) modes, so
the kernel switches to the serial reduced solve below
|lambda| = min_k(a_k + c_k).  This test shows, on the channel grids the
benchmarks use, that the hybrid keeps the residual within a small factor of a
serial Thomas solve of the whole column, for 16, 32, 48 and 64 chunks (48:
cyclic reduction on a chunk count that is not a power of two).
"""
import numpy as np
import pytest

import oracle as O
from dnsproj import workloads as W


def thomas(a, b, c, d):
    n = len(b)
    cp = np.zeros(n)
    dd = d.copy()
    cp[0] = c[0] / b[0]
    dd[0] = dd[0] / b[0]
    for k in range(1, n):
        den = b[k] - a[k] * cp[k - 1]
        cp[k] = c[k] / den
        dd[k] = (dd[k] - a[k] * dd[k - 1]) / den
    for k in range(n - 2, -1, -1):
        dd[k] = dd[k] - dd[k + 1] * cp[k]
    return dd


def partitioned(a, b, c, d, MR, P, reduced, two_pass=False):
    ny = len(b)
    n = MR * P
    A, B, C, D = np.zeros(n), -np.ones(n), np.zeros(n), np.zeros(n)
    A[:ny], B[:ny], C[:ny], D[:ny] = a, b, c, d
    srow, erow, keep = [], [], []
    for p in range(P):
        s = p * MR
        dl = D[s:s + MR].copy()
        Ah, Ch = np.zeros(MR), np.zeros(MR)
        Ak, Bk, Dk, cprev = A[s + 1], B[s + 1], dl[1], C[s + 1]
        Pp, Qq, Rr, pi = 0.0, 0.0, -1.0, 1.0
        for l in range(2, MR):
            r = 1.0 / Bk
            if two_pass:
                Pp, Qq, pi = Pp + pi * (Dk * r), Qq + pi * (Ak * r), -pi * (cprev * r)
            else:
                dl[l - 1], Ah[l - 1], Ch[l - 1] = Dk * r, Ak * r, cprev * r
            f = A[s + l] * r
            Ak, Bk, Dk, cprev = -f * Ak, B[s + l] - f * cprev, dl[l] - f * Dk, C[s + l]
        if two_pass:
            Rr = -pi
        else:
            for l in range(MR - 2, 0, -1):
                Pp, Qq, Rr = dl[l] - Ch[l] * Pp, Ah[l] - Ch[l] * Qq, -Ch[l] * Rr
        cs = C[s]
        sb = B[s] - cs * Qq
        # row s: sa x_{s-1} + sb x_s + sc x_e = sd;  row e: Ae x_s + Be x_e + ce x_{e+1} = De
        srow.append((A[s], sb, 1.0 / sb, -cs * Rr, dl[0] - cs * Pp))
        erow.append((Ak, Bk, C[s + MR - 1], Dk))
        keep.append((dl, Ah, Ch))
    ra, rb, rc, rd = (np.zeros(P) for _ in range(4))
    for p in range(P):
        sa, sb, sib, sc, sd = srow[p]
        Ae, Be, ce, De = erow[p]
        al = -Ae * sib
        ra[p] = al * sa if p > 0 else 0.0
        rb[p] = Be + al * sc
        rd[p] = De + al * sd
        if p + 1 < P:
            na, nb, nib, nc, nd = srow[p + 1]
            g = -ce * nib
            rc[p] = g * nc
            rb[p] += g * na
            rd[p] += g * nd
    if reduced == "serial":
        y = thomas(ra, rb, rc, rd)
    else:
        ib = 1.0 / rb
        st = 1
        while st < P:
            na, nb, nc, nd = (np.zeros(P) for _ in range(4))
            for q in range(P):
                a2, b2, c2, d2 = 0.0, rb[q], 0.0, rd[q]
                if q - st >= 0:
                    f = -ra[q] * ib[q - st]
                    a2, b2, d2 = f * ra[q - st], b2 + f * rc[q - st], d2 + f * rd[q - st]
                if q + st < P:
                    g = -rc[q] * ib[q + st]
                    c2, b2, d2 = g * rc[q + st], b2 + g * ra[q + st], d2 + g * rd[q + st]
                na[q], nb[q], nc[q], nd[q] = a2, b2, c2, d2
            ra, rb, rc, rd = na, nb, nc, nd
            ib = 1.0 / rb
            st *= 2
        y = rd * ib
    x = np.zeros(n)
    for p in range(P):
        s = p * MR
        dl, Ah, Ch = keep[p]
        sa, sb, sib, sc, sd = srow[p]
        ym = y[p - 1] if p > 0 else 0.0
        xs, xe = (sd - sa * ym - sc * y[p]) * sib, y[p]
        x[s], x[s + MR - 1] = xs, xe
        if two_pass:
            if MR > 2:
                rhs = dl[1:MR - 1].copy()
                rhs[0] -= A[s + 1] * xs
                rhs[-1] -= C[s + MR - 2] * xe
                x[s + 1:s + MR - 1] = thomas(A[s + 1:s + MR - 1], B[s + 1:s + MR - 1], C[s + 1:s + MR - 1], rhs)
            continue
        xn = xe
        for l in range(MR - 2, 0, -1):
            xn = dl[l] - Ah[l] * xs - Ch[l] * xn
            x[s + l] = xn
    return x[:ny]


@pytest.mark.parametrize("two_pass", [False, True])
@pytest.mark.parametrize("chunks", [16, 32, 48, 64])
@pytest.mark.parametrize("grid", ["config1", "config2"])
def test_hybrid_reduced_solve_matches_serial_thomas(grid, chunks, two_pass):
    yu = W.config1()["yu"] if grid == "config1" else W.config2_grid()["yu"]
    ny = len(yu) - 1
    MR = max(2, -(-ny // chunks))
    bands = O.bands_from_yu(yu)
    a, b, c = bands[0], bands[1], bands[2]
    pcr_min = np.min(np.abs(a[1:-1]) + np.abs(c[1:-1]))
    rng = np.random.default_rng(7)
    for lam, use00 in [(0.0, True), (-1.0, False), (-0.1 * pcr_min, False), (-pcr_min, False), (-30 * pcr_min, False)]:
        A, B, C = (bands[3], bands[4], bands[5]) if use00 else (a, b, c)
        d = rng.uniform(-1, 1, ny)
        M = np.diag(B + lam) + np.diag(A[1:], -1) + np.diag(C[:-1], 1)
        res = lambda x: np.abs(M @ x - d).max() / np.abs(d).max()
        r_serial = res(thomas(A, B + lam, C, d))
        reduced = "serial" if abs(lam) < pcr_min else "pcr"
        r_hybrid = res(partitioned(A, B + lam, C, d, MR, chunks, reduced, two_pass))
        assert r_hybrid <= 10 * r_serial + 2e-14, (grid, lam, r_hybrid, r_serial)


def test_partitioned_solution_matches_dense_solve():
    # every piece of the elimination (both chunk ends, the exchange of start
    # rows, the reduced solve) against a dense LU solve on a small system,
    # including padded chunks (ny not a multiple of the chunk count)
    rng = np.random.default_rng(3)
    for ny, MR, P in [(37, 2, 32), (64, 2, 32), (100, 4, 32), (97, 2, 64), (384, 12, 32), (384, 8, 48),
                      (300, 8, 48), (41, 2, 48), (384, 24, 16), (500, 32, 16)]:
        a = rng.uniform(0.5, 1.5, ny)
        c = rng.uniform(0.5, 1.5, ny)
        a[0] = 0.0
        c[-1] = 0.0
        b = -(a + c) - rng.uniform(0.1, 1.0, ny)
        d = rng.uniform(-1, 1, ny)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        ref = np.linalg.solve(M, d)
        for reduced in ("pcr", "serial"):
            for two_pass in (False, True):
                x = partitioned(a, b, c, d, MR, P, reduced, two_pass)
                assert np.abs(x - ref).max() <= 1e-13 * np.abs(ref).max(), (ny, MR, P, reduced, two_pass)

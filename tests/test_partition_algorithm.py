"""Numerical design check of the partitioned tridiagonal kernel
(kt_tri_part in csrc/poisson_ct.h), emulated in numpy on single columns:
32 chunks, chunk elimination with fill-in, the 64-row reduced system solved
by parallel cyclic reduction or by a serial Thomas sweep, back-substitution.

Cyclic reduction loses digits on nearly singular (low |lambda|) modes, so the
kernel switches to the serial reduced solve below |lambda| = min_k(a_k + c_k).
This test shows both facts on the channel grids the benchmarks use: the
hybrid keeps the residual within a small factor of a serial Thomas solve of
the whole column, and pure PCR would not at lambda = 0.
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


def partitioned(a, b, c, d, MR, P, reduced):
    ny = len(b)
    n = MR * P
    A, B, C, D = np.zeros(n), np.ones(n), np.zeros(n), np.zeros(n)
    A[:ny], B[:ny], C[:ny], D[:ny] = a, b, c, d
    ra, rb, rc, rd = (np.zeros(2 * P) for _ in range(4))
    keep = []
    for p in range(P):
        s = p * MR
        dl = D[s:s + MR].copy()
        Ah, Ch = np.zeros(MR), np.zeros(MR)
        Ak, Bk, Dk, cprev = A[s + 1], B[s + 1], dl[1], C[s + 1]
        for l in range(2, MR):
            r = 1.0 / Bk
            dl[l - 1], Ah[l - 1], Ch[l - 1] = Dk * r, Ak * r, cprev * r
            f = A[s + l] * r
            Ak, Bk, Dk, cprev = -f * Ak, B[s + l] - f * cprev, dl[l] - f * Dk, C[s + l]
        Pp, Qq, Rr = 0.0, 0.0, -1.0
        for l in range(MR - 2, 0, -1):
            Pp, Qq, Rr = dl[l] - Ch[l] * Pp, Ah[l] - Ch[l] * Qq, -Ch[l] * Rr
        cs = C[s]
        ra[2 * p], rb[2 * p], rc[2 * p], rd[2 * p] = A[s], B[s] - cs * Qq, -cs * Rr, dl[0] - cs * Pp
        ra[2 * p + 1], rb[2 * p + 1], rc[2 * p + 1], rd[2 * p + 1] = Ak, Bk, C[s + MR - 1], Dk
        keep.append((dl, Ah, Ch))
    NR = 2 * P
    if reduced == "serial":
        xr = thomas(ra, rb, rc, rd)
    else:
        st = 1
        while st < NR:
            na, nb, nc, nd = (np.zeros(NR) for _ in range(4))
            for q in range(NR):
                a2, b2, c2, d2 = 0.0, rb[q], 0.0, rd[q]
                if q - st >= 0:
                    al = -ra[q] / rb[q - st]
                    a2, b2, d2 = al * ra[q - st], b2 + al * rc[q - st], d2 + al * rd[q - st]
                if q + st < NR:
                    ga = -rc[q] / rb[q + st]
                    c2, b2, d2 = ga * rc[q + st], b2 + ga * ra[q + st], d2 + ga * rd[q + st]
                na[q], nb[q], nc[q], nd[q] = a2, b2, c2, d2
            ra, rb, rc, rd = na, nb, nc, nd
            st *= 2
        xr = rd / rb
    x = np.zeros(n)
    for p in range(P):
        s = p * MR
        dl, Ah, Ch = keep[p]
        xs, xe = xr[2 * p], xr[2 * p + 1]
        x[s], x[s + MR - 1] = xs, xe
        xn = xe
        for l in range(MR - 2, 0, -1):
            xn = dl[l] - Ah[l] * xs - Ch[l] * xn
            x[s + l] = xn
    return x[:ny]


@pytest.mark.parametrize("grid", ["config1", "config2"])
def test_hybrid_reduced_solve_matches_serial_thomas(grid):
    yu = W.config1()["yu"] if grid == "config1" else W.config2_grid()["yu"]
    ny = len(yu) - 1
    MR = max(2, -(-ny // 32))
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
        r_hybrid = res(partitioned(A, B + lam, C, d, MR, 32, reduced))
        assert r_hybrid <= 10 * r_serial + 2e-14, (grid, lam, r_hybrid, r_serial)
    # pure cyclic reduction at lambda = 0 would not be
    d = rng.uniform(-1, 1, ny)
    M = np.diag(bands[4]) + np.diag(bands[3][1:], -1) + np.diag(bands[5][:-1], 1)
    r_pcr = np.abs(M @ partitioned(bands[3], bands[4], bands[5], d, MR, 32, "pcr") - d).max()
    r_ser = np.abs(M @ thomas(bands[3], bands[4], bands[5], d) - d).max()
    assert r_pcr > 20 * r_ser

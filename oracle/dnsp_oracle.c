/*
 * dnsp_oracle.c -- CPU restatement of the reference's single-rank pressure
 * Poisson path (FFTW branch).  TEST INFRASTRUCTURE ONLY: this file is the
 * parity checker and the timed CPU baseline.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it.  The product library never
 * calls into it.
 *
 * Array layout (same as the reference's Fortran arrays, x fastest):
 *   P(i, j, k) at P[i + nx*(j + nz*k)],  i in [0,nx) x, j in [0,nz) z, k in [0,ny) y.
 * Spectral layout: S(a, b, k) at the same flat position, where a is the
 * FFTW halfcomplex index along x and b the halfcomplex index along z
 * (exactly nx*nz*ny reals; the reference's two padding entries per line,
 * n_fft = 2*(n/2+1), are never read by its inverse transform and are not
 * stored here).
 *
 * Functions and the reference lines they restate:
 *   orc_yp_ext            src/pycode/calc_bands.py:226-227
 *   orc_mk_P_abc          src/pycode/calc_bands.py:181-201
 *   orc_eigen             src/poisson_solver_mgpu_dtdma/poisson_solver_multigpu.f90:132-155 (FFTW layout)
 *   orc_r2hc / orc_hc2r   FFTW_R2HC / FFTW_HC2R definitions used at poisson_solver_multigpu.f90:63-68,89-104
 *                         (own mixed-radix Stockham FFT; FFTW itself is absent from this image)
 *   orc_poisson_forward   poisson_solver_multigpu.f90:360-393 (x R2HC, transpose, z R2HC)
 *   orc_poisson_tridiag   poisson_solver_multigpu.f90:162-179 (poi_C) and :395-422 (sweeps)
 *   orc_poisson_backward  poisson_solver_multigpu.f90:501-551 (z HC2R, transpose, x HC2R, scale)
 *   orc_min_pivot         poisson_solver_multigpu.f90:181-197
 *   orc_get_lo_n_bounds_1d poisson_solver_multigpu.f90:556-575
 *   orc_apply_operator    the discrete operator the solve inverts, in physical space:
 *                         periodic 3-point second differences in x and z (the eigenvalues
 *                         at :147-148 are exactly their symbols) plus the y bands, with the
 *                         (0,0)-mode band set applied to the xz-plane means.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI_L 3.141592653589793238462643383279502884L

typedef struct { double re, im; } cpx;

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, any length                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, nf;
    int fac[64];
    cpx* tw; /* tw[t] = exp(-2*pi*i*t/n) */
} orc_plan;

static int orc_factorize(int n, int* fac) {
    int nf = 0;
    while (n % 4 == 0) { fac[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { fac[nf++] = 2; n /= 2; }
    for (int p = 3; (long)p * p <= n; p += 2)
        while (n % p == 0) { fac[nf++] = p; n /= p; }
    if (n > 1) fac[nf++] = n;
    return nf;
}

static void orc_plan_init(orc_plan* pl, int n) {
    pl->n = n;
    pl->nf = orc_factorize(n, pl->fac);
    pl->tw = (cpx*)malloc(sizeof(cpx) * (size_t)(n > 0 ? n : 1));
    for (int t = 0; t < n; ++t) {
        long double ang = 2.0L * ORC_PI_L * (long double)t / (long double)n;
        pl->tw[t].re = (double)cosl(ang);
        pl->tw[t].im = (double)(-sinl(ang));
    }
}

static void orc_plan_free(orc_plan* pl) { free(pl->tw); pl->tw = NULL; }

/* unnormalised complex DFT; inverse uses exp(+2 pi i ...) */
static void orc_cfft(const orc_plan* pl, cpx* x, cpx* work, int inverse) {
    const int n = pl->n;
    int Ns = 1;
    cpx* src = x;
    cpx* dst = work;
    const double sg = inverse ? -1.0 : 1.0; /* multiplies tw.im */
    for (int s = 0; s < pl->nf; ++s) {
        const int R = pl->fac[s];
        const int m = n / R;
        const int tstep = n / (Ns * R);
        cpx v[R];
        cpx o[R];
        /* j = jb Ns + k in increasing order, k = j mod Ns */
        for (int jb = 0; jb < m / Ns; ++jb)
        for (int k = 0; k < Ns; ++k) {
            const int j = jb * Ns + k;
            for (int r = 0; r < R; ++r) {
                cpx a = src[j + r * m];
                int t = r * k * tstep; /* < n */
                double wr = pl->tw[t].re, wi = sg * pl->tw[t].im;
                v[r].re = a.re * wr - a.im * wi;
                v[r].im = a.re * wi + a.im * wr;
            }
            if (R == 2) {
                o[0].re = v[0].re + v[1].re; o[0].im = v[0].im + v[1].im;
                o[1].re = v[0].re - v[1].re; o[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                double t0r = v[0].re + v[2].re, t0i = v[0].im + v[2].im;
                double t1r = v[0].re - v[2].re, t1i = v[0].im - v[2].im;
                double t2r = v[1].re + v[3].re, t2i = v[1].im + v[3].im;
                double t3r = v[1].re - v[3].re, t3i = v[1].im - v[3].im;
                /* forward: multiply t3 by -i; inverse: by +i */
                double ur = sg * t3i, ui = -sg * t3r;
                o[0].re = t0r + t2r; o[0].im = t0i + t2i;
                o[2].re = t0r - t2r; o[2].im = t0i - t2i;
                o[1].re = t1r + ur;  o[1].im = t1i + ui;
                o[3].re = t1r - ur;  o[3].im = t1i - ui;
            } else {
                const int rs = n / R;
                for (int q = 0; q < R; ++q) {
                    double sr = 0.0, si = 0.0;
                    int qr = 0; /* (q r) mod R */
                    for (int r = 0; r < R; ++r) {
                        int t = qr * rs;
                        double wr = pl->tw[t].re, wi = sg * pl->tw[t].im;
                        sr += v[r].re * wr - v[r].im * wi;
                        si += v[r].re * wi + v[r].im * wr;
                        qr += q;
                        if (qr >= R) qr -= R;
                    }
                    o[q].re = sr; o[q].im = si;
                }
            }
            const int idxD = jb * Ns * R + k;
            for (int q = 0; q < R; ++q) dst[idxD + q * Ns] = o[q];
        }
        cpx* tmp = src; src = dst; dst = tmp;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(cpx) * (size_t)n);
}

/* FFTW_R2HC: out = r0 r1 .. r_{n/2} i_{(n+1)/2-1} .. i1 (unnormalised, e^{-2 pi i}) */
static void orc_r2hc_plan(const orc_plan* pl, const double* in, double* out, cpx* buf, cpx* work) {
    const int n = pl->n;
    for (int t = 0; t < n; ++t) { buf[t].re = in[t]; buf[t].im = 0.0; }
    orc_cfft(pl, buf, work, 0);
    for (int k = 0; k <= n / 2; ++k) out[k] = buf[k].re;
    for (int k = 1; k < (n + 1) / 2; ++k) out[n - k] = buf[k].im;
}

/* FFTW_HC2R: inverse of R2HC without the 1/n factor */
static void orc_hc2r_plan(const orc_plan* pl, const double* in, double* out, cpx* buf, cpx* work) {
    const int n = pl->n;
    buf[0].re = in[0]; buf[0].im = 0.0;
    for (int k = 1; k < (n + 1) / 2; ++k) {
        buf[k].re = in[k];     buf[k].im = in[n - k];
        buf[n - k].re = in[k]; buf[n - k].im = -in[n - k];
    }
    if (n % 2 == 0 && n > 0) { buf[n / 2].re = in[n / 2]; buf[n / 2].im = 0.0; }
    orc_cfft(pl, buf, work, 1);
    for (int t = 0; t < n; ++t) out[t] = buf[t].re;
}

#ifdef ORC_FAST_REAL
/* Two real transforms per complex one (baseline build only; same results up
 * to rounding): z = in0 + i in1, Z = DFT(z), X0(k) = (Z(k) + conj Z(n-k))/2,
 * X1(k) = (Z(k) - conj Z(n-k))/(2i); halfcomplex out as orc_r2hc_plan. */
static void orc_r2hc_pair(const orc_plan* pl, const double* in0, const double* in1, double* out0, double* out1,
                          cpx* buf, cpx* work) {
    const int n = pl->n;
    for (int t = 0; t < n; ++t) { buf[t].re = in0[t]; buf[t].im = in1[t]; }
    orc_cfft(pl, buf, work, 0);
    for (int k = 0; k <= n / 2; ++k) {
        const cpx a = buf[k], b = buf[(n - k) % n];
        out0[k] = 0.5 * (a.re + b.re);
        out1[k] = 0.5 * (a.im + b.im);
        if (k > 0 && k < (n + 1) / 2) {
            out0[n - k] = 0.5 * (a.im - b.im);
            out1[n - k] = 0.5 * (b.re - a.re);
        }
    }
}

/* inverse: z(t) = x0(t) + i x1(t) from the two halfcomplex spectra */
static void orc_hc2r_pair(const orc_plan* pl, const double* in0, const double* in1, double* out0, double* out1,
                          cpx* buf, cpx* work) {
    const int n = pl->n;
    buf[0].re = in0[0]; buf[0].im = in1[0];
    for (int k = 1; k < (n + 1) / 2; ++k) {
        /* X0(k) = in0[k] + i in0[n-k], X1(k) = in1[k] + i in1[n-k]; Z = X0 + i X1 */
        buf[k].re = in0[k] - in1[n - k];     buf[k].im = in0[n - k] + in1[k];
        buf[n - k].re = in0[k] + in1[n - k]; buf[n - k].im = in1[k] - in0[n - k];
    }
    if (n % 2 == 0 && n > 0) { buf[n / 2].re = in0[n / 2]; buf[n / 2].im = in1[n / 2]; }
    orc_cfft(pl, buf, work, 1);
    for (int t = 0; t < n; ++t) { out0[t] = buf[t].re; out1[t] = buf[t].im; }
}
#endif

void orc_r2hc(int n, const double* in, double* out) {
    orc_plan pl; orc_plan_init(&pl, n);
    cpx* buf = (cpx*)malloc(sizeof(cpx) * (size_t)n * 2);
    orc_r2hc_plan(&pl, in, out, buf, buf + n);
    free(buf); orc_plan_free(&pl);
}

void orc_hc2r(int n, const double* in, double* out) {
    orc_plan pl; orc_plan_init(&pl, n);
    cpx* buf = (cpx*)malloc(sizeof(cpx) * (size_t)n * 2);
    orc_hc2r_plan(&pl, in, out, buf, buf + n);
    free(buf); orc_plan_free(&pl);
}

/* complex DFT exposed for tests: x interleaved (re,im), length n */
void orc_cfft_inplace(int n, double* x, int inverse) {
    orc_plan pl; orc_plan_init(&pl, n);
    cpx* work = (cpx*)malloc(sizeof(cpx) * (size_t)n);
    orc_cfft(&pl, (cpx*)x, work, inverse);
    free(work); orc_plan_free(&pl);
}

/* ------------------------------------------------------------------------ */
/* Grid, bands and eigenvalues                                              */
/* ------------------------------------------------------------------------ */

/* calc_bands.py:226-227: yp = cell centres of the faces yu, yp_ext = [yu0, yp..., yu_ny] */
void orc_yp_ext(int ny, const double* yu, double* yp_ext) {
    yp_ext[0] = yu[0];
    for (int i = 0; i < ny; ++i) yp_ext[i + 1] = (yu[i] + yu[i + 1]) / 2;
    yp_ext[ny + 1] = yu[ny];
}

/* calc_bands.py:181-201 */
void orc_mk_P_abc(int ny, const double* yp_ext, int is_dirich_first, double* a, double* b, double* c) {
    double* yp_pad = (double*)malloc(sizeof(double) * (size_t)(ny + 2));
    for (int i = 0; i < ny + 2; ++i) yp_pad[i] = yp_ext[i];
    yp_pad[0] = 2 * yp_ext[0] - yp_ext[1];
    yp_pad[ny + 1] = 2 * yp_ext[ny + 1] - yp_ext[ny];
    for (int i = 0; i < ny; ++i) {
        double x0 = yp_pad[i], x1 = yp_pad[i + 1], x2 = yp_pad[i + 2];
        double dy = 0.5 * (x2 - x0);
        a[i] = 1.0 / ((x1 - x0) * dy);
        c[i] = 1.0 / ((x2 - x1) * dy);
        b[i] = -(a[i] + c[i]);
    }
    if (is_dirich_first) b[0] -= a[0];
    else                 b[0] += a[0];
    b[ny - 1] += c[ny - 1];
    a[0] = 0;
    c[ny - 1] = 0;
    free(yp_pad);
}

/* poisson_solver_multigpu.f90:139-152 (FFTW halfcomplex index -> wavenumber ii) */
static int orc_hc_wavenumber(int i, int n) {
    if (i <= n / 2) return i;
    int ii = (n + 1) / 2 - 1 - (i - (n / 2 + 1));
    return ii > 1 ? ii : 1;
}

void orc_eigen(int n, double inv_d, double* lam, int* is00) {
    const double const_pi = 4.0 * atan(1.0);
    for (int i = 0; i < n; ++i) {
        int ii = orc_hc_wavenumber(i, n);
        double s = sin(ii * const_pi / n);
        lam[i] = -4 * (inv_d * inv_d) * (s * s);
        if (is00) is00[i] = (ii == 0);
    }
}

/* poisson_solver_multigpu.f90:556-575 */
void orc_get_lo_n_bounds_1d(int mpi_divs, int n_glob, int* lo_blocks, int* n_blocks) {
    int n_base = n_glob / mpi_divs;
    for (int i = 0; i < mpi_divs; ++i) n_blocks[i] = n_base;
    for (int i = 0; i < n_glob - n_base * mpi_divs; ++i) n_blocks[i % mpi_divs] += 1;
    int total = 0;
    for (int i = 0; i < mpi_divs; ++i) { lo_blocks[i] = total; total += n_blocks[i]; }
}

/* ------------------------------------------------------------------------ */
/* Poisson solve stages                                                     */
/* ------------------------------------------------------------------------ */

static int orc_nthreads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int orc_tid(void) {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* z transforms of every (a, k) column, R2HC (inverse = 0) or HC2R (1), in
 * place.  Columns are gathered ORC_ZB at a time (adjacent a), so each row j
 * is read and written as one contiguous run; every column goes through the
 * same 1-D transform as alone. */
#define ORC_ZB 8
static void orc_zcols(const orc_plan* pz, int nx, int nz, int ny, double* P, int inverse) {
    const int nt = orc_nthreads();
    const size_t per = (size_t)ORC_ZB * nz * 2 + (size_t)nz * 5;
    double* scratch = (double*)malloc(sizeof(double) * (size_t)nt * per);
    const int nblk = (nx + ORC_ZB - 1) / ORC_ZB;
    const long nwork = (long)nblk * ny;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nwork; ++q) {
        double* s = scratch + (size_t)orc_tid() * per;
        double* in = s;                              /* ORC_ZB columns */
        double* out = s + (size_t)ORC_ZB * nz;      /* ORC_ZB columns */
        double* w = s + (size_t)ORC_ZB * nz * 2;    /* buf (2 nz) + work (2 nz) */
        const int a0 = (int)(q % nblk) * ORC_ZB;
        const long k = q / nblk;
        const int nb = nx - a0 < ORC_ZB ? nx - a0 : ORC_ZB;
        double* base = P + (size_t)k * nx * nz + a0;
        for (int j = 0; j < nz; ++j)
            for (int u = 0; u < nb; ++u) in[(size_t)u * nz + j] = base[(size_t)j * nx + u];
        int u = 0;
#ifdef ORC_FAST_REAL
        for (; u + 1 < nb; u += 2) {
            double *i0 = in + (size_t)u * nz, *o0 = out + (size_t)u * nz;
            if (inverse)
                orc_hc2r_pair(pz, i0, i0 + nz, o0, o0 + nz, (cpx*)w, (cpx*)(w + 2 * nz));
            else
                orc_r2hc_pair(pz, i0, i0 + nz, o0, o0 + nz, (cpx*)w, (cpx*)(w + 2 * nz));
        }
#endif
        for (; u < nb; ++u) {
            if (inverse)
                orc_hc2r_plan(pz, in + (size_t)u * nz, out + (size_t)u * nz, (cpx*)w, (cpx*)(w + 2 * nz));
            else
                orc_r2hc_plan(pz, in + (size_t)u * nz, out + (size_t)u * nz, (cpx*)w, (cpx*)(w + 2 * nz));
        }
        for (int j = 0; j < nz; ++j)
            for (int u = 0; u < nb; ++u) base[(size_t)j * nx + u] = out[(size_t)u * nz + j];
    }
    free(scratch);
}

/* S1-S3: x R2HC of every (j,k) line, then z R2HC of every (a,k) column. In place. */
void orc_poisson_forward(int nx, int nz, int ny, double* P) {
    orc_plan px, pz;
    orc_plan_init(&px, nx);
    orc_plan_init(&pz, nz);
    const int nmax = nx > nz ? nx : nz;
    const int nt = orc_nthreads();
    double* scratch = (double*)malloc(sizeof(double) * (size_t)nt * (size_t)nmax * 6);
    const long nlines_x = (long)nz * ny;
#ifdef ORC_FAST_REAL
    /* line pairs; the scratch holds both lines (2 nmax), buf and work (4 nmax) */
    double* scratch2 = (double*)malloc(sizeof(double) * (size_t)nt * (size_t)nmax * 6);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nlines_x - 1; l += 2) {
        double* s = scratch2 + (size_t)orc_tid() * nmax * 6;
        double* line = P + (size_t)l * nx;
        memcpy(s, line, sizeof(double) * (size_t)nx * 2);
        orc_r2hc_pair(&px, s, s + nx, line, line + nx, (cpx*)(s + 2 * nmax), (cpx*)(s + 4 * nmax));
    }
    free(scratch2);
    for (long l = nlines_x & ~1L; l < nlines_x; ++l) {
#else
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nlines_x; ++l) {
#endif
        double* s = scratch + (size_t)orc_tid() * nmax * 6;
        double* line = P + (size_t)l * nx;
        memcpy(s, line, sizeof(double) * (size_t)nx);
        orc_r2hc_plan(&px, s, line, (cpx*)(s + nmax), (cpx*)(s + 3 * nmax));
    }
    orc_zcols(&pz, nx, nz, ny, P, 0);
    free(scratch);
    orc_plan_free(&px);
    orc_plan_free(&pz);
}

/* S5-S8: z HC2R, x HC2R, then scale by 1/(nx*nz). In place. */
void orc_poisson_backward(int nx, int nz, int ny, double* P) {
    orc_plan px, pz;
    orc_plan_init(&px, nx);
    orc_plan_init(&pz, nz);
    const int nmax = nx > nz ? nx : nz;
    const int nt = orc_nthreads();
    double* scratch = (double*)malloc(sizeof(double) * (size_t)nt * (size_t)nmax * 6);
    orc_zcols(&pz, nx, nz, ny, P, 1);
    const long nlines_x = (long)nz * ny;
    const double inv_nxnz = 1.0 / ((double)nx * nz + 0.0);
#ifdef ORC_FAST_REAL
    double* scratch2 = (double*)malloc(sizeof(double) * (size_t)nt * (size_t)nmax * 6);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nlines_x - 1; l += 2) {
        double* s = scratch2 + (size_t)orc_tid() * nmax * 6;
        double* line = P + (size_t)l * nx;
        memcpy(s, line, sizeof(double) * (size_t)nx * 2);
        orc_hc2r_pair(&px, s, s + nx, line, line + nx, (cpx*)(s + 2 * nmax), (cpx*)(s + 4 * nmax));
        for (int i = 0; i < 2 * nx; ++i) line[i] = line[i] * inv_nxnz;
    }
    free(scratch2);
    for (long l = nlines_x & ~1L; l < nlines_x; ++l) {
#else
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nlines_x; ++l) {
#endif
        double* s = scratch + (size_t)orc_tid() * nmax * 6;
        double* line = P + (size_t)l * nx;
        memcpy(s, line, sizeof(double) * (size_t)nx);
        orc_hc2r_plan(&px, s, line, (cpx*)(s + nmax), (cpx*)(s + 3 * nmax));
        for (int i = 0; i < nx; ++i) line[i] = line[i] * inv_nxnz;
    }
    free(scratch);
    orc_plan_free(&px);
    orc_plan_free(&pz);
}

#define ORC_MB 8
/* S4: per-mode Thomas, poisson_solver_multigpu.f90:162-179 and :395-422.
 * Bands are selected per mode: the *_00 set for the (0,0) mode, else the plain set. */
void orc_poisson_tridiag(int nx, int nz, int ny, double inv_dx, double inv_dz,
                         const double* ba, const double* bb, const double* bc,
                         const double* ba00, const double* bb00, const double* bc00,
                         double* S) {
    double* ax = (double*)malloc(sizeof(double) * (size_t)nx);
    double* az = (double*)malloc(sizeof(double) * (size_t)nz);
    int* is0x = (int*)malloc(sizeof(int) * (size_t)nx);
    int* is0z = (int*)malloc(sizeof(int) * (size_t)nz);
    orc_eigen(nx, inv_dx, ax, is0x);
    orc_eigen(nz, inv_dz, az, is0z);
    const long plane = (long)nx * nz;
    const int nt = orc_nthreads();
    /* modes go ORC_MB at a time (adjacent i of one j), so each row k is one
     * contiguous run; per mode the operations and their order are those of
     * the reference's sweep */
    double* cpall = (double*)malloc(sizeof(double) * (size_t)nt * (size_t)ny * ORC_MB);
    const long nblk = (plane + ORC_MB - 1) / ORC_MB;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nblk; ++q) {
        double* cp = cpall + (size_t)orc_tid() * ny * ORC_MB;
        const long m0 = q * ORC_MB;
        const int nb = plane - m0 < ORC_MB ? (int)(plane - m0) : ORC_MB;
        double axz[ORC_MB];
        const double* Au[ORC_MB];
        const double* Bu[ORC_MB];
        const double* Cu[ORC_MB];
        for (int u = 0; u < nb; ++u) {
            const long m = m0 + u;
            const int i = (int)(m % nx);
            const int j = (int)(m / nx);
            axz[u] = ax[i] + az[j];
            const int is00 = is0x[i] && is0z[j];
            Au[u] = is00 ? ba00 : ba;
            Bu[u] = is00 ? bb00 : bb;
            Cu[u] = is00 ? bc00 : bc;
        }
        double* d = S + m0;
        /* poi_C: init_spec_poisson_multigpu :167-176 */
        for (int u = 0; u < nb; ++u) cp[u] = Cu[u][0] / (Bu[u][0] + axz[u]);
        for (int k = 1; k < ny; ++k)
            for (int u = 0; u < nb; ++u)
                cp[(size_t)k * ORC_MB + u] =
                    Cu[u][k] / (Bu[u][k] + axz[u] - Au[u][k] * cp[(size_t)(k - 1) * ORC_MB + u]);
        /* run_spec_poisson_multigpu :396-421 */
        for (int u = 0; u < nb; ++u) d[u] = d[u] / (Bu[u][0] + axz[u]);
        for (int k = 1; k < ny; ++k) {
            double* dk = d + (size_t)k * plane;
            const double* dm = d + (size_t)(k - 1) * plane;
            for (int u = 0; u < nb; ++u)
                dk[u] = (dk[u] - Au[u][k] * dm[u]) /
                        (Bu[u][k] + axz[u] - Au[u][k] * cp[(size_t)(k - 1) * ORC_MB + u]);
        }
        for (int k = ny - 2; k >= 0; --k) {
            double* dk = d + (size_t)k * plane;
            const double* dp = d + (size_t)(k + 1) * plane;
            for (int u = 0; u < nb; ++u) dk[u] = dk[u] - dp[u] * cp[(size_t)k * ORC_MB + u];
        }
    }
    free(cpall); free(ax); free(az); free(is0x); free(is0z);
}

/* init diagnostic min_abs_div_bacp, poisson_solver_multigpu.f90:181-197 */
double orc_min_pivot(int nx, int nz, int ny, double inv_dx, double inv_dz,
                     const double* ba, const double* bb, const double* bc,
                     const double* ba00, const double* bb00, const double* bc00) {
    double* ax = (double*)malloc(sizeof(double) * (size_t)nx);
    double* az = (double*)malloc(sizeof(double) * (size_t)nz);
    int* is0x = (int*)malloc(sizeof(int) * (size_t)nx);
    int* is0z = (int*)malloc(sizeof(int) * (size_t)nz);
    orc_eigen(nx, inv_dx, ax, is0x);
    orc_eigen(nz, inv_dz, az, is0z);
    double mn = 1e20;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            const double a_xz = ax[i] + az[j];
            const int is00 = is0x[i] && is0z[j];
            const double* A = is00 ? ba00 : ba;
            const double* B = is00 ? bb00 : bb;
            const double* C = is00 ? bc00 : bc;
            double cp = C[0] / (B[0] + a_xz);
            double v = fabs(B[0] + a_xz);
            if (v < mn) mn = v;
            for (int k = 1; k < ny; ++k) {
                double den = B[k] + a_xz - A[k] * cp;
                v = fabs(den);
                if (v < mn) mn = v;
                cp = C[k] / den;
            }
        }
    free(ax); free(az); free(is0x); free(is0z);
    return mn;
}

/* whole single-rank solve, in place: P holds the RHS on entry, the solution on exit */
int orc_poisson_solve(int nx, int nz, int ny, double inv_dx, double inv_dz,
                      const double* ba, const double* bb, const double* bc,
                      const double* ba00, const double* bb00, const double* bc00,
                      double* P) {
    if (nx < 1 || nz < 1 || ny < 2) return 1; /* alloc_poisson_solver_multigpu.f90:34-35 */
    orc_poisson_forward(nx, nz, ny, P);
    orc_poisson_tridiag(nx, nz, ny, inv_dx, inv_dz, ba, bb, bc, ba00, bb00, bc00, P);
    orc_poisson_backward(nx, nz, ny, P);
    return 0;
}

/* L p in physical space: the operator whose inverse the spectral solve applies.
 *   (L p)(i,j,k) = inv_dx^2 (p[i+1]-2p[i]+p[i-1]) + inv_dz^2 (p[j+1]-2p[j]+p[j-1])   (periodic)
 *                + a_k p[k-1] + b_k p[k] + c_k p[k+1]
 *                + sum over the xz-plane means m_k of ((a00-a)_k m[k-1] + (b00-b)_k m[k] + (c00-c)_k m[k+1])
 * The last line applies the (0,0)-mode band set to the (0,0) mode only. */
void orc_apply_operator(int nx, int nz, int ny, double inv_dx, double inv_dz,
                        const double* ba, const double* bb, const double* bc,
                        const double* ba00, const double* bb00, const double* bc00,
                        const double* p, double* out) {
    const long plane = (long)nx * nz;
    double* mean = (double*)calloc((size_t)ny, sizeof(double));
    for (int k = 0; k < ny; ++k) {
        long double s = 0.0L;
        for (long m = 0; m < plane; ++m) s += p[(size_t)k * plane + m];
        mean[k] = (double)(s / (long double)plane);
    }
    const double ix2 = inv_dx * inv_dx, iz2 = inv_dz * inv_dz;
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < (long)ny * nz; ++kk) {
        const int k = (int)(kk / nz), j = (int)(kk % nz);
        const int jm = (j + nz - 1) % nz, jp = (j + 1) % nz;
        const double* pk = p + (size_t)k * plane;
        double corr = (bb00[k] - bb[k]) * mean[k];
        if (k > 0) corr += (ba00[k] - ba[k]) * mean[k - 1];
        if (k < ny - 1) corr += (bc00[k] - bc[k]) * mean[k + 1];
        for (int i = 0; i < nx; ++i) {
            const int im = (i + nx - 1) % nx, ip = (i + 1) % nx;
            const double c0 = pk[(size_t)j * nx + i];
            double v = ix2 * (pk[(size_t)j * nx + ip] - 2 * c0 + pk[(size_t)j * nx + im]) +
                       iz2 * (pk[(size_t)jp * nx + i] - 2 * c0 + pk[(size_t)jm * nx + i]) + bb[k] * c0;
            if (k > 0) v += ba[k] * p[(size_t)(k - 1) * plane + (size_t)j * nx + i];
            if (k < ny - 1) v += bc[k] * p[(size_t)(k + 1) * plane + (size_t)j * nx + i];
            out[(size_t)k * plane + (size_t)j * nx + i] = v + corr;
        }
    }
    free(mean);
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) { return orc_nthreads(); }

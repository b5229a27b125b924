// IBM forcing coefficients of a rough wall (host only): the reference's
// coefficient engine (src/pycode/ibm_coeffs_engine.f90 with its driver
// main_runner_ibm_coeffs_engine.f90), for a height function given as a sum of
// Fourier modes (the examples' rough_surf_define.f90, Mod_Height_Function):
//   H(x, z) = sum_i a_i cos(t_i) + b_i sin(t_i),  t_i = 2 pi (m_i x/Lx + n_i z/Lz),
// bottom wall at H, top wall at Ly - H.
//
// Ghost mode (make_rough_surface_ghost :250-345): for each of the 26
// neighbour directions in the reference's order, every fluid point whose
// neighbour in that direction is solid and not yet forced makes that
// neighbour a target; sources are the fluid point and the next one away from
// the target.  apply_Fadlun (:91-131): the wall point on the segment target ->
// second source by bisection (search_mid_h :6-76), the target's mirror image
// through it, and linear interpolation of the image from the wall and the
// sources: target = ca*u1 + cb*u2 + cw*u_wall.
// Quadratic mode (make_rough_surface_quady :133-189, apply_quady :191-210):
// the first fluid point of every column is the target, its value the
// quadratic through the wall and the next two fluid points.
// The generator then repeats the records over cc_x x cc_z surface tiles,
// sorts them by (k, j, i) and writes the records and the column slabs.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

using namespace dnsp;

namespace {

struct Surf {
    const dnsp_surface* s;
    double h(double x, double z, bool bottom) const {
        const double two_pi = 8.0 * std::atan(1.0);
        double H = 0.0;
        for (int i = 0; i < s->n_modes; ++i) {
            const double m = s->mnab[4 * i], n = s->mnab[4 * i + 1], a = s->mnab[4 * i + 2], b = s->mnab[4 * i + 3];
            const double th = two_pi * (m * (x / s->Lx) + n * (z / s->Lz));
            H = H + a * std::cos(th) + b * std::sin(th);
        }
        return bottom ? H : s->Ly - H;
    }
};

struct Rec {
    int t[3], p1[3], p2[3];
    double ca, cb, cw;
};

// Ghost-mode record before its coefficients: coordinates of target and
// sources and the column heights search_mid_h starts from.
struct Cand {
    double xt[3], x1[3], x2[3], ht, h2;
};

// search_mid_h: the wall point between a (target) and b (second source).
int search_mid_h(const double* oa, const double* ob, double oha, double ohb, const Surf& S, bool bottom, double* n) {
    double a[3], b[3], ha = oha, hb = ohb;
    std::memcpy(a, oa, sizeof(a));
    std::memcpy(b, ob, sizeof(b));
    if ((a[2] - ha) > (b[2] - hb)) {
        std::memcpy(a, ob, sizeof(a));
        std::memcpy(b, oa, sizeof(b));
        ha = ohb;
        hb = oha;
    }
    const double dx = b[0] - a[0], dz = b[1] - a[1], dy = b[2] - a[2];
    if (std::fabs(dx) < 1e-10 && std::fabs(dz) < 1e-10) {
        n[0] = 0.5 * (a[0] + b[0]);
        n[1] = 0.5 * (a[1] + b[1]);
        n[2] = 0.5 * (ha + hb);
        return DNSP_OK;
    }
    const double x0 = a[0], z0 = a[1], y0 = a[2];
    if ((hb - b[2] - 1e-10) > (ha - a[2])) {
        set_error("IBM generation: eval_h(1) > eval_h(0) + 1e-10");
        return DNSP_ERR_ARG;
    }
    if ((b[2] + 1e-10) < hb) {
        set_error("IBM generation: second source below the surface");
        return DNSP_ERR_ARG;
    }
    if ((a[2] - 1e-10) > ha) {
        set_error("IBM generation: target above the surface");
        return DNSP_ERR_ARG;
    }
    double lo = 0.0, hi = 1.0, mid;
    int iters = 0;
    while ((hi - lo) > 1e-10) {
        if (++iters >= 10000) {
            set_error("IBM generation: bisection did not converge");
            return DNSP_ERR_ARG;
        }
        mid = 0.5 * (lo + hi);
        const double nh = S.h(dx * mid + x0, dz * mid + z0, bottom) - (dy * mid + y0);
        if ((nh - 1e-10) <= 0) hi = mid;
        if ((nh + 1e-10) >= 0) lo = mid;
    }
    mid = 0.5 * (lo + hi);
    n[0] = dx * mid + x0;
    n[1] = dz * mid + z0;
    n[2] = S.h(n[0], n[1], bottom);
    return DNSP_OK;
}

void contained_alpha(const double* T, const double* A, const double* B, const double* ee, const double* p1,
                     bool* contained, double* alpha) {
    double nA = 0, nB = 0, nT = 0;
    for (int d = 0; d < 3; ++d) {
        nA += (A[d] - p1[d]) * ee[d];
        nB += (B[d] - p1[d]) * ee[d];
        nT += (T[d] - p1[d]) * ee[d];
    }
    *contained = ((nA - 1e-10) <= nT) && (nT <= (nB + 1e-10));
    *alpha = 1 - (nT - nA) / (nB - nA);
}

int apply_fadlun(const double* t, const double* p1, const double* p2, double ht, double h2, const Surf& S,
                 bool bottom, double* ca, double* cb, double* cw) {
    double n[3];
    int rc = search_mid_h(t, p2, ht, h2, S, bottom, n);
    if (rc) return rc;
    double g[3], ee[3];
    double nrm = 0;
    for (int d = 0; d < 3; ++d) {
        g[d] = 2 * n[d] - t[d];
        nrm += (p2[d] - p1[d]) * (p2[d] - p1[d]);
    }
    for (int d = 0; d < 3; ++d) ee[d] = (p2[d] - p1[d]) / nrm;
    bool c_n1, c_12;
    double a_n, a_p1;
    contained_alpha(g, n, p1, ee, p1, &c_n1, &a_n);
    contained_alpha(g, p1, p2, ee, p1, &c_12, &a_p1);
    if (c_n1 && c_12 && !(std::fabs(a_n) < 1e-10 && std::fabs(a_p1 - 1) < 1e-10)) {
        set_error("IBM generation: image on both segments with inconsistent weights");
        return DNSP_ERR_ARG;
    }
    if (c_n1) {
        *ca = -(1 - a_n);
        *cb = 0;
        *cw = 2 - a_n;
    } else {
        if (!c_12) {
            set_error("IBM generation: image outside the source segment");
            return DNSP_ERR_ARG;
        }
        *ca = -a_p1;
        *cb = -(1 - a_p1);
        *cw = 2;
    }
    return DNSP_OK;
}

void apply_quady(double y0, double ya, double yb, double hw, double* ca, double* cb, double* cw) {
    const double xw = hw - y0, xa = ya - y0, xb = yb - y0;
    const double den = xa * xa * xb - xa * xa * xw - xa * xb * xb + xa * xw * xw + xb * xb * xw - xb * xw * xw;
    *cw = (xa * xa * xb - xa * xb * xb) / den;
    *ca = (xb * xb * xw - xb * xw * xw) / den;
    *cb = (-xa * xa * xw + xa * xw * xw) / den;
}

// the 26 directions of make_rough_surface_ghost, in its order
const int kDirs[26][3] = {{0, 0, -1},  {0, 0, 1},   {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},  {-1, 0, -1},
                          {-1, 0, 1},  {1, 0, -1},  {1, 0, 1},   {0, -1, -1}, {0, -1, 1},  {0, 1, -1}, {0, 1, 1},
                          {-1, -1, 0}, {-1, 1, 0},  {1, -1, 0},  {1, 1, 0},   {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1},
                          {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};

int ghost_pass(const Surf& S, int nx, int nz, int ny, const double* x, const double* z, const double* y,
               const std::vector<double>& hxz, bool bottom, double bc_w, std::vector<Rec>& out,
               std::vector<int>& slab) {
    const double mag_x = (x[nx - 1] - x[0]) / (nx - 1.0);
    const double mag_z = (z[nz - 1] - z[0]) / (nz - 1.0);
    if (std::fabs(mag_x - (x[1] - x[0])) > 1e-10 || std::fabs(mag_z - (z[1] - z[0])) > 1e-10) {
        set_error("IBM generation: Mismatch in mag_x, mag_z (x and z must be uniform)");
        return DNSP_ERR_ARG;
    }
    auto H = [&](int i, int j) { return hxz[(size_t)i * nz + j]; };
    auto fluid = [&](int i, int j, int k) { return bottom ? (H(i, j) < y[k]) : (H(i, j) > y[k]); };
    std::vector<char> ok((size_t)nx * nz * ny);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j)
            for (int k = 0; k < ny; ++k) ok[((size_t)i * nz + j) * ny + k] = fluid(i, j, k);
    std::fill(slab.begin(), slab.end(), bottom ? -1 : ny);
    const size_t first = out.size();
    std::vector<Cand> cand;
    for (const auto& dv : kDirs) {
        const int dx = dv[0], dz = dv[1], dy = dv[2];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < nz; ++j)
                for (int k = 0; k < ny; ++k) {
                    if (!fluid(i, j, k)) continue;
                    const int i2 = ((i + dx) % nx + nx) % nx, j2 = ((j + dz) % nz + nz) % nz, k2 = k + dy;
                    if (k2 < 0 || k2 >= ny || ok[((size_t)i2 * nz + j2) * ny + k2]) continue;
                    if (k - dy < 0 || k - dy >= ny) {
                        set_error("IBM generation: second source outside the grid (i=%d j=%d k=%d)", i, j, k);
                        return DNSP_ERR_ARG;
                    }
                    const int sx = (i + dx) < 0 ? nx : ((i + dx) >= nx ? -nx : 0);
                    const int sz = (j + dz) < 0 ? nz : ((j + dz) >= nz ? -nz : 0);
                    Rec r;
                    r.t[0] = i + dx + sx;
                    r.t[1] = j + dz + sz;
                    r.t[2] = k + dy;
                    r.p1[0] = i + sx;
                    r.p1[1] = j + sz;
                    r.p1[2] = k;
                    r.p2[0] = i - dx + sx;
                    r.p2[1] = j - dz + sz;
                    r.p2[2] = k - dy;
                    Cand cd = {{x[i] + dx * mag_x, z[j] + dz * mag_z, y[k + dy]},
                               {x[i], z[j], y[k]},
                               {x[i] - dx * mag_x, z[j] - dz * mag_z, y[k - dy]},
                               H(i2, j2),
                               H(((i - dx) % nx + nx) % nx, ((j - dz) % nz + nz) % nz)};
                    r.ca = r.cb = r.cw = 0.0;
                    int& sl = slab[(size_t)r.t[0] * nz + r.t[1]];
                    sl = bottom ? std::max(r.t[2], sl) : std::min(r.t[2], sl);
                    out.push_back(r);
                    cand.push_back(cd);
                    ok[((size_t)i2 * nz + j2) * ny + k2] = 1;
                }
    }
    // coefficients: independent per record, computed on all host threads
    const size_t n = cand.size();
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<int> err(nt, DNSP_OK);
    std::vector<size_t> err_at(nt, n);
    std::vector<std::string> msg(nt);
    auto work = [&](unsigned w) {
        for (size_t r = first + w; r < first + n; r += nt) {
            const Cand& cd = cand[r - first];
            double ca, cb, cw;
            const int rc = apply_fadlun(cd.xt, cd.x1, cd.x2, cd.ht, cd.h2, S, bottom, &ca, &cb, &cw);
            if (rc) {
                err[w] = rc;
                err_at[w] = r;
                msg[w] = dnsp_last_error();
                return;
            }
            out[r].ca = ca;
            out[r].cb = cb;
            out[r].cw = cw * bc_w;
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nt; ++w) th.emplace_back(work, w);
        for (auto& t : th) t.join();
    }
    size_t bad = n;
    int rc = DNSP_OK;
    for (unsigned w = 0; w < nt; ++w)
        if (err[w] && err_at[w] < bad) {
            bad = err_at[w];
            rc = err[w];
            set_error("%s (record %zu)", msg[w].c_str(), bad);
        }
    return rc;
}

void quady_pass(int nx, int nz, int ny, const double* y, const std::vector<double>& hxz, bool bottom, double bc_w,
                std::vector<Rec>& out, std::vector<int>& slab) {
    std::fill(slab.begin(), slab.end(), bottom ? -1 : ny);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            const double h = hxz[(size_t)i * nz + j];
            for (int k = 0; k < ny; ++k) {
                // y(-1) and y(ny) lie outside the grid: taken as -inf / +inf
                if (bottom && h < y[k] && (k == 0 || y[k - 1] <= h) && k + 2 < ny) {
                    Rec r = {{i, j, k}, {i, j, k + 1}, {i, j, k + 2}, 0, 0, 0};
                    apply_quady(y[k], y[k + 1], y[k + 2], h, &r.ca, &r.cb, &r.cw);
                    r.cw *= bc_w;
                    slab[(size_t)i * nz + j] = std::max(k - 1, slab[(size_t)i * nz + j]);
                    out.push_back(r);
                }
                if (!bottom && y[k] < h && (k == ny - 1 || h <= y[k + 1]) && k - 2 >= 0) {
                    Rec r = {{i, j, k}, {i, j, k - 1}, {i, j, k - 2}, 0, 0, 0};
                    apply_quady(y[k], y[k - 1], y[k - 2], h, &r.ca, &r.cb, &r.cw);
                    r.cw *= bc_w;
                    slab[(size_t)i * nz + j] = std::min(k + 1, slab[(size_t)i * nz + j]);
                    out.push_back(r);
                }
            }
        }
}

}  // namespace

extern "C" {

int dnsp_ibm_generate(const dnsp_surface* surf, int nx, int nz, int ny, const double* x, const double* z,
                      const double* y, int cc_x, int cc_z, int quady, double bc_bot, double bc_top, int cap, int* n,
                      int* ijk, double* c, int* slab_bot, int* slab_top) {
    clear_error();
    DNSP_ARG_CHECK(surf && (surf->n_modes == 0 || surf->mnab) && x && z && y && n, "NULL argument");
    DNSP_ARG_CHECK(nx >= 2 && nz >= 2 && ny >= 3 && cc_x >= 1 && cc_z >= 1, "invalid sizes");
    DNSP_ARG_CHECK(surf->Lx > 0 && surf->Lz > 0, "surface periods must be > 0");
    for (int k = 1; k < ny; ++k) DNSP_ARG_CHECK(y[k] > y[k - 1], "y must increase");
    const Surf S{surf};
    std::vector<double> hb((size_t)nx * nz), ht((size_t)nx * nz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            hb[(size_t)i * nz + j] = S.h(x[i], z[j], true);
            ht[(size_t)i * nz + j] = S.h(x[i], z[j], false);
        }
    std::vector<Rec> recs;
    std::vector<int> sb((size_t)nx * nz), st((size_t)nx * nz);
    if (quady) {
        quady_pass(nx, nz, ny, y, hb, true, bc_bot, recs, sb);
        quady_pass(nx, nz, ny, y, ht, false, bc_top, recs, st);
    } else {
        int rc = ghost_pass(S, nx, nz, ny, x, z, y, hb, true, bc_bot, recs, sb);
        if (!rc) rc = ghost_pass(S, nx, nz, ny, x, z, y, ht, false, bc_top, recs, st);
        if (rc) return rc;
    }
    // repetitions (copy block ii*cc_z + jj), then a stable sort by (k, j, i)
    const size_t n0 = recs.size();
    std::vector<Rec> all(n0 * cc_x * cc_z);
    for (int ii = 0; ii < cc_x; ++ii)
        for (int jj = 0; jj < cc_z; ++jj)
            for (size_t r = 0; r < n0; ++r) {
                Rec q = recs[r];
                for (int* p : {q.t, q.p1, q.p2}) {
                    p[0] += nx * ii;
                    p[1] += nz * jj;
                }
                all[r + n0 * ((size_t)ii * cc_z + jj)] = q;
            }
    std::vector<size_t> order(all.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const Rec& A = all[a];
        const Rec& B = all[b];
        if (A.t[2] != B.t[2]) return A.t[2] < B.t[2];
        if (A.t[1] != B.t[1]) return A.t[1] < B.t[1];
        return A.t[0] < B.t[0];
    });
    if (all.size() > (size_t)0x7fffffff) {
        set_error("too many IBM records");
        return DNSP_ERR_ARG;
    }
    *n = (int)all.size();
    if (!ijk) return DNSP_OK;  // size query
    DNSP_ARG_CHECK(c && cap >= *n, "capacity %d < %d records", cap, *n);
    for (size_t r = 0; r < order.size(); ++r) {
        const Rec& q = all[order[r]];
        int* o = ijk + 9 * r;
        for (int d = 0; d < 3; ++d) {
            o[d] = q.t[d];
            o[3 + d] = q.p1[d];
            o[6 + d] = q.p2[d];
        }
        c[3 * r] = q.cw;
        c[3 * r + 1] = q.ca;
        c[3 * r + 2] = q.cb;
    }
    const int NX = nx * cc_x, NZ = nz * cc_z;
    for (int I = 0; I < NX; ++I)
        for (int J = 0; J < NZ; ++J) {
            const size_t src = (size_t)(I % nx) * nz + (J % nz);
            if (slab_bot) slab_bot[(size_t)J * NX + I] = sb[src];
            if (slab_top) slab_top[(size_t)J * NX + I] = st[src];
        }
    return DNSP_OK;
}

int dnsp_ibm_write_coeffs(const char* path, int n, const int* ijk, const double* c) {
    clear_error();
    DNSP_ARG_CHECK(path && n >= 0 && (n == 0 || (ijk && c)), "invalid arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error("cannot create %s", path);
        return DNSP_ERR_ARG;
    }
    bool ok = std::fwrite(&n, sizeof(int), 1, f) == 1;
    for (int r = 0; r < n && ok; ++r) {
        // t(3 int32) c1 p1(3) c2 p2(3) cw: 60 bytes (main_runner_ibm_coeffs_engine.f90:148-156)
        ok = std::fwrite(ijk + 9 * r, sizeof(int), 3, f) == 3 && std::fwrite(c + 3 * r + 1, sizeof(double), 1, f) == 1 &&
             std::fwrite(ijk + 9 * r + 3, sizeof(int), 3, f) == 3 &&
             std::fwrite(c + 3 * r + 2, sizeof(double), 1, f) == 1 &&
             std::fwrite(ijk + 9 * r + 6, sizeof(int), 3, f) == 3 && std::fwrite(c + 3 * r, sizeof(double), 1, f) == 1;
    }
    std::fclose(f);
    if (!ok) {
        set_error("%s: write failed", path);
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

int dnsp_ibm_write_slab(const char* path, int nx, int nz, const int* slab) {
    clear_error();
    DNSP_ARG_CHECK(path && slab && nx >= 1 && nz >= 1, "invalid arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error("cannot create %s", path);
        return DNSP_ERR_ARG;
    }
    const int hdr[2] = {nx, nz};
    const bool ok = std::fwrite(hdr, sizeof(int), 2, f) == 2 &&
                    std::fwrite(slab, sizeof(int), (size_t)nx * nz, f) == (size_t)nx * nz;
    std::fclose(f);
    if (!ok) {
        set_error("%s: write failed", path);
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

}  // extern "C"

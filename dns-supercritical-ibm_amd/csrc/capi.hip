// Host-only part of the C ABI: error state, version, and the two host
// helpers of the reference path (block partition and y bands).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"

namespace dnsp {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

}  // namespace dnsp

using namespace dnsp;

extern "C" {

const char* dnsp_last_error(void) { return g_err; }

const char* dnsp_version(void) { return "dnsproj 0.1.0 (gfx950)"; }

int dnsp_stream_synchronize(void* stream) {
    clear_error();
    DNSP_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return DNSP_OK;
}

// poisson_solver_multigpu.f90:556-575: n_glob split in mpi_divs blocks, the
// remainder going one each to the first blocks.
int dnsp_get_lo_n_bounds_1d(int mpi_divs, int n_glob, int* lo_blocks_mpi, int* n_blocks_mpi) {
    clear_error();
    DNSP_ARG_CHECK(mpi_divs >= 1 && n_glob >= 0, "mpi_divs=%d n_glob=%d", mpi_divs, n_glob);
    DNSP_ARG_CHECK(lo_blocks_mpi && n_blocks_mpi, "NULL output");
    const int n_base = n_glob / mpi_divs;
    for (int i = 0; i < mpi_divs; ++i) n_blocks_mpi[i] = n_base;
    for (int i = 0; i < n_glob - n_base * mpi_divs; ++i) n_blocks_mpi[i % mpi_divs] += 1;
    int total = 0;
    for (int i = 0; i < mpi_divs; ++i) {
        lo_blocks_mpi[i] = total;
        total += n_blocks_mpi[i];
    }
    if (total != n_glob) {
        set_error("sum(n_blocks_mpi).ne.n_glob");
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

// calc_bands.py:181-201.  The cell centres yp and yp_ext = [yu0, yp, yu_ny]
// follow calc_bands.py:226-227; the ghost centres mirror the first and last
// centres about the walls.
int dnsp_mk_P_abc(int ny, const double* yu, int is_dirich_first, double* a, double* b, double* c) {
    clear_error();
    DNSP_ARG_CHECK(ny >= 1 && yu && a && b && c, "invalid arguments to dnsp_mk_P_abc");
    std::vector<double> ext((size_t)ny + 2);
    ext[0] = yu[0];
    for (int i = 0; i < ny; ++i) ext[(size_t)i + 1] = (yu[i] + yu[i + 1]) / 2;
    ext[(size_t)ny + 1] = yu[ny];
    std::vector<double> pad(ext);
    pad[0] = 2 * ext[0] - ext[1];
    pad[(size_t)ny + 1] = 2 * ext[(size_t)ny + 1] - ext[(size_t)ny];
    for (int i = 0; i < ny; ++i) {
        const double x0 = pad[(size_t)i], x1 = pad[(size_t)i + 1], x2 = pad[(size_t)i + 2];
        const double dy = 0.5 * (x2 - x0);
        a[i] = 1.0 / ((x1 - x0) * dy);
        c[i] = 1.0 / ((x2 - x1) * dy);
        b[i] = -(a[i] + c[i]);
    }
    if (is_dirich_first) b[0] -= a[0];
    else b[0] += a[0];
    b[ny - 1] += c[ny - 1];
    a[0] = 0;
    c[ny - 1] = 0;
    for (int i = 0; i < ny; ++i)
        if (!std::isfinite(a[i]) || !std::isfinite(b[i]) || !std::isfinite(c[i])) {
            set_error("non-finite band at row %d (coincident y faces?)", i);
            return DNSP_ERR_ARG;
        }
    return DNSP_OK;
}

}  // extern "C"

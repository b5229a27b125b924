// Pressure-projection step on one GPU: the part of the reference's time step
// between the predictor and the next predictor (src/main.f90 time loop):
//   rho_min (mod_cfd_subroutines.f90:470-499), Poisson right-hand side
//   (build_vars :527-574), pressure Poisson solve (run_spec_poisson_multigpu),
//   gradient correction update_vars (:1048-1245), then on rough walls the
//   part of the next build_vars' opening that acts on U, V, W (:98-113): the
//   rough-wall columns (ibm_part_2, src/ibm_setup/mod_cfd_ibm.f90:47-141)
//   with cfd_periodicity where column slabs are set, then the IBM forcing
//   and wall planes (ibm_part_1 :10-45).
// The object owns every field at its reference layout on the device and
// enqueues the kernels of the existing entry points on one stream; it can
// record the sequence once into a graph and replay it.  Nothing in the
// sequence allocates, copies through the host or synchronises.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"

using namespace dnsp;

struct dnsp_step {
    int nx = 0, nz = 0, ny = 0;
    double inv_dx = 0, inv_dz = 0, dtmax = 0;
    int incompressible = 0;
    int device = 0;
    dnsp_poisson* poisson = nullptr;
    double* f[DNSP_STEP_NFIELDS] = {};  // halo fields, then P (interior)
    double* d_geom = nullptr;           // inv_dy_p (ny), inv_dy_p_pad (ny+1), hi (ny), lo (ny)
    double* d_rho_min = nullptr;        // 1
    double* d_walls = nullptr;          // 6: U_bot, V_bot, W_bot, U_top, V_top, W_top
    double walls[6] = {0, 0, 0, 0, 0, 0};
    dnsp_ibm* ibm[3] = {nullptr, nullptr, nullptr};  // U, V, W (borrowed)
    int* d_bot[3] = {};                 // rough-wall column slabs of U, V, W (nx*nz each), or NULL
    int* d_top[3] = {};
    size_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

size_t halo_count(const dnsp_step* s) { return (size_t)(s->nx + 2) * (s->nz + 2) * (s->ny + 2); }

void release(dnsp_step* s) {
    if (!s) return;
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (double*& p : s->f) (void)hipFree(p);
    (void)hipFree(s->d_geom);
    (void)hipFree(s->d_rho_min);
    (void)hipFree(s->d_walls);
    for (int c = 0; c < 3; ++c) {
        (void)hipFree(s->d_bot[c]);
        (void)hipFree(s->d_top[c]);
    }
    if (s->poisson) dnsp_poisson_free(s->poisson);
    delete s;
}

// The sequence; refinements < 0: FP64 solve, else the mixed form with that
// many refinement sweeps (0 or 1: no right-hand-side copy).
int enqueue(dnsp_step* s, int refinements, hipStream_t st) {
    const int nx = s->nx, nz = s->nz, ny = s->ny;
    double* const* F = s->f;
    const double* inv_dy_p = s->d_geom;
    const double* inv_dy_p_pad = s->d_geom + ny;
    const double* hi = s->d_geom + 2 * ny + 1;
    const double* lo = s->d_geom + 3 * ny + 1;
    void* vs = reinterpret_cast<void*>(st);
    double* P = F[DNSP_STEP_P];
    int rc = DNSP_OK;
    if (s->incompressible) {
        rc = dnsp_build_p_rhs(P, F[DNSP_STEP_U_HAT], F[DNSP_STEP_W_HAT], F[DNSP_STEP_V_HAT], inv_dy_p, nx, nz, ny,
                              1.0 / s->dtmax, s->inv_dx, s->inv_dz, vs);
    } else {
        rc = dnsp_rho_min(F[DNSP_STEP_RHO], nx, nz, ny, s->d_rho_min, vs);
        if (!rc)
            rc = dnsp_build_p_rhs_varprop(P, F[DNSP_STEP_U_HAT], F[DNSP_STEP_W_HAT], F[DNSP_STEP_V_HAT],
                                          F[DNSP_STEP_P_NOW], F[DNSP_STEP_P_OLD], F[DNSP_STEP_RHO],
                                          F[DNSP_STEP_DIV_NOW], inv_dy_p, inv_dy_p_pad, s->d_rho_min, nx, nz, ny,
                                          1.0 / s->dtmax, s->inv_dx, s->inv_dz, s->inv_dx * s->inv_dx,
                                          s->inv_dz * s->inv_dz, vs);
    }
    if (rc) return rc;
    rc = refinements < 0 ? dnsp_poisson_run(s->poisson, P, vs) : dnsp_poisson_run_mixed(s->poisson, P, refinements, vs);
    if (rc) return rc;
    rc = dnsp_update_vars(P, F[DNSP_STEP_P_PER], F[DNSP_STEP_P_NOW], F[DNSP_STEP_P_OLD], F[DNSP_STEP_U],
                          F[DNSP_STEP_V], F[DNSP_STEP_W], F[DNSP_STEP_U_HAT], F[DNSP_STEP_V_HAT], F[DNSP_STEP_W_HAT],
                          F[DNSP_STEP_RHO], hi, lo, s->d_rho_min, s->d_walls, nx, nz, ny, 1, s->dtmax, s->inv_dx,
                          s->inv_dz, s->incompressible, vs);
    if (rc) return rc;
    // Rough walls: the opening of the next build_vars on U, V, W
    // (mod_cfd_subroutines.f90:98-113), so the step ends on the velocities the
    // next predictor reads.  With column slabs (the quadratic-wall ordering,
    // use_rough_surf_quady) ibm_part_2 fills each component's columns and
    // cfd_periodicity restores halos and walls first; then ibm_part_1 forces
    // U, V, W and sets their wall ghost planes (mod_cfd_ibm.f90:26-38).  The
    // forcing comes last, so no column fill overwrites a forced value.
    const int comp[3] = {DNSP_STEP_U, DNSP_STEP_V, DNSP_STEP_W};
    const bool columns = s->d_bot[0] || s->d_bot[1] || s->d_bot[2];
    const bool forcing = s->ibm[0] || s->ibm[1] || s->ibm[2];
    if (columns) {
        // each component's own columns; kmax = ny-1 for U, W and ny for V
        // (mod_cfd_ibm.f90:47-141)
        for (int c = 0; c < 3 && !rc; ++c)
            if (s->d_bot[c])
                rc = dnsp_ibm_part2(F[comp[c]], s->d_bot[c], s->d_top[c], nx, nz, ny, c == 1 ? ny : ny - 1,
                                    s->walls[c], s->walls[3 + c], vs);
        // cfd_periodicity (:860-899): halos, then U, W at k = -1, V at k = -1
        // and 0, and all three at k = ny take the wall values
        for (int c = 0; c < 3 && !rc; ++c) rc = dnsp_halo_periodic_xz(F[comp[c]], nx, nz, ny, vs);
        for (int c = 0; c < 3 && !rc; ++c) {
            rc = dnsp_set_val_slice(F[comp[c]], nx, nz, ny, 0, c == 1 ? 1 : 0, s->walls[c], vs);
            if (!rc) rc = dnsp_set_val_slice(F[comp[c]], nx, nz, ny, -1, -1, s->walls[3 + c], vs);
        }
        if (rc) return rc;
    }
    if (forcing) {
        for (int c = 0; c < 3 && !rc; ++c)
            if (s->ibm[c]) rc = dnsp_ibm_exec(s->ibm[c], F[comp[c]], vs);
        // ibm_part_1's wall planes: k = -1 and k = ny of U, V, W
        for (int c = 0; c < 3 && !rc; ++c) {
            rc = dnsp_set_val_slice(F[comp[c]], nx, nz, ny, 0, 0, s->walls[c], vs);
            if (!rc) rc = dnsp_set_val_slice(F[comp[c]], nx, nz, ny, -1, -1, s->walls[3 + c], vs);
        }
    }
    return rc;
}

}  // namespace

extern "C" {

int dnsp_step_create(int nx, int nz, int ny, const double* yu, double inv_dx, double inv_dz, double dtmax,
                     int use_incompressible, int precision, dnsp_step** out) {
    clear_error();
    DNSP_ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    DNSP_ARG_CHECK(nx >= 1 && nz >= 1 && ny >= 2 && yu, "invalid sizes or NULL faces");
    DNSP_ARG_CHECK(std::isfinite(dtmax) && dtmax > 0, "dtmax must be finite and > 0");
    dnsp_step* s = new (std::nothrow) dnsp_step();
    if (!s) {
        set_error("host allocation failed");
        return DNSP_ERR_NOMEM;
    }
    s->nx = nx;
    s->nz = nz;
    s->ny = ny;
    s->inv_dx = inv_dx;
    s->inv_dz = inv_dz;
    s->dtmax = dtmax;
    s->incompressible = use_incompressible ? 1 : 0;
    int rc = DNSP_OK;
    do {
        if (hipGetDevice(&s->device) != hipSuccess) {
            set_error("hipGetDevice failed");
            rc = DNSP_ERR_HIP;
            break;
        }
        // y bands of both sets from the faces (calc_bands.py:181-201)
        std::vector<double> b((size_t)6 * ny);
        rc = dnsp_mk_P_abc(ny, yu, 0, b.data(), b.data() + ny, b.data() + 2 * ny);
        if (!rc) rc = dnsp_mk_P_abc(ny, yu, 1, b.data() + 3 * ny, b.data() + 4 * ny, b.data() + 5 * ny);
        if (!rc)
            rc = dnsp_poisson_init(nx, nz, ny, inv_dx, inv_dz, b.data(), b.data() + ny, b.data() + 2 * ny,
                                   b.data() + 3 * ny, b.data() + 4 * ny, b.data() + 5 * ny, precision, &s->poisson);
        if (rc) break;
        std::vector<double> g((size_t)4 * ny + 1);
        rc = dnsp_geom_coeffs(ny, yu, g.data(), nullptr, g.data() + ny, g.data() + 3 * ny + 1, g.data() + 2 * ny + 1);
        if (rc) break;
        const size_t hc = halo_count(s), pc = (size_t)nx * nz * ny;
        bool ok = true;
        for (int i = 0; i < DNSP_STEP_NFIELDS && ok; ++i) {
            const size_t n = i == DNSP_STEP_P ? pc : hc;
            ok = hipMalloc(&s->f[i], sizeof(double) * n) == hipSuccess &&
                 hipMemset(s->f[i], 0, sizeof(double) * n) == hipSuccess;
            s->bytes += sizeof(double) * n;
        }
        ok = ok && hipMalloc(&s->d_geom, sizeof(double) * g.size()) == hipSuccess &&
             hipMemcpy(s->d_geom, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice) == hipSuccess &&
             hipMalloc(&s->d_rho_min, sizeof(double)) == hipSuccess &&
             hipMemset(s->d_rho_min, 0, sizeof(double)) == hipSuccess &&
             hipMalloc(&s->d_walls, sizeof(double) * 6) == hipSuccess &&
             hipMemset(s->d_walls, 0, sizeof(double) * 6) == hipSuccess;
        if (!ok) {
            set_error("device allocation of the step fields failed");
            rc = DNSP_ERR_NOMEM;
            break;
        }
        s->bytes += sizeof(double) * (g.size() + 7);
    } while (0);
    if (rc) {
        release(s);
        return rc;
    }
    *out = s;
    return DNSP_OK;
}

int dnsp_step_info(const dnsp_step* s, int* nx, int* nz, int* ny, double* inv_dx, double* inv_dz, double* dtmax,
                   int* use_incompressible) {
    clear_error();
    DNSP_ARG_CHECK(s, "NULL step");
    if (nx) *nx = s->nx;
    if (nz) *nz = s->nz;
    if (ny) *ny = s->ny;
    if (inv_dx) *inv_dx = s->inv_dx;
    if (inv_dz) *inv_dz = s->inv_dz;
    if (dtmax) *dtmax = s->dtmax;
    if (use_incompressible) *use_incompressible = s->incompressible;
    return DNSP_OK;
}

int dnsp_step_field(dnsp_step* s, int which, double** ptr) {
    clear_error();
    DNSP_ARG_CHECK(s && ptr, "NULL argument");
    DNSP_ARG_CHECK(which >= 0 && which < DNSP_STEP_NFIELDS, "field %d", which);
    *ptr = s->f[which];
    return DNSP_OK;
}

static size_t field_count(const dnsp_step* s, int which) {
    return which == DNSP_STEP_P ? (size_t)s->nx * s->nz * s->ny : halo_count(s);
}

int dnsp_step_write(dnsp_step* s, int which, const double* src, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(s && src, "NULL argument");
    DNSP_ARG_CHECK(which >= 0 && which < DNSP_STEP_NFIELDS, "field %d", which);
    DNSP_HIP_TRY(hipMemcpyAsync(s->f[which], src, sizeof(double) * field_count(s, which), hipMemcpyDeviceToDevice,
                                as_stream(stream)));
    return DNSP_OK;
}

int dnsp_step_read(dnsp_step* s, int which, double* dst, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(s && dst, "NULL argument");
    DNSP_ARG_CHECK(which >= 0 && which < DNSP_STEP_NFIELDS, "field %d", which);
    DNSP_HIP_TRY(hipMemcpyAsync(dst, s->f[which], sizeof(double) * field_count(s, which), hipMemcpyDeviceToDevice,
                                as_stream(stream)));
    return DNSP_OK;
}

int dnsp_step_set_walls(dnsp_step* s, const double* wall_bc) {
    clear_error();
    DNSP_ARG_CHECK(s && wall_bc, "NULL argument");
    std::memcpy(s->walls, wall_bc, sizeof(s->walls));
    DNSP_HIP_TRY(hipMemcpy(s->d_walls, wall_bc, sizeof(double) * 6, hipMemcpyHostToDevice));
    return DNSP_OK;
}

int dnsp_step_set_ibm(dnsp_step* s, int component, dnsp_ibm* ibm) {
    clear_error();
    DNSP_ARG_CHECK(s && component >= 0 && component < 3, "NULL step or component %d", component);
    s->ibm[component] = ibm;
    return DNSP_OK;
}

int dnsp_step_set_component_columns(dnsp_step* s, int component, const int* bot, const int* top) {
    clear_error();
    DNSP_ARG_CHECK(s && component >= 0 && component < 3, "NULL step or component %d", component);
    return set_slab_pair(&s->d_bot[component], &s->d_top[component], bot, top, (size_t)s->nx * s->nz, &s->bytes);
}

int dnsp_step_set_columns(dnsp_step* s, const int* bot, const int* top) {
    clear_error();
    DNSP_ARG_CHECK(s, "NULL step");
    for (int c = 0; c < 3; ++c) {
        const int rc = dnsp_step_set_component_columns(s, c, bot, top);
        if (rc) return rc;
    }
    return DNSP_OK;
}

int dnsp_step_poisson(dnsp_step* s, dnsp_poisson** h) {
    clear_error();
    DNSP_ARG_CHECK(s && h, "NULL argument");
    *h = s->poisson;
    return DNSP_OK;
}

int dnsp_step_footprint(const dnsp_step* s, size_t* bytes) {
    clear_error();
    DNSP_ARG_CHECK(s && bytes, "NULL argument");
    // the step's fields, geometry and slabs, plus every array of its solver
    *bytes = s->bytes + poisson_device_bytes(s->poisson);
    return DNSP_OK;
}

int dnsp_step_enqueue(dnsp_step* s, int refinements, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(s, "NULL step");
    DNSP_ARG_CHECK(refinements >= -1 && refinements <= 1, "refinements=%d: -1 (FP64), 0 or 1", refinements);
    return enqueue(s, refinements, as_stream(stream));
}

int dnsp_step_capture(dnsp_step* s, int refinements, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(s, "NULL step");
    DNSP_ARG_CHECK(refinements >= -1 && refinements <= 1, "refinements=%d: -1 (FP64), 0 or 1", refinements);
    hipStream_t st = as_stream(stream);
    DNSP_ARG_CHECK(st != nullptr, "capture needs a non-null stream");
    if (s->exec) {
        (void)hipGraphExecDestroy(s->exec);
        s->exec = nullptr;
    }
    if (s->graph) {
        (void)hipGraphDestroy(s->graph);
        s->graph = nullptr;
    }
    DNSP_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(s, refinements, st);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    DNSP_HIP_TRY(ec);
    s->graph = g;
    DNSP_HIP_TRY(hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0));
    return DNSP_OK;
}

int dnsp_step_launch(dnsp_step* s, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(s && s->exec, "no captured step (dnsp_step_capture)");
    DNSP_HIP_TRY(hipGraphLaunch(s->exec, as_stream(stream)));
    return DNSP_OK;
}

int dnsp_step_free(dnsp_step* s) {
    clear_error();
    release(s);
    return DNSP_OK;
}

}  // extern "C"

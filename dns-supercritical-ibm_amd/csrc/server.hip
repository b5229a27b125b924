// Serving engine: projection-step requests for an ensemble of channel cases,
// continuously batched on one GPU.
//
// A request names a case and a seed.  Its inputs (predictor velocities,
// density, previous pressures, divergence) are generated on the device from
// (case seed, request seed) by a counter-based hash, so the result of a
// request depends on nothing else: not on what ran before or beside it.
// The host loop (dnsp_server_run) replays an arrival trace against the wall
// clock: arrived requests wait in a FIFO; whenever a stream slot is free the
// first waiting request whose case is not busy is admitted (out of order
// across cases).  A case's device state (a dnsp_step object, its IBM handles
// and its captured step graph) is created on first admission if its HBM
// footprint fits the budget, evicting least-recently-used idle cases if
// needed, and otherwise stays resident.  Completion is detected by polling
// an event per slot; latency = completion time - arrival time.  Each
// request's output is reduced to a fixed-order checksum on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <thread>
#include <vector>

#include "common.h"

using namespace dnsp;

namespace {

// SplitMix64 output function (Steele, Lea, Flood 2014).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// uniform in [-1, 1) from (key, index)
__device__ __forceinline__ double unit(unsigned long long key, unsigned long long i) {
    const unsigned long long h = mix64(key ^ mix64(i));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Request inputs on the halo fields, y from the padded cell centres ypad
// (ny+2): U_hat = 1.5 (1 - (y-1)^2) + 0.1 u, V_hat, W_hat = 0.1 u,
// rho = 1 - 0.45 (1 + tanh((T - 0.5)/0.05)) with T = y/2 + 0.05 u,
// P_now, P_old, div_now = 0.01 u; u from distinct keys per field.
__global__ __launch_bounds__(256) void k_gen_inputs(double* uh, double* vh, double* wh, double* rho, double* pn,
                                                    double* po, double* dn, const double* __restrict__ ypad,
                                                    int nx, int nz, int ny, unsigned long long key) {
    const long n = (long)(nx + 2) * (nz + 2) * (ny + 2);
    const long sy = (long)(nx + 2) * (nz + 2);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double y = ypad[i / sy];
        uh[i] = 1.5 * (1.0 - (y - 1.0) * (y - 1.0)) + 0.1 * unit(key + 1, i);
        vh[i] = 0.1 * unit(key + 2, i);
        wh[i] = 0.1 * unit(key + 3, i);
        const double T = 0.5 * y + 0.05 * unit(key + 4, i);
        rho[i] = 1.0 - 0.45 * (1.0 + tanh((T - 0.5) / 0.05));
        pn[i] = 0.01 * unit(key + 5, i);
        po[i] = 0.01 * unit(key + 6, i);
        dn[i] = 0.01 * unit(key + 7, i);
    }
}

// Fixed-order checksum of the request output: for P (interior) and U (halo),
// sum v and sum v * w(i) with w(i) = 1 + (i mod 97)/97 (order-sensitive).
constexpr int kSumBlocks = 256;

__global__ __launch_bounds__(256) void k_checksum_part(const double* __restrict__ a, long n, double* part) {
    __shared__ double s0[256], s1[256];
    double x0 = 0.0, x1 = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = a[i];
        x0 += v;
        x1 += v * (1.0 + (double)(i % 97) / 97.0);
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

__global__ __launch_bounds__(64) void k_checksum_final(const double* __restrict__ part, double* out) {
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < kSumBlocks; ++i) {
            a += part[2 * i];
            b += part[2 * i + 1];
        }
        out[0] = a;
        out[1] = b;
    }
}

struct CaseDesc {
    int nx, nz, ny;
    std::vector<double> yu, ypad;
    double inv_dx, inv_dz, dtmax;
    int incompressible;
    double walls[6];
    std::vector<int> ijk[3];
    std::vector<double> c[3];
    std::vector<int> bot, top;
    unsigned long long seed;
    size_t footprint = 0;  // upper bound before the first build, the measured bytes after it
    bool measured = false;
    // form of the solve: -1 FP64, 0/1 mixed refinements (the server's, or
    // per case the faster one when the server picks)
    int form = -1;
    double ms_fp64 = -1.0, ms_mixed = -1.0;  // step times measured when the server picks
    // device state while resident
    dnsp_step* step = nullptr;
    dnsp_ibm* ibm[3] = {nullptr, nullptr, nullptr};
    double* d_ypad = nullptr;
    double* f[DNSP_STEP_NFIELDS] = {};
    bool busy = false;
    long last_use = -1;
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct dnsp_server {
    size_t budget = 0;
    size_t resident = 0;
    int max_slots = 1;
    int refinements = 1;
    int precision = 0;
    std::vector<CaseDesc> cases;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;
    double* d_part = nullptr;  // kSumBlocks*2 per slot
    double* d_sums = nullptr;  // 4 per slot
    long tick = 0;
    int admissions = 0, evictions = 0;
    size_t peak_resident = 0;
};

namespace {

void drop_case(dnsp_server* sv, CaseDesc& c, bool count = true) {
    if (!c.step) return;
    dnsp_step_free(c.step);
    for (auto& h : c.ibm) {
        if (h) dnsp_ibm_free(h);
        h = nullptr;
    }
    (void)hipFree(c.d_ypad);
    c.d_ypad = nullptr;
    c.step = nullptr;
    sv->resident -= c.footprint;
    if (count) ++sv->evictions;
}

// Times the case's step in both forms (eagerly on `st`, on whatever its
// fields hold: the work does not depend on the values) and keeps the faster.
int pick_form(CaseDesc& c, hipStream_t st) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        set_error("event creation failed");
        return DNSP_ERR_HIP;
    }
    int rc = DNSP_OK;
    double ms[2] = {0.0, 0.0};
    const int forms[2] = {-1, 1};
    for (int f = 0; f < 2 && !rc; ++f) {
        rc = dnsp_step_enqueue(c.step, forms[f], reinterpret_cast<void*>(st));  // warm-up
        if (hipEventRecord(e0, st) != hipSuccess) rc = DNSP_ERR_HIP;
        for (int r = 0; r < 3 && !rc; ++r) rc = dnsp_step_enqueue(c.step, forms[f], reinterpret_cast<void*>(st));
        if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) {
            set_error("timing the case's step failed");
            rc = DNSP_ERR_HIP;
        }
        float t = 0.f;
        if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
        ms[f] = t / 3.0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    c.ms_fp64 = ms[0];
    c.ms_mixed = ms[1];
    c.form = ms[1] < ms[0] ? 1 : -1;
    return DNSP_OK;
}

// Creates the case's device state; the graph is recorded on `st`.
int build_case(dnsp_server* sv, CaseDesc& c, hipStream_t st) {
    int rc = dnsp_step_create(c.nx, c.nz, c.ny, c.yu.data(), c.inv_dx, c.inv_dz, c.dtmax, c.incompressible,
                              sv->precision, &c.step);
    if (rc) return rc;
    rc = dnsp_step_set_walls(c.step, c.walls);
    for (int k = 0; k < 3 && !rc; ++k) {
        if (c.c[k].empty()) continue;
        rc = dnsp_ibm_create(c.nx, c.nz, c.ny, 1, 0, (int)c.c[k].size() / 3, c.ijk[k].data(), c.c[k].data(),
                             &c.ibm[k]);
        if (!rc) rc = dnsp_step_set_ibm(c.step, k, c.ibm[k]);
    }
    if (!rc && !c.bot.empty()) rc = dnsp_step_set_columns(c.step, c.bot.data(), c.top.data());
    for (int k = 0; k < DNSP_STEP_NFIELDS && !rc; ++k) rc = dnsp_step_field(c.step, k, &c.f[k]);
    if (!rc) {
        if (hipMalloc(&c.d_ypad, sizeof(double) * c.ypad.size()) != hipSuccess ||
            hipMemcpy(c.d_ypad, c.ypad.data(), sizeof(double) * c.ypad.size(), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("device allocation for a case failed");
            rc = DNSP_ERR_NOMEM;
        }
    }
    if (!rc && sv->refinements == DNSP_SERVER_PICK_FORM && c.ms_fp64 < 0) rc = pick_form(c, st);
    if (!rc) rc = dnsp_step_capture(c.step, c.form, reinterpret_cast<void*>(st));
    if (rc) {
        // leave no half-built state behind
        if (c.step) dnsp_step_free(c.step);
        for (auto& h : c.ibm) {
            if (h) dnsp_ibm_free(h);
            h = nullptr;
        }
        (void)hipFree(c.d_ypad);
        c.d_ypad = nullptr;
        c.step = nullptr;
        return rc;
    }
    // the bytes the case really holds replace the admission bound
    size_t bytes = 0;
    (void)dnsp_step_footprint(c.step, &bytes);
    for (dnsp_ibm* h : c.ibm) bytes += ibm_device_bytes(h);
    bytes += sizeof(double) * c.ypad.size();
    c.footprint = bytes;
    c.measured = true;
    sv->resident += c.footprint;
    sv->peak_resident = std::max(sv->peak_resident, sv->resident);
    ++sv->admissions;
    return DNSP_OK;
}

// Makes case `ci` resident if the budget allows (evicting idle cases, least
// recently used first).  Returns 1 when resident, 0 when it must wait.
int admit_case(dnsp_server* sv, int ci, hipStream_t st, int* rc_out) {
    CaseDesc& c = sv->cases[(size_t)ci];
    *rc_out = DNSP_OK;
    if (c.step) return 1;
    // evict only when the idle cases' bytes make room: otherwise dropping
    // them would only force their rebuild later
    size_t idle = 0;
    for (const CaseDesc& o : sv->cases)
        if (o.step && !o.busy) idle += o.footprint;
    if (sv->resident - idle + c.footprint > sv->budget) return 0;
    while (sv->resident + c.footprint > sv->budget) {
        int victim = -1;
        for (size_t k = 0; k < sv->cases.size(); ++k) {
            const CaseDesc& o = sv->cases[k];
            if (o.step && !o.busy && (victim < 0 || o.last_use < sv->cases[(size_t)victim].last_use))
                victim = (int)k;
        }
        if (victim < 0) return 0;
        drop_case(sv, sv->cases[(size_t)victim]);
    }
    *rc_out = build_case(sv, c, st);
    return *rc_out ? 0 : 1;
}

int enqueue_request(dnsp_server* sv, int slot, int ci, unsigned long long rseed) {
    CaseDesc& c = sv->cases[(size_t)ci];
    hipStream_t st = sv->streams[(size_t)slot];
    const long nh = (long)(c.nx + 2) * (c.nz + 2) * (c.ny + 2);
    const int gb = (int)std::min<long>((nh + 255) / 256, 4096);
    const unsigned long long key = (c.seed * 0x100000001b3ULL) ^ (rseed * 0x9e3779b97f4a7c15ULL);
    k_gen_inputs<<<gb, 256, 0, st>>>(c.f[DNSP_STEP_U_HAT], c.f[DNSP_STEP_V_HAT], c.f[DNSP_STEP_W_HAT],
                                     c.f[DNSP_STEP_RHO], c.f[DNSP_STEP_P_NOW], c.f[DNSP_STEP_P_OLD],
                                     c.f[DNSP_STEP_DIV_NOW], c.d_ypad, c.nx, c.nz, c.ny, key);
    DNSP_LAUNCH_CHECK();
    int rc = dnsp_step_launch(c.step, reinterpret_cast<void*>(st));
    if (rc) return rc;
    double* part = sv->d_part + (size_t)slot * 2 * kSumBlocks;
    double* sums = sv->d_sums + (size_t)slot * 4;
    k_checksum_part<<<kSumBlocks, 256, 0, st>>>(c.f[DNSP_STEP_P], (long)c.nx * c.nz * c.ny, part);
    k_checksum_final<<<1, 64, 0, st>>>(part, sums);
    k_checksum_part<<<kSumBlocks, 256, 0, st>>>(c.f[DNSP_STEP_U], nh, part);
    k_checksum_final<<<1, 64, 0, st>>>(part, sums + 2);
    DNSP_LAUNCH_CHECK();
    DNSP_HIP_TRY(hipEventRecord(sv->done[(size_t)slot], st));
    return DNSP_OK;
}

}  // namespace

extern "C" {

int dnsp_server_create(size_t hbm_budget, int max_slots, int precision, int refinements, dnsp_server** out) {
    clear_error();
    DNSP_ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    DNSP_ARG_CHECK(max_slots >= 1 && max_slots <= 16, "max_slots=%d not in [1, 16]", max_slots);
    DNSP_ARG_CHECK(refinements >= DNSP_SERVER_PICK_FORM && refinements <= 1,
                   "refinements=%d: -1 (FP64), 0 or 1 (mixed), or DNSP_SERVER_PICK_FORM", refinements);
    DNSP_ARG_CHECK(refinements == -1 || precision == DNSP_PRECISION_MIXED,
                   "the mixed form (and picking a form) needs precision 1");
    dnsp_server* sv = new (std::nothrow) dnsp_server();
    if (!sv) {
        set_error("host allocation failed");
        return DNSP_ERR_NOMEM;
    }
    sv->budget = hbm_budget;
    sv->max_slots = max_slots;
    sv->refinements = refinements;
    sv->precision = precision;
    int rc = DNSP_OK;
    for (int i = 0; i < max_slots && !rc; ++i) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            set_error("stream or event creation failed");
            rc = DNSP_ERR_HIP;
        }
        if (s) sv->streams.push_back(s);
        if (e) sv->done.push_back(e);
    }
    if (!rc && (hipMalloc(&sv->d_part, sizeof(double) * 2 * kSumBlocks * max_slots) != hipSuccess ||
                hipMalloc(&sv->d_sums, sizeof(double) * 4 * max_slots) != hipSuccess)) {
        set_error("device allocation failed");
        rc = DNSP_ERR_NOMEM;
    }
    if (rc) {
        dnsp_server_free(sv);
        return rc;
    }
    *out = sv;
    return DNSP_OK;
}

int dnsp_server_add_case(dnsp_server* sv, const dnsp_case_desc* d, int* case_id) {
    clear_error();
    DNSP_ARG_CHECK(sv && d && case_id, "NULL argument");
    DNSP_ARG_CHECK(d->nx >= 1 && d->nz >= 1 && d->ny >= 2 && d->yu, "invalid case sizes or faces");
    CaseDesc c;
    c.nx = d->nx;
    c.nz = d->nz;
    c.ny = d->ny;
    c.yu.assign(d->yu, d->yu + d->ny + 1);
    c.ypad.resize((size_t)d->ny + 2);
    for (int k = 0; k < d->ny; ++k) c.ypad[(size_t)k + 1] = 0.5 * (c.yu[(size_t)k] + c.yu[(size_t)k + 1]);
    c.ypad[0] = 2 * c.yu[0] - c.ypad[1];
    c.ypad[(size_t)d->ny + 1] = 2 * c.yu[(size_t)d->ny] - c.ypad[(size_t)d->ny];
    c.inv_dx = d->inv_dx;
    c.inv_dz = d->inv_dz;
    c.dtmax = d->dtmax;
    c.incompressible = d->incompressible;
    std::memcpy(c.walls, d->walls, sizeof(c.walls));
    for (int k = 0; k < 3; ++k) {
        DNSP_ARG_CHECK(d->n_ibm[k] >= 0 && (d->n_ibm[k] == 0 || (d->ibm_ijk[k] && d->ibm_c[k])),
                       "IBM component %d: bad records", k);
        c.ijk[k].assign(d->ibm_ijk[k], d->ibm_ijk[k] + 9 * (size_t)d->n_ibm[k]);
        c.c[k].assign(d->ibm_c[k], d->ibm_c[k] + 3 * (size_t)d->n_ibm[k]);
    }
    if (d->bot && d->top) {
        const size_t n = (size_t)d->nx * d->nz;
        c.bot.assign(d->bot, d->bot + n);
        c.top.assign(d->top, d->top + n);
    }
    c.seed = d->seed;
    c.form = sv->refinements == DNSP_SERVER_PICK_FORM ? -1 : sv->refinements;
    // Admission bound until the first build measures the case: 11 halo fields
    // and P; the solver's arrays (a full field of workspace: serial Thomas
    // where the partitioned kernel does not apply, the blocked solve where nx
    // is a multiple of 16; the mixed form's two FP32 fields,
    // line sums, group lists, tables); IBM records (56 bytes each, at most 7
    // periodic images of 16 bytes per target); slabs, geometry.
    const size_t nx = (size_t)c.nx, nz = (size_t)c.nz, ny = (size_t)c.ny;
    const size_t hc = (nx + 2) * (nz + 2) * (ny + 2), pc = nx * nz * ny;
    size_t fp = sizeof(double) * (11 * hc + pc);
    if (ny < 32 || ny > 1024) fp += sizeof(double) * pc;  // serial Thomas workspace
    else if (nx % 16 == 0) fp += sizeof(double) * pc;       // blocked-solve workspace
    fp += sizeof(double) * (2 * nz * ny + 16 * ny + 2048) + 64 * (nx + nz);
    if (sv->precision == DNSP_PRECISION_MIXED) fp += 2 * sizeof(float) * pc + sizeof(int) * (nx * nz / 4 + 8);
    for (int k = 0; k < 3; ++k) fp += (size_t)d->n_ibm[k] * (56 + 7 * 16);
    fp += 2 * sizeof(int) * c.bot.size();
    c.footprint = fp;
    DNSP_ARG_CHECK(c.footprint <= sv->budget, "case footprint %zu bytes exceeds the HBM budget %zu", c.footprint,
                   sv->budget);
    sv->cases.push_back(std::move(c));
    *case_id = (int)sv->cases.size() - 1;
    return DNSP_OK;
}

int dnsp_server_case_footprint(const dnsp_server* sv, int case_id, size_t* bytes) {
    clear_error();
    DNSP_ARG_CHECK(sv && bytes && case_id >= 0 && case_id < (int)sv->cases.size(), "bad argument");
    *bytes = sv->cases[(size_t)case_id].footprint;
    return DNSP_OK;
}

int dnsp_server_case_form(const dnsp_server* sv, int case_id, int* refinements, double* ms_fp64,
                          double* ms_mixed) {
    clear_error();
    DNSP_ARG_CHECK(sv && case_id >= 0 && case_id < (int)sv->cases.size(), "bad argument");
    const CaseDesc& c = sv->cases[(size_t)case_id];
    if (refinements) *refinements = c.form;
    if (ms_fp64) *ms_fp64 = c.ms_fp64;
    if (ms_mixed) *ms_mixed = c.ms_mixed;
    return DNSP_OK;
}

// Replays n requests (case_id[i], seed[i]) arriving at arrival_s[i] seconds
// (non-decreasing) after the call starts.  Outputs per request: latency_ms
// (completion - arrival), start_ms (admission time since the call started)
// and checksum[4] (sum and weighted sum of P, then of U).  stats (optional,
// 4 doubles): admissions, evictions, peak resident bytes, wall seconds.
int dnsp_server_run(dnsp_server* sv, int n, const int* case_id, const double* arrival_s,
                    const unsigned long long* seed, double* latency_ms, double* start_ms, double* checksum,
                    double* stats) {
    clear_error();
    DNSP_ARG_CHECK(sv && n >= 0 && (n == 0 || (case_id && arrival_s && seed && latency_ms)), "NULL argument");
    for (int i = 0; i < n; ++i) {
        DNSP_ARG_CHECK(case_id[i] >= 0 && case_id[i] < (int)sv->cases.size(), "request %d: case %d", i, case_id[i]);
        DNSP_ARG_CHECK(i == 0 || arrival_s[i] >= arrival_s[i - 1], "arrivals must be non-decreasing");
    }
    const int S = sv->max_slots;
    std::vector<int> slot_req((size_t)S, -1);
    std::deque<int> waiting;
    int next = 0, completed = 0, rc = DNSP_OK;
    const auto t0 = Clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::vector<double> host_sums(4);
    while (completed < n && rc == DNSP_OK) {
        bool progress = false;
        const double t = now_s();
        while (next < n && arrival_s[next] <= t) waiting.push_back(next++);
        // completions
        for (int s = 0; s < S && rc == DNSP_OK; ++s) {
            const int r = slot_req[(size_t)s];
            if (r < 0) continue;
            const hipError_t q = hipEventQuery(sv->done[(size_t)s]);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                set_error("request %d failed on the device: %s", r, hipGetErrorString(q));
                rc = DNSP_ERR_HIP;
                break;
            }
            latency_ms[r] = (now_s() - arrival_s[r]) * 1e3;
            if (checksum) {
                if (hipMemcpy(host_sums.data(), sv->d_sums + (size_t)s * 4, sizeof(double) * 4,
                              hipMemcpyDeviceToHost) != hipSuccess) {
                    set_error("checksum copy failed");
                    rc = DNSP_ERR_HIP;
                    break;
                }
                std::memcpy(checksum + 4 * (size_t)r, host_sums.data(), sizeof(double) * 4);
            }
            sv->cases[(size_t)case_id[r]].busy = false;
            slot_req[(size_t)s] = -1;
            ++completed;
            progress = true;
        }
        if (rc) break;
        // admissions: first waiting request whose case is idle, into a free slot
        for (int s = 0; s < S && rc == DNSP_OK; ++s) {
            if (slot_req[(size_t)s] >= 0) continue;
            for (auto it = waiting.begin(); it != waiting.end(); ++it) {
                const int r = *it;
                CaseDesc& c = sv->cases[(size_t)case_id[r]];
                if (c.busy) continue;
                int brc = DNSP_OK;
                if (!admit_case(sv, case_id[r], sv->streams[(size_t)s], &brc)) {
                    rc = brc;
                    if (!rc) continue;  // no memory yet: try a request of a resident case
                    break;
                }
                c.busy = true;
                c.last_use = ++sv->tick;
                if (start_ms) start_ms[r] = now_s() * 1e3;
                rc = enqueue_request(sv, s, case_id[r], seed[r]);
                slot_req[(size_t)s] = r;
                waiting.erase(it);
                progress = true;
                break;
            }
        }
        if (!progress && rc == DNSP_OK) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (rc) {
        for (int s = 0; s < S; ++s) (void)hipStreamSynchronize(sv->streams[(size_t)s]);
        for (auto& c : sv->cases) c.busy = false;
        return rc;
    }
    if (stats) {
        stats[0] = sv->admissions;
        stats[1] = sv->evictions;
        stats[2] = (double)sv->peak_resident;
        stats[3] = now_s();
    }
    return DNSP_OK;
}

int dnsp_server_free(dnsp_server* sv) {
    clear_error();
    if (!sv) return DNSP_OK;
    for (hipStream_t s : sv->streams) (void)hipStreamSynchronize(s);
    for (auto& c : sv->cases) drop_case(sv, c, false);
    for (hipEvent_t e : sv->done) (void)hipEventDestroy(e);
    for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
    (void)hipFree(sv->d_part);
    (void)hipFree(sv->d_sums);
    delete sv;
    return DNSP_OK;
}

}  // extern "C"

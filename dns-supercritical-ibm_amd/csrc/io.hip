// On-disk formats of the reference, host side (rank-0 I/O on one GPU):
//  * field files (src/mod_io.f90 cfd_read_arr :7-79, cfd_write_arr :81-150):
//    stream-unformatted int32 nx, nz, ny, then nx*nz*ny FP64 values, x
//    fastest; a header that differs from the expected sizes is an error;
//  * coefficient text files: a count line, then that many rows of
//    whitespace-separated numbers (writer calc_bands.py:358-364; reader
//    read_column_r8, src/mod_utils.f90:14-35, which takes column icol of
//    rows n_prev .. n_prev+n-1);
//  * yu.txt: the ny+1 wall-normal faces, one per line (calc_bands.py:211).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

using namespace dnsp;

namespace {

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

size_t field_index(int nx, int nz, int halo, int i, int j, int k) {
    if (!halo) return (size_t)i + (size_t)nx * ((size_t)j + (size_t)nz * k);
    return (size_t)(i + 1) + (size_t)(nx + 2) * ((size_t)(j + 1) + (size_t)(nz + 2) * (size_t)(k + 1));
}

// Numbers of one text line; Fortran exponents (1.0d0) are accepted.
bool parse_numbers(const std::string& line, std::vector<double>& out) {
    out.clear();
    std::string t(line);
    for (char& ch : t)
        if (ch == 'd' || ch == 'D') ch = 'e';
        else if (ch == ',') ch = ' ';
    const char* p = t.c_str();
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
        if (!*p) break;
        char* end = nullptr;
        errno = 0;
        const double v = std::strtod(p, &end);
        if (end == p || errno == ERANGE) return false;
        out.push_back(v);
        p = end;
    }
    return true;
}

bool read_lines(const char* path, std::vector<std::string>& lines) {
    File fh(path, "r");
    if (!fh.f) return false;
    lines.clear();
    std::string cur;
    int ch;
    while ((ch = std::fgetc(fh.f)) != EOF) {
        if (ch == '\n') {
            lines.push_back(cur);
            cur.clear();
        } else {
            cur.push_back((char)ch);
        }
    }
    if (!cur.empty()) lines.push_back(cur);
    return true;
}

bool blank(const std::string& s) { return s.find_first_not_of(" \t\r") == std::string::npos; }

}  // namespace

extern "C" {

int dnsp_io_field_shape(const char* path, int* nx, int* nz, int* ny) {
    clear_error();
    DNSP_ARG_CHECK(path && nx && nz && ny, "NULL argument");
    File fh(path, "rb");
    if (!fh.f) {
        set_error("cannot open %s", path);
        return DNSP_ERR_ARG;
    }
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, fh.f) != 3) {
        set_error("%s: short header", path);
        return DNSP_ERR_ARG;
    }
    *nx = hdr[0];
    *nz = hdr[1];
    *ny = hdr[2];
    return DNSP_OK;
}

int dnsp_io_read_field(const char* path, int nx, int nz, int ny, int halo, double* A) {
    clear_error();
    DNSP_ARG_CHECK(path && A && nx >= 1 && nz >= 1 && ny >= 1, "invalid arguments");
    File fh(path, "rb");
    if (!fh.f) {
        set_error("cannot open %s", path);
        return DNSP_ERR_ARG;
    }
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, fh.f) != 3) {
        set_error("%s: short header", path);
        return DNSP_ERR_ARG;
    }
    if (hdr[0] != nx || hdr[1] != nz || hdr[2] != ny) {
        set_error("Mismatch in grid params: expected %d %d %d, file %s has %d %d %d", nx, nz, ny, path, hdr[0], hdr[1],
                  hdr[2]);
        return DNSP_ERR_ARG;
    }
    std::vector<double> line((size_t)nx);
    for (int k = 0; k < ny; ++k)
        for (int j = 0; j < nz; ++j) {
            if (std::fread(line.data(), sizeof(double), (size_t)nx, fh.f) != (size_t)nx) {
                set_error("%s: truncated at line (j=%d, k=%d)", path, j, k);
                return DNSP_ERR_ARG;
            }
            if (!halo) {
                std::memcpy(A + field_index(nx, nz, 0, 0, j, k), line.data(), sizeof(double) * nx);
            } else {
                for (int i = 0; i < nx; ++i) A[field_index(nx, nz, 1, i, j, k)] = line[(size_t)i];
            }
        }
    return DNSP_OK;
}

int dnsp_io_write_field(const char* path, int nx, int nz, int ny, int halo, const double* A) {
    clear_error();
    DNSP_ARG_CHECK(path && A && nx >= 1 && nz >= 1 && ny >= 1, "invalid arguments");
    File fh(path, "wb");
    if (!fh.f) {
        set_error("cannot create %s", path);
        return DNSP_ERR_ARG;
    }
    const int hdr[3] = {nx, nz, ny};
    bool ok = std::fwrite(hdr, sizeof(int), 3, fh.f) == 3;
    std::vector<double> line((size_t)nx);
    for (int k = 0; k < ny && ok; ++k)
        for (int j = 0; j < nz && ok; ++j) {
            for (int i = 0; i < nx; ++i) line[(size_t)i] = A[field_index(nx, nz, halo, i, j, k)];
            ok = std::fwrite(line.data(), sizeof(double), (size_t)nx, fh.f) == (size_t)nx;
        }
    if (!ok) {
        set_error("%s: write failed", path);
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

int dnsp_io_read_column(const char* path, int icol, int n_prev, int n, double* out) {
    clear_error();
    DNSP_ARG_CHECK(path && out && n >= 0 && n_prev >= 0, "invalid arguments");
    if (icol < 0 || icol > 3) {
        set_error("Unrecognized icol value (%d)", icol);
        return DNSP_ERR_ARG;
    }
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) {
        set_error("cannot open %s", path);
        return DNSP_ERR_ARG;
    }
    std::vector<double> nums;
    size_t at = 0;
    while (at < lines.size() && blank(lines[at])) ++at;
    if (at >= lines.size() || !parse_numbers(lines[at], nums) || nums.size() != 1 || nums[0] < 0) {
        set_error("%s: first line must hold the row count", path);
        return DNSP_ERR_ARG;
    }
    const long nb = (long)nums[0];
    if ((long)n_prev + n - 1 >= nb) {
        set_error("mismatch in (n_prev, n, nb) = (%d, %d, %ld)", n_prev, n, nb);
        return DNSP_ERR_ARG;
    }
    long row = 0;
    for (size_t li = at + 1; li < lines.size() && row < nb; ++li) {
        if (blank(lines[li])) continue;
        if (!parse_numbers(lines[li], nums) || (int)nums.size() <= icol) {
            set_error("%s: row %ld has no column %d", path, row, icol);
            return DNSP_ERR_ARG;
        }
        if (row >= n_prev && row < (long)n_prev + n) out[row - n_prev] = nums[(size_t)icol];
        ++row;
    }
    if (row < nb) {
        set_error("%s: %ld rows announced, %ld found", path, nb, row);
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

int dnsp_io_write_table(const char* path, int nrows, int ncols, const double* rows) {
    clear_error();
    DNSP_ARG_CHECK(path && nrows >= 0 && ncols >= 1 && (nrows == 0 || rows), "invalid arguments");
    File fh(path, "w");
    if (!fh.f) {
        set_error("cannot create %s", path);
        return DNSP_ERR_ARG;
    }
    bool ok = std::fprintf(fh.f, "%d\n", nrows) > 0;
    for (int r = 0; r < nrows && ok; ++r)
        for (int c = 0; c < ncols && ok; ++c)
            ok = std::fprintf(fh.f, c + 1 < ncols ? "%.17g " : "%.17g\n", rows[(size_t)r * ncols + c]) > 0;
    if (!ok) {
        set_error("%s: write failed", path);
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

int dnsp_io_read_faces(const char* path, int* ny, double* yu) {
    clear_error();
    DNSP_ARG_CHECK(path && ny, "NULL argument");
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) {
        set_error("cannot open %s", path);
        return DNSP_ERR_ARG;
    }
    std::vector<double> vals, nums;
    for (const std::string& l : lines) {
        if (blank(l)) continue;
        if (!parse_numbers(l, nums) || nums.size() != 1) {
            set_error("%s: one face per line expected", path);
            return DNSP_ERR_ARG;
        }
        vals.push_back(nums[0]);
    }
    if (vals.size() < 3) {
        set_error("%s: at least 3 faces needed", path);
        return DNSP_ERR_ARG;
    }
    for (size_t i = 1; i < vals.size(); ++i)
        if (!(vals[i] > vals[i - 1])) {
            set_error("%s: faces must increase (line %zu)", path, i);
            return DNSP_ERR_ARG;
        }
    *ny = (int)vals.size() - 1;
    if (yu) std::memcpy(yu, vals.data(), sizeof(double) * vals.size());
    return DNSP_OK;
}

}  // extern "C"

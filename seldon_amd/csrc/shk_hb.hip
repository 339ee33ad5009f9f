// shk_hb.hip -- Harwell-Boeing files for device matrices
// (matrix_sparse/IOMatrixMarket.cxx, which holds Seldon's Harwell-Boeing
// reader and writer).
//
// Writing (WriteHarwellBoeing, IOMatrixMarket.cxx:397-510): the stored arrays
// come back from the device and are put in compressed-column order the way
// ConvertToCSC does it -- the coordinates row by row in stored order (the
// symmetric storage adds the mirrored entry after each off-diagonal one,
// Matrix_Conversions.cxx:113-160), then a stable sort by column (its
// MergeSort, Functions_Arrays.cxx:412-480; for the symmetric storage the
// stable sorts by row, by column within rows, then by column amount to a
// stable sort by (column, row)) -- and printed with Seldon's layout: left-
// aligned fields, widths from the entry and column counts, values at 15
// digits in scientific notation, three per line.
//
// Reading (ReadHarwellBoeing, IOMatrixMarket.cxx:80-360): the header and the
// fixed-width fields are parsed as Seldon parses them (the value format's
// last three numbers give the count per line and the width; Seldon splits it
// on " (ABEFGILOPZ*.)", so a D-format such as its own "(3D23.15)" cannot be
// read), giving a compressed-column matrix.  The row-compressed matrix a
// Seldon program gets from it (ConvertMatrix_to_Coordinates, then
// ConvertMatrix_from_Coordinates into a matrix of the same size) is the
// entries stably sorted by (row, column): that is the device assembly from
// coordinates, which runs on the column-ordered entries.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "shk_internal.hpp"

namespace shk {

// ---- writer ----------------------------------------------------------------
static void put_int(std::string &out, int width, int v)
{
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%-*d", width, v);
    out += buf;
}

template <typename T>
static int hb_write_typed(shk_csr A, const char *path)
{
    std::vector<int> ptr, ind;
    std::vector<T> val;
    SHK_TRY(stored_arrays<T>(A, ptr, ind, val));
    const int m = A->m, n = A->n;
    const bool sym = A->storage == SHK_ROWSYMSPARSE;
    std::vector<int> row, col;
    std::vector<T> v;
    for (int i = 0; i < m; i++)
        for (int k = ptr[i]; k < ptr[(size_t)i + 1]; k++) {
            row.push_back(i);
            col.push_back(ind[k]);
            v.push_back(val[k]);
            if (sym && ind[k] != i) {
                row.push_back(ind[k]);
                col.push_back(i);
                v.push_back(val[k]);
            }
        }
    if (sym) {
        // Seldon sizes the expanded list as twice the stored entries less one
        // per row whose first entry is its diagonal; any other layout makes
        // it write entries it never set
        int64_t expect = 2 * (int64_t)ind.size();
        for (int i = 0; i < m; i++)
            if (ptr[i] < ptr[(size_t)i + 1] && ind[ptr[i]] == i) expect--;
        SHK_CHECK(expect == (int64_t)row.size(), SHK_ERR_ARG,
                  "hb write: a RowSymSparse row holds its diagonal other than first");
    }
    const int nnz = (int)row.size();
    // the format's line counts divide by the number of entries
    SHK_CHECK(nnz > 0, SHK_ERR_ARG, "hb write: a matrix without entries");
    std::vector<int> order(nnz);
    std::iota(order.begin(), order.end(), 0);
    if (sym)
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return col[a] != col[b] ? col[a] < col[b] : row[a] < row[b]; });
    else
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return col[a] < col[b]; });
    std::vector<int> cptr((size_t)n + 1, 0);
    for (int k = 0; k < nnz; k++) cptr[(size_t)col[k] + 1]++;
    for (int j = 0; j < n; j++) cptr[(size_t)j + 1] += cptr[j];

    const int ptr_len = int(std::ceil(std::log10(0.1 + nnz + 1))) + 1;
    const int ptr_nperline = std::min(80 / ptr_len, n + 1);
    const int ptrcrd = n / ptr_nperline + 1;
    const int ind_len = int(std::ceil(std::log10(0.1 + n + 1))) + 1;
    const int ind_nperline = std::min(80 / ind_len, nnz);
    const int indcrd = (nnz - 1) / ind_nperline + 1;
    const int valcrd = (nnz - 1) / 3 + 1;
    const std::string ptrfmt = "(" + std::to_string(ptr_nperline) + "I" + std::to_string(ptr_len) + ")";
    const std::string indfmt = "(" + std::to_string(ind_nperline) + "I" + std::to_string(ind_len) + ")";
    const std::string valfmt = "(3D23.15)";

    std::string out;
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%-72s%-8s\n", "Seldon", "S0000008");
    out += buf;
    for (int x : {ptrcrd + indcrd + valcrd, ptrcrd, indcrd, valcrd, 0}) put_int(out, 14, x);
    out += '\n';
    out += m == n ? "RUA" : "RRA";
    out += "           ";
    for (int x : {m, n, nnz, 0}) put_int(out, 14, x);
    out += '\n';
    std::snprintf(buf, sizeof(buf), "%-16s%-16s%-20s%-20s\n", ptrfmt.c_str(), indfmt.c_str(), valfmt.c_str(),
                  valfmt.c_str());
    out += buf;
    for (int i = 0; i <= n; i += ptr_nperline) {
        for (int j = i; j < std::min(i + ptr_nperline, n + 1); j++) put_int(out, ptr_len, cptr[j] + 1);
        out += '\n';
    }
    for (int i = 0; i < nnz; i += ind_nperline) {
        for (int j = i; j < std::min(i + ind_nperline, nnz); j++) put_int(out, ind_len, row[order[j]] + 1);
        out += '\n';
    }
    for (int i = 0; i < nnz; i += 3) {
        for (int j = i; j < std::min(i + 3, nnz); j++) {
            std::snprintf(buf, sizeof(buf), "%-23.15e", (double)v[order[j]]);
            out += buf;
        }
        out += '\n';
    }
    FILE *f = std::fopen(path, "wb");
    SHK_CHECK(f, SHK_ERR_ARG, std::string("hb write: cannot open ") + path);
    bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    ok = (std::fclose(f) == 0) && ok;
    SHK_CHECK(ok, SHK_ERR_ARG, std::string("hb write: write failed: ") + path);
    return SHK_OK;
}

// ---- reader ----------------------------------------------------------------
// Extraction of an int from a field the way `istringstream >> int` reads it:
// leading blanks, an optional sign, digits; false when no digit follows.
static bool field_int(const std::string &s, int &v)
{
    const char *p = s.c_str();
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') p++;
    const char *q = p;
    if (*q == '+' || *q == '-') q++;
    if (*q < '0' || *q > '9') return false;
    char *end = nullptr;
    const long x = std::strtol(p, &end, 10);
    if (x > 2147483647L || x < -2147483647L - 1) return false;
    v = (int)x;
    return true;
}

// Extraction of a value the way `istringstream >> double` (or float) reads
// it: blanks, then sign, digits, a point, digits and an exponent with its own
// digits; the text stops at any other character (a Fortran 'D' exponent is
// not read).  The characters taken are converted with strtod / strtof.
template <typename T>
static bool field_value(const std::string &s, T &v)
{
    const char *p = s.c_str();
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') p++;
    std::string tok;
    const char *q = p;
    if (*q == '+' || *q == '-') tok += *q++;
    int digits = 0;
    while (*q >= '0' && *q <= '9') tok += *q++, digits++;
    if (*q == '.') {
        tok += *q++;
        while (*q >= '0' && *q <= '9') tok += *q++, digits++;
    }
    if (!digits) return false;
    if (*q == 'e' || *q == 'E') {
        std::string e(1, *q++);
        if (*q == '+' || *q == '-') e += *q++;
        int ed = 0;
        while (*q >= '0' && *q <= '9') e += *q++, ed++;
        if (!ed) return false;
        tok += e;
    }
    errno = 0;
    char *end = nullptr;
    if (sizeof(T) == 8)
        v = (T)std::strtod(tok.c_str(), &end);
    else
        v = (T)std::strtof(tok.c_str(), &end);
    return errno != ERANGE;
}

// Reads count fixed-width fields, per_line to a line, into out (Seldon's
// loops at IOMatrixMarket.cxx:283-320); false on a missing or bad field.
template <typename V, typename Parse>
static bool read_fields(std::ifstream &in, int nlines, int per_line, int width, int count, std::vector<V> &out,
                        Parse parse)
{
    out.assign((size_t)std::max(count, 0), V(0));
    int index = 0;
    std::string line;
    for (int i = 0; i < nlines && index < count; i++) {
        if (!std::getline(in, line)) return false;
        size_t k = 0;
        for (int j = 0; j < per_line && index < count; j++) {
            if (k > line.size()) return false;
            if (!parse(line.substr(k, (size_t)width), out[(size_t)index])) return false;
            index++;
            k += (size_t)width;
        }
    }
    return index == count;
}

struct HbHeader {
    int nrow = 0, ncol = 0, nnz = 0;
    int nl_ptr = 0, nl_ind = 0, nl_val = 0, nl_rhs = 0;
    int lw_ptr = 0, ew_ptr = 0, lw_ind = 0, ew_ind = 0, lw_val = 0, ew_val = 0;
    bool pattern = false;
};

static int hb_header(std::ifstream &in, const char *path, HbHeader &h)
{
    const std::string where = std::string("hb read: ") + path + ": ";
    std::string line;
    SHK_CHECK(std::getline(in, line) && line.size() >= 72, SHK_ERR_ARG, where + "title line shorter than 72 characters");
    int total = 0;
    in >> total >> h.nl_ptr >> h.nl_ind >> h.nl_val >> h.nl_rhs;
    std::string type;
    int nelemental = 0;
    in >> type >> h.nrow >> h.ncol >> h.nnz >> nelemental;
    SHK_CHECK(!in.fail(), SHK_ERR_ARG, where + "unreadable line counts or sizes");
    SHK_CHECK(type.size() == 3, SHK_ERR_ARG, where + "the matrix type must have three characters");
    SHK_CHECK(type[0] == 'R' || type[0] == 'C' || type[0] == 'P', SHK_ERR_ARG,
              where + "the matrix type must start with R, C or P");
    SHK_CHECK(type[0] != 'C', SHK_ERR_ARG, where + "complex-valued matrix, which is not supported");
    SHK_CHECK(type[1] != 'H', SHK_ERR_ARG, where + "Hermitian matrix, which is not supported");
    SHK_CHECK(type[1] != 'Z', SHK_ERR_ARG, where + "skew symmetric matrix, which is not supported");
    SHK_CHECK(type[1] != 'S', SHK_ERR_ARG, where + "symmetric matrix, which is not supported");
    SHK_CHECK(nelemental == 0 && type[2] != 'E', SHK_ERR_ARG, where + "elemental matrix, which is not supported");
    SHK_CHECK(h.nrow >= 0 && h.ncol >= 0 && h.nnz >= 0, SHK_ERR_ARG, where + "negative size");
    h.pattern = type[0] == 'P';
    std::getline(in, line);
    SHK_CHECK(std::getline(in, line) && line.size() >= 16, SHK_ERR_ARG, where + "missing format line");
    SHK_CHECK(std::sscanf(line.substr(0, 16).c_str(), "(%dI%d)", &h.lw_ptr, &h.ew_ptr) == 2, SHK_ERR_ARG,
              where + "unreadable pointer format");
    SHK_CHECK(line.size() >= 16 &&
                  std::sscanf(line.substr(16, 16).c_str(), "(%dI%d)", &h.lw_ind, &h.ew_ind) == 2,
              SHK_ERR_ARG, where + "unreadable index format");
    if (!h.pattern) {
        SHK_CHECK(line.size() >= 32, SHK_ERR_ARG, where + "missing value format");
        const std::string element = line.substr(32, 16), delim = " (ABEFGILOPZ*.)";
        std::vector<int> vect;
        size_t b = element.find_first_not_of(delim);
        while (b != std::string::npos) {
            const size_t e = element.find_first_of(delim, b);
            int x = 0;
            if (!field_int(element.substr(b, e == std::string::npos ? std::string::npos : e - b), x)) x = 0;
            vect.push_back(x);
            b = element.find_first_not_of(delim, e);
        }
        SHK_CHECK(vect.size() >= 3, SHK_ERR_ARG,
                  where + "values written in format \"" + element + "\", which could not be parsed");
        h.lw_val = vect[vect.size() - 3];
        h.ew_val = vect[vect.size() - 2];
    }
    SHK_CHECK(h.lw_ptr > 0 && h.ew_ptr > 0 && h.lw_ind > 0 && h.ew_ind > 0 &&
                  (h.pattern || (h.lw_val > 0 && h.ew_val > 0)),
              SHK_ERR_ARG, where + "a field count or width is not positive");
    if (h.nl_rhs != 0) std::getline(in, line);
    return SHK_OK;
}

template <typename T>
static int hb_read_typed(shk_ctx ctx, const char *path, int layout, shk_csr *out)
{
    std::ifstream in(path);
    SHK_CHECK(in.good(), SHK_ERR_ARG, std::string("hb read: unable to read ") + path);
    HbHeader h;
    SHK_TRY(hb_header(in, path, h));
    const std::string where = std::string("hb read: ") + path + ": ";
    std::vector<int> cptr, rind;
    std::vector<T> val;
    const auto one_based = [](const std::string &s, int &x) {
        if (!field_int(s, x)) return false;
        x--;
        return true;
    };
    SHK_CHECK(read_fields(in, h.nl_ptr, h.lw_ptr, h.ew_ptr, h.ncol + 1, cptr, one_based), SHK_ERR_ARG,
              where + "missing or unreadable column pointers");
    SHK_CHECK(read_fields(in, h.nl_ind, h.lw_ind, h.ew_ind, h.nnz, rind, one_based), SHK_ERR_ARG,
              where + "missing or unreadable row indices");
    if (h.pattern)
        val.assign((size_t)h.nnz, (T)1);
    else
        SHK_CHECK(read_fields(in, h.nl_val, h.lw_val, h.ew_val, h.nnz, val,
                              [](const std::string &s, T &x) { return field_value<T>(s, x); }),
                  SHK_ERR_ARG, where + "missing or unreadable values");
    SHK_CHECK(cptr[0] == 0 && cptr[(size_t)h.ncol] == h.nnz, SHK_ERR_ARG, where + "inconsistent column pointers");
    for (int j = 0; j < h.ncol; j++)
        SHK_CHECK(cptr[j] <= cptr[(size_t)j + 1], SHK_ERR_ARG, where + "decreasing column pointers");
    std::vector<int> ccol((size_t)h.nnz);
    for (int j = 0; j < h.ncol; j++)
        for (int k = cptr[j]; k < cptr[(size_t)j + 1]; k++) ccol[k] = j;
    for (int k = 0; k < h.nnz; k++)
        SHK_CHECK(rind[k] >= 0, SHK_ERR_ARG, where + "a row index below 1");
    const shk_dtype dt = sizeof(T) == 8 ? SHK_F64 : SHK_F32;
    if (h.nnz == 0) {
        std::vector<int> zp((size_t)h.nrow + 1, 0);
        return shk_csr_create_layout(ctx, SHK_ROWSPARSE, dt, h.nrow, h.ncol, 0, zp.data(), nullptr, nullptr, 0,
                                     layout, out);
    }
    // coordinates (column order) to the device, assembled there in Seldon's
    // (row, column) order; rows beyond the declared count widen the matrix as
    // ConvertMatrix_from_Coordinates does
    int maxrow = *std::max_element(rind.begin(), rind.end());
    const int mcap = std::max(h.nrow, maxrow + 1);
    const size_t nz = (size_t)h.nnz;
    int *d_row = nullptr, *d_col = nullptr, *d_ptr = nullptr, *d_ind = nullptr;
    T *d_val = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(&d_row, nz * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_col, nz * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_val, nz * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&d_ptr, ((size_t)mcap + 1) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_ind, nz * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_out, nz * sizeof(T));
    hipStream_t s = ctx->stream;
    if (e == hipSuccess) e = hipMemcpyAsync(d_row, rind.data(), nz * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_col, ccol.data(), nz * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_val, val.data(), nz * sizeof(T), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    int rc = e == hipSuccess ? SHK_OK : fail(SHK_ERR_NOMEM, std::string("hb read: device memory: ") + hipGetErrorString(e));
    int m = 0, n = 0;
    if (rc == SHK_OK)
        rc = shk_coo_to_csr(ctx, dt, (int64_t)nz, d_row, d_col, d_val, 0, h.nrow, h.ncol, mcap, &m, &n, d_ptr, d_ind,
                            d_out);
    if (rc == SHK_OK) rc = shk_csr_create_layout(ctx, SHK_ROWSPARSE, dt, m, n, (int64_t)nz, d_ptr, d_ind, d_out, 1,
                                                 layout, out);
    hipStreamSynchronize(s);
    hipFree(d_row);
    hipFree(d_col);
    hipFree(d_val);
    hipFree(d_ptr);
    hipFree(d_ind);
    hipFree(d_out);
    return rc;
}

}  // namespace shk

using namespace shk;

extern "C" {

int shk_csr_write_hb(shk_ctx ctx, shk_csr A, const char *path)
{
    SHK_CHECK(ctx && A && path && A->ctx == ctx, SHK_ERR_ARG, "hb write: null or foreign argument");
    return A->dtype == SHK_F64 ? hb_write_typed<double>(A, path) : hb_write_typed<float>(A, path);
}

int shk_csr_read_hb(shk_ctx ctx, const char *path, shk_dtype dtype, int layout, shk_csr *out)
{
    SHK_CHECK(ctx && path && out, SHK_ERR_ARG, "hb read: null argument");
    SHK_CHECK(dtype == SHK_F64 || dtype == SHK_F32, SHK_ERR_ARG, "hb read: bad dtype");
    return dtype == SHK_F64 ? hb_read_typed<double>(ctx, path, layout, out)
                            : hb_read_typed<float>(ctx, path, layout, out);
}

}  // extern "C"

// shk_sor.hip -- Seldon's SOR / SSOR sweeps on the device, level-scheduled.
//
// SOR(A, X, B, omega, iter, type_ssor) (matrix_sparse/Relaxation_MatVect.cxx:
// 41-95 RowSparse, :106-167 ArrayRowSparse) updates X row after row:
//   temp = sum_k a_jk X(ind_k)          (the whole row, stored order)
//   temp = B(j) - temp + ajj * X(j)
//   X(j) = (1 - omega) * X(j) + omega * temp / ajj
// with the rows in increasing order (forward sweep, type_ssor 2), decreasing
// (backward, 3), or all forward sweeps then all backward ones (0).  A row
// reads the new X of the rows swept before it and the old X of the others.
// Rows are grouped in levels such that a row comes after every earlier-swept
// row whose X it reads and before every later-swept row that reads its X
// (both directions of each stored entry count); rows of one level then
// neither read nor write each other's X, so a level runs as one launch, one
// thread per row, and every row sees exactly the operands of the sequential
// sweep: the results are bitwise Seldon's.
#include <algorithm>
#include <string>
#include <vector>

#include "shk_kernels.hpp"

namespace shk {

struct SorPlan {
    int m = 0, storage = 0, dtype = 0;
    int *ptr = nullptr, *ind = nullptr;
    void *val = nullptr, *diag = nullptr;
    int *rows[2] = {nullptr, nullptr};  // forward, backward: rows grouped by level
    std::vector<int> off[2];            // level offsets into rows[d] (host)
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sor_rows(const int *__restrict__ rows, int count, const int *__restrict__ ptr,
                                                     const int *__restrict__ ind, const T *__restrict__ val,
                                                     const T *__restrict__ diag, const T *__restrict__ B, T *X,
                                                     T omega)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const int j = rows[t];
    T temp = (T)0;
    for (int k = ptr[j]; k < ptr[j + 1]; k++) temp += val[k] * X[ind[k]];
    const T ajj = diag[j];
    const T xj = X[j];
    temp = B[j] - temp + ajj * xj;
    X[j] = ((T)1 - omega) * xj + omega * temp / ajj;
}

// Levels of one sweep direction (dir 0: rows 0..m-1, dir 1: m-1..0).
static void sor_levels(int m, const std::vector<int> &ptr, const std::vector<int> &ind, int dir,
                       std::vector<int> &rows, std::vector<int> &off)
{
    std::vector<int> lev(m, 0), minlev(m, 0);
    int nlev = 0;
    for (int q = 0; q < m; q++) {
        const int j = dir == 0 ? q : m - 1 - q;
        int L = minlev[j];
        for (int k = ptr[j]; k < ptr[j + 1]; k++) {
            const int c = ind[k];
            const bool before = dir == 0 ? c < j : c > j;  // swept before j: read new
            if (before) L = std::max(L, lev[c] + 1);
        }
        lev[j] = L;
        nlev = std::max(nlev, L + 1);
        for (int k = ptr[j]; k < ptr[j + 1]; k++) {
            const int c = ind[k];
            const bool after = dir == 0 ? c > j : c < j;  // swept after j: read old
            if (after) minlev[c] = std::max(minlev[c], L + 1);
        }
    }
    off.assign((size_t)nlev + 1, 0);
    for (int j = 0; j < m; j++) off[(size_t)lev[j] + 1]++;
    for (int l = 0; l < nlev; l++) off[(size_t)l + 1] += off[l];
    std::vector<int> next(off.begin(), off.end() - 1);
    rows.assign(m, 0);
    for (int j = 0; j < m; j++) rows[next[lev[j]]++] = j;
}

template <typename T>
static int sor_build(shk_csr A, SorPlan **out)
{
    std::vector<int> ptr, ind;
    std::vector<T> val;
    SHK_TRY(stored_arrays<T>(A, ptr, ind, val));
    const int m = A->m;
    // ajj: RowSparse reads A(j, j) (the stored entry, 0 when absent,
    // Matrix_Sparse operator()); ArrayRowSparse adds the row's diagonal
    // entries onto 0 (Relaxation_MatVect.cxx:136-140)
    std::vector<T> diag(m, (T)0);
    for (int j = 0; j < m; j++) {
        bool found = false;
        for (int k = ptr[j]; k < ptr[(size_t)j + 1]; k++)
            if (ind[k] == j) {
                if (A->storage == SHK_ARRAYROWSPARSE) diag[j] += val[k];
                else if (!found) diag[j] = val[k];
                found = true;
            }
    }
    SorPlan *P = new SorPlan;
    P->m = m;
    P->storage = A->storage;
    P->dtype = A->dtype;
    std::vector<int> rows[2];
    for (int d = 0; d < 2; d++) sor_levels(m, ptr, ind, d, rows[d], P->off[d]);
    const size_t nz = ind.size();
    hipError_t e = hipMalloc(&P->ptr, sizeof(int) * ((size_t)m + 1));
    if (e == hipSuccess) e = hipMalloc(&P->ind, sizeof(int) * std::max<size_t>(nz, 1));
    if (e == hipSuccess) e = hipMalloc(&P->val, sizeof(T) * std::max<size_t>(nz, 1));
    if (e == hipSuccess) e = hipMalloc(&P->diag, sizeof(T) * std::max(m, 1));
    for (int d = 0; d < 2 && e == hipSuccess; d++) e = hipMalloc(&P->rows[d], sizeof(int) * std::max(m, 1));
    hipStream_t s = A->ctx->stream;
    if (e == hipSuccess) e = hipMemcpyAsync(P->ptr, ptr.data(), sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nz) e = hipMemcpyAsync(P->ind, ind.data(), sizeof(int) * nz, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nz) e = hipMemcpyAsync(P->val, val.data(), sizeof(T) * nz, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && m) e = hipMemcpyAsync(P->diag, diag.data(), sizeof(T) * m, hipMemcpyHostToDevice, s);
    for (int d = 0; d < 2 && e == hipSuccess && m; d++)
        e = hipMemcpyAsync(P->rows[d], rows[d].data(), sizeof(int) * m, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        sor_free(P);
        return fail(SHK_ERR_NOMEM, std::string("sor: device memory: ") + hipGetErrorString(e));
    }
    *out = P;
    return SHK_OK;
}

void sor_free(SorPlan *P)
{
    if (!P) return;
    hipFree(P->ptr);
    hipFree(P->ind);
    hipFree(P->val);
    hipFree(P->diag);
    hipFree(P->rows[0]);
    hipFree(P->rows[1]);
    delete P;
}

template <typename T>
static int sor_run(shk_ctx ctx, const SorPlan *P, T *X, const T *B, T omega, int iter, int type_ssor)
{
    hipStream_t s = ctx->stream;
    auto sweep = [&](int d) {
        const std::vector<int> &off = P->off[d];
        for (size_t l = 0; l + 1 < off.size(); l++) {
            const int count = off[l + 1] - off[l];
            hipLaunchKernelGGL(k_sor_rows<T>, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               P->rows[d] + off[l], count, P->ptr, P->ind, (const T *)P->val, (const T *)P->diag, B,
                               X, omega);
        }
    };
    if (type_ssor % 2 == 0)
        for (int i = 0; i < iter; i++) sweep(0);
    if (type_ssor % 3 == 0)
        for (int i = 0; i < iter; i++) sweep(1);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

extern "C" {

int shk_sor(shk_ctx ctx, shk_csr A, void *dx, const void *db, double omega, int iter, int type_ssor)
{
    SHK_CHECK(ctx && A && A->ctx == ctx, SHK_ERR_ARG, "sor: null or foreign argument");
    SHK_CHECK(A->storage == SHK_ROWSPARSE || A->storage == SHK_ARRAYROWSPARSE, SHK_ERR_ARG,
              "sor: RowSparse and ArrayRowSparse only");
    SHK_CHECK(A->m == A->n, SHK_ERR_DIM, "sor: the matrix must be square");
    SHK_CHECK(iter >= 0, SHK_ERR_ARG, "sor: negative iteration count");
    if (A->m == 0 || iter == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(db, "b"));
    if (!A->sor) {
        SorPlan *P = nullptr;
        SHK_TRY(A->dtype == SHK_F64 ? sor_build<double>(A, &P) : sor_build<float>(A, &P));
        A->sor = P;
    }
    const SorPlan *P = A->sor;
    if (A->dtype == SHK_F64)
        return sor_run<double>(ctx, P, (double *)dx, (const double *)db, omega, iter, type_ssor);
    return sor_run<float>(ctx, P, (float *)dx, (const float *)db, (float)omega, iter, type_ssor);
}

}  // extern "C"

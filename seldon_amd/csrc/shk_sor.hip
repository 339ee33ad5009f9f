// shk_sor.hip -- Seldon's SOR / SSOR sweeps on the device, level-scheduled.
//
// SOR(A, X, B, omega, iter, type_ssor) runs iter forward sweeps when
// type_ssor % 2 == 0, then iter backward sweeps when type_ssor % 3 == 0.
//
// General storages (matrix_sparse/Relaxation_MatVect.cxx:41-95 RowSparse,
// :106-167 ArrayRowSparse) update X row after row:
//   temp = sum_k a_jk X(ind_k)          (the whole row, stored order)
//   temp = B(j) - temp + ajj * X(j)
//   X(j) = (1 - omega) * X(j) + omega * temp / ajj
// with the rows in increasing order (forward) or decreasing (backward).  A
// row reads the new X of the rows swept before it and the old X of the
// others.  Rows are grouped in levels such that a row comes after every
// earlier-swept row whose X it reads and before every later-swept row that
// reads its X (both directions of each stored entry count); rows of one level
// then neither read nor write each other's X, so a level runs in parallel,
// one thread per row, and every row sees exactly the operands of the
// sequential sweep: the results are bitwise Seldon's.
//
// RowSymSparse (:179-300, upper triangle stored) splits each sweep in two
// passes.  Forward: Z(j) = (1-w)/w ajj X(j) + (B(j) - sum_{p>j} a_jp X(p))
// from the old X (one launch over all rows), then the lower solve
// X(p) = (Z(p) - sum_{j<p} a_jp X(j)) * (w / ajj), the sum over column p of
// the stored triangle in Seldon's scatter order (rows increasing), by levels.
// Backward: Y(j) = sum_{i<j} a_ij X(i) (old X, same order) and
// Z(j) = (1-w)/w ajj X(j) + B(j) - Y(j) in one launch, then the upper solve
// X(j) = (Z(j) - sum_{p>j} a_jp X(p)) w / ajj in decreasing levels.
//
// omega is applied in the type C++ gives (value, omega): a double omega on
// float values makes each relaxation step double, rounded once on the store.
#include <algorithm>
#include <string>
#include <vector>

#include "shk_kernels.hpp"

namespace shk {

// Consecutive levels of at most kRunRows rows are run by one work-group of
// kRunThreads threads, level after level with a barrier between them (the
// barrier orders each level's X stores before the next level's loads within
// the work-group); larger levels get a launch each.
constexpr int kRunThreads = 1024;
constexpr int kRunRows = 1024;

struct LevelSeg {
    int l0, l1;  // levels [l0, l1)
    bool run;    // one work-group steps through them; else one level, one launch
};

struct SorPlan {
    int m = 0, storage = 0, dtype = 0;
    int *ptr = nullptr, *ind = nullptr;  // the stored rows
    void *val = nullptr, *diag = nullptr;
    int *tptr = nullptr, *tind = nullptr;  // RowSymSparse: strict upper entries by column
    void *tval = nullptr, *z = nullptr;    // (rows increasing), and the half-step vector
    int *rows[2] = {nullptr, nullptr};     // forward, backward: rows grouped by level
    int *doff[2] = {nullptr, nullptr};     // level offsets into rows[d] (device)
    std::vector<int> off[2];               // the same (host)
    std::vector<LevelSeg> segs[2];         // launches of direction d
};

// ---- row operations ------------------------------------------------------
// general storages: one row of the sweep
template <typename T, typename W>
struct SorRow {
    const int *ptr, *ind;
    const T *val, *diag, *B;
    T *X;
    W omega;
    __device__ __forceinline__ void operator()(int j) const
    {
        T temp = (T)0;
        for (int k = ptr[j]; k < ptr[j + 1]; k++) temp += val[k] * X[ind[k]];
        const T ajj = diag[j];
        const T xj = X[j];
        temp = B[j] - temp + ajj * xj;
        X[j] = (T)(((W)1 - omega) * (W)xj + omega * (W)temp / (W)ajj);
    }
};

// RowSymSparse forward solve: X(p) = (Z(p) - sum_j a_jp X(j)) * (w / ajj)
template <typename T, typename W>
struct SymForward {
    const int *tptr, *tind;
    const T *tval, *diag, *Z;
    T *X;
    W omega;
    __device__ __forceinline__ void operator()(int p) const
    {
        T x = Z[p];
        for (int k = tptr[p]; k < tptr[p + 1]; k++) x -= tval[k] * X[tind[k]];
        X[p] = (T)((W)x * (omega / (W)diag[p]));
    }
};

// RowSymSparse backward solve: X(j) = (Z(j) - sum_{p != j} a_jp X(p)) w / ajj
template <typename T, typename W>
struct SymBackward {
    const int *ptr, *ind;
    const T *val, *diag, *Z;
    T *X;
    W omega;
    __device__ __forceinline__ void operator()(int j) const
    {
        T temp = (T)0;
        for (int k = ptr[j]; k < ptr[j + 1]; k++)
            if (ind[k] != j) temp += val[k] * X[ind[k]];
        X[j] = (T)((W)(Z[j] - temp) * omega / (W)diag[j]);
    }
};

// one level, one thread per row
template <class Op>
__global__ __launch_bounds__(kBlock) void k_level(const int *__restrict__ rows, int count, Op op)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) op(rows[t]);
}

// levels [l0, l1) in one work-group
template <class Op>
__global__ __launch_bounds__(kRunThreads) void k_level_run(const int *__restrict__ rows, const int *__restrict__ off,
                                                           int l0, int l1, Op op)
{
    for (int l = l0; l < l1; l++) {
        const int end = off[l + 1];
        for (int t = off[l] + (int)threadIdx.x; t < end; t += kRunThreads) op(rows[t]);
        __syncthreads();
    }
}

// ---- RowSymSparse explicit halves ----------------------------------------
// forward: Z(j) = (1-w)/w ajj X(j) + (B(j) - sum_{p != j} a_jp X(p))
template <typename T, typename W>
__global__ __launch_bounds__(kBlock) void k_sor_sym_f1(int m, const int *__restrict__ ptr, const int *__restrict__ ind,
                                                       const T *__restrict__ val, const T *__restrict__ diag,
                                                       const T *__restrict__ B, const T *__restrict__ X,
                                                       T *__restrict__ Z, W omega)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= m) return;
    T temp = (T)0;
    for (int k = ptr[j]; k < ptr[j + 1]; k++)
        if (ind[k] != j) temp += val[k] * X[ind[k]];
    temp = B[j] - temp;
    Z[j] = (T)(((W)1 - omega) / omega * (W)diag[j] * (W)X[j] + (W)temp);
}

// backward: Z(j) = (1-w)/w ajj X(j) + B(j) - sum_{i<j} a_ij X(i)
template <typename T, typename W>
__global__ __launch_bounds__(kBlock) void k_sor_sym_b1(int m, const int *__restrict__ tptr,
                                                       const int *__restrict__ tind, const T *__restrict__ tval,
                                                       const T *__restrict__ diag, const T *__restrict__ B,
                                                       const T *__restrict__ X, T *__restrict__ Z, W omega)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= m) return;
    T y = (T)0;
    for (int k = tptr[j]; k < tptr[j + 1]; k++) y += tval[k] * X[tind[k]];
    Z[j] = (T)(((W)1 - omega) / omega * (W)diag[j] * (W)X[j] + (W)B[j] - (W)y);
}

// ---- plan ----------------------------------------------------------------
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

// Launches of one direction: runs of consecutive small levels, and the
// larger levels one by one.
static void level_segments(const std::vector<int> &off, std::vector<LevelSeg> &segs)
{
    segs.clear();
    const int nlev = (int)off.size() - 1;
    for (int l = 0; l < nlev;) {
        int e = l;
        while (e < nlev && off[(size_t)e + 1] - off[e] <= kRunRows) e++;
        if (e - l >= 2) {
            segs.push_back({l, e, true});
            l = e;
        } else {
            segs.push_back({l, l + 1, false});
            l++;
        }
    }
}

// Seldon's A(j, j) on RowSparse (matrix_sparse/Matrix_Sparse.cxx:1054-1084):
// the scan moves on while the column is below j and stops at the row's last
// entry; the entry it stops on counts if its column is j, else a_jj = 0.
template <typename T>
static T rowsparse_diag(const std::vector<int> &ptr, const std::vector<int> &ind, const std::vector<T> &val, int j)
{
    const int first = ptr[j], last = ptr[(size_t)j + 1] - 1;
    if (last < first) return (T)0;
    int k = first;
    while (k < last && ind[k] < j) k++;
    return ind[k] == j ? val[k] : (T)0;
}

template <typename V>
static hipError_t put(V **dst, const std::vector<V> &src, hipStream_t s)
{
    hipError_t e = hipMalloc((void **)dst, sizeof(V) * std::max<size_t>(src.size(), 1));
    if (e == hipSuccess && !src.empty())
        e = hipMemcpyAsync(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice, s);
    return e;
}

template <typename T>
static int sor_build(shk_csr A, SorPlan **out)
{
    std::vector<int> ptr, ind;
    std::vector<T> val;
    SHK_TRY(stored_arrays<T>(A, ptr, ind, val));
    const int m = A->m;
    const bool sym = A->storage == SHK_ROWSYMSPARSE;
    // ajj: RowSparse reads A(j, j); ArrayRowSparse and RowSymSparse add the
    // row's diagonal entries onto 0 (Relaxation_MatVect.cxx:136-140, :217-225)
    std::vector<T> diag(m, (T)0);
    for (int j = 0; j < m; j++) {
        if (A->storage == SHK_ROWSPARSE) {
            diag[j] = rowsparse_diag(ptr, ind, val, j);
            continue;
        }
        for (int k = ptr[j]; k < ptr[(size_t)j + 1]; k++) {
            SHK_CHECK(!sym || ind[k] >= j, SHK_ERR_ARG, "sor: a RowSymSparse entry below the diagonal");
            if (ind[k] == j) diag[j] += val[k];
        }
    }
    std::vector<int> rows[2], tptr, tind;
    std::vector<T> tval;
    std::vector<int> off[2];
    if (sym) {
        // strict upper entries gathered by column, rows increasing (a stable
        // counting sort of the row-major entries)
        tptr.assign((size_t)m + 1, 0);
        for (int j = 0; j < m; j++)
            for (int k = ptr[j]; k < ptr[(size_t)j + 1]; k++)
                if (ind[k] != j) tptr[(size_t)ind[k] + 1]++;
        for (int p = 0; p < m; p++) tptr[(size_t)p + 1] += tptr[p];
        tind.resize(tptr[m]);
        tval.resize(tptr[m]);
        std::vector<int> next(tptr.begin(), tptr.end() - 1);
        for (int j = 0; j < m; j++)
            for (int k = ptr[j]; k < ptr[(size_t)j + 1]; k++)
                if (ind[k] != j) {
                    const int q = next[ind[k]]++;
                    tind[q] = j;
                    tval[q] = val[k];
                }
        // the forward solve's row p reads the X(j) of its column entries, the
        // backward solve's row j the X(p) of its row entries
        sor_levels(m, tptr, tind, 0, rows[0], off[0]);
        sor_levels(m, ptr, ind, 1, rows[1], off[1]);
    } else {
        for (int d = 0; d < 2; d++) sor_levels(m, ptr, ind, d, rows[d], off[d]);
    }
    SorPlan *P = new SorPlan;
    P->m = m;
    P->storage = A->storage;
    P->dtype = A->dtype;
    P->off[0] = std::move(off[0]);
    P->off[1] = std::move(off[1]);
    for (int d = 0; d < 2; d++) level_segments(P->off[d], P->segs[d]);
    hipStream_t s = A->ctx->stream;
    hipError_t e = put(&P->ptr, ptr, s);
    if (e == hipSuccess) e = put(&P->ind, ind, s);
    if (e == hipSuccess) e = put((T **)&P->val, val, s);
    if (e == hipSuccess) e = put((T **)&P->diag, diag, s);
    for (int d = 0; d < 2 && e == hipSuccess; d++) e = put(&P->rows[d], rows[d], s);
    for (int d = 0; d < 2 && e == hipSuccess; d++) e = put(&P->doff[d], P->off[d], s);
    if (sym) {
        if (e == hipSuccess) e = put(&P->tptr, tptr, s);
        if (e == hipSuccess) e = put(&P->tind, tind, s);
        if (e == hipSuccess) e = put((T **)&P->tval, tval, s);
        if (e == hipSuccess) e = hipMalloc(&P->z, sizeof(T) * std::max(m, 1));
    }
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
    hipFree(P->tptr);
    hipFree(P->tind);
    hipFree(P->tval);
    hipFree(P->z);
    hipFree(P->rows[0]);
    hipFree(P->rows[1]);
    hipFree(P->doff[0]);
    hipFree(P->doff[1]);
    delete P;
}

static inline dim3 grid_for(int count) { return dim3((count + kBlock - 1) / kBlock); }

template <class Op>
static void run_levels(const SorPlan *P, int d, const Op &op, hipStream_t s)
{
    for (const LevelSeg &g : P->segs[d]) {
        if (g.run) {
            hipLaunchKernelGGL(k_level_run<Op>, dim3(1), dim3(kRunThreads), 0, s, (const int *)P->rows[d],
                               (const int *)P->doff[d], g.l0, g.l1, op);
        } else {
            const int first = P->off[d][g.l0], count = P->off[d][g.l1] - first;
            hipLaunchKernelGGL(k_level<Op>, grid_for(count), dim3(kBlock), 0, s, (const int *)P->rows[d] + first,
                               count, op);
        }
    }
}

template <typename T, typename W>
static int sor_run(shk_ctx ctx, const SorPlan *P, T *X, const T *B, W omega, int iter, int type_ssor)
{
    hipStream_t s = ctx->stream;
    const T *val = (const T *)P->val, *diag = (const T *)P->diag, *tval = (const T *)P->tval;
    T *Z = (T *)P->z;
    auto sweep = [&](int d) {
        if (P->storage != SHK_ROWSYMSPARSE) {
            run_levels(P, d, SorRow<T, W>{P->ptr, P->ind, val, diag, B, X, omega}, s);
        } else if (d == 0) {
            hipLaunchKernelGGL((k_sor_sym_f1<T, W>), grid_for(P->m), dim3(kBlock), 0, s, P->m, P->ptr, P->ind, val,
                               diag, B, (const T *)X, Z, omega);
            run_levels(P, 0, SymForward<T, W>{P->tptr, P->tind, tval, diag, Z, X, omega}, s);
        } else {
            hipLaunchKernelGGL((k_sor_sym_b1<T, W>), grid_for(P->m), dim3(kBlock), 0, s, P->m, P->tptr, P->tind,
                               tval, diag, B, (const T *)X, Z, omega);
            run_levels(P, 1, SymBackward<T, W>{P->ptr, P->ind, val, diag, Z, X, omega}, s);
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

int shk_sor(shk_ctx ctx, shk_csr A, void *dx, const void *db, double omega, shk_dtype omega_type, int iter,
            int type_ssor)
{
    SHK_CHECK(ctx && A && A->ctx == ctx, SHK_ERR_ARG, "sor: null or foreign argument");
    SHK_CHECK(A->storage == SHK_ROWSPARSE || A->storage == SHK_ARRAYROWSPARSE || A->storage == SHK_ROWSYMSPARSE,
              SHK_ERR_ARG, "sor: unknown storage");
    SHK_CHECK(omega_type == SHK_F64 || omega_type == SHK_F32, SHK_ERR_ARG, "sor: bad omega type");
    SHK_CHECK(A->m == A->n, SHK_ERR_DIM, "sor: the matrix must be square");
    SHK_CHECK(iter >= 0, SHK_ERR_ARG, "sor: negative iteration count");
    if (A->m == 0 || iter == 0 || (type_ssor % 2 != 0 && type_ssor % 3 != 0)) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(db, "b"));
    if (!A->sor) {
        SorPlan *P = nullptr;
        SHK_TRY(A->dtype == SHK_F64 ? sor_build<double>(A, &P) : sor_build<float>(A, &P));
        A->sor = P;
    }
    const SorPlan *P = A->sor;
    if (A->dtype == SHK_F64)
        return sor_run<double, double>(ctx, P, (double *)dx, (const double *)db, omega, iter, type_ssor);
    if (omega_type == SHK_F64)
        return sor_run<float, double>(ctx, P, (float *)dx, (const float *)db, omega, iter, type_ssor);
    return sor_run<float, float>(ctx, P, (float *)dx, (const float *)db, (float)omega, iter, type_ssor);
}

}  // extern "C"

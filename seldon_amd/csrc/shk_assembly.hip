// shk_assembly.hip -- CSR assembly from coordinates on the device
// (Seldon's ConvertMatrix_from_Coordinates for RowSparse).
//
// Seldon sorts the entries by row with its bottom-up MergeSort, counts the
// rows into Ptr, then sorts each row by column with the same MergeSort
// (matrix_sparse/Matrix_Conversions.cxx:1288-1335).  That merge takes from
// the right run only on a strictly smaller key (vector/Functions_Arrays.cxx:
// 412-480), so it is stable: the result is the entries ordered by
// (row, column), entries with equal (row, column) -- duplicates, which this
// conversion keeps -- in their input order.  A stable LSD radix sort of the
// packed (row, column) key with the entry index as payload gives exactly that
// order; the values then move by one gather.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "shk_internal.hpp"

namespace shk {

// Packed keys, the entry index as payload, and the largest row / column
// (integer atomics: max is order-independent) and a flag for an index below
// zero after the shift.
__global__ __launch_bounds__(kBlock) void k_coo_scan(int64_t nnz, const int *row, const int *col, int index,
                                                     int *stats /* [0] max row, [1] max col, [2] bad */)
{
    int mr = -1, mc = -1, bad = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kBlock) {
        const int r = row[i] - index, c = col[i] - index;
        bad |= (r < 0) | (c < 0);
        mr = max(mr, r);
        mc = max(mc, c);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mr = max(mr, __shfl_xor(mr, off, kWave));
        mc = max(mc, __shfl_xor(mc, off, kWave));
        bad |= __shfl_xor(bad, off, kWave);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicMax(&stats[0], mr);
        atomicMax(&stats[1], mc);
        if (bad) atomicOr(&stats[2], 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_coo_keys(int64_t nnz, const int *row, const int *col, int index,
                                                     int colbits, uint64_t *key, uint32_t *idx)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kBlock) {
        key[i] = ((uint64_t)(uint32_t)(row[i] - index) << colbits) | (uint32_t)(col[i] - index);
        idx[i] = (uint32_t)i;
    }
}

// Sorted entries -> column indices, values (gathered by entry index) and the
// row pointer: the first entry of each run of equal rows writes the pointers
// of its row and of the empty rows before it; the last entry closes the rest.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_coo_fill(int64_t nnz, int m, int colbits, const uint64_t *key,
                                                     const uint32_t *idx, const T *val, int *ptr, int *ind,
                                                     T *val_out)
{
    const uint64_t cmask = (colbits >= 64) ? ~0ull : ((1ull << colbits) - 1);
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kBlock) {
        const uint64_t k = key[i];
        const int r = (int)(k >> colbits);
        ind[i] = (int)(k & cmask);
        val_out[i] = val[idx[i]];
        const int rprev = i == 0 ? -1 : (int)(key[i - 1] >> colbits);
        for (int q = rprev + 1; q <= r; q++) ptr[q] = (int)i;
        if (i == nnz - 1)
            for (int q = r + 1; q <= m; q++) ptr[q] = (int)nnz;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_int(int64_t n, int v, int *x)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) x[i] = v;
}

static int bits_for(int v)
{
    int b = 0;
    while (b < 31 && (1 << b) <= v) b++;
    return b;
}

template <typename T>
static int coo_to_csr_typed(shk_ctx ctx, int64_t nnz, const int *row, const int *col, const T *val, int index,
                            int m_min, int n_min, int m_cap, int *m_out, int *n_out, int *ptr, int *ind,
                            T *val_out)
{
    hipStream_t st = ctx->stream;
    int *stats = (int *)ctx->scalars;  // 3 ints
    const int init[3] = {-1, -1, 0};
    SHK_HIP(hipMemcpyAsync(stats, init, sizeof(init), hipMemcpyHostToDevice, st));
    const int nb = red_blocks(nnz);
    if (nnz > 0) hipLaunchKernelGGL(k_coo_scan, dim3(nb), dim3(kBlock), 0, st, nnz, row, col, index, stats);
    int h[3];
    SHK_HIP(hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, st));
    SHK_HIP(hipStreamSynchronize(st));
    SHK_CHECK(h[2] == 0, SHK_ERR_ARG, "coo_to_csr: an index is below the base index");
    // m = max(row) + 1 and at least the matrix's previous size; with no
    // entries Seldon's GetNormInf gives 0, so m >= 1 (Matrix_Conversions.cxx:
    // 1310-1316)
    const int m = std::max(std::max(h[0], 0) + 1, m_min);
    const int n = std::max(std::max(h[1], 0) + 1, n_min);
    *m_out = m;
    *n_out = n;
    if (!ptr) return SHK_OK;  // dimensions only
    SHK_CHECK(m <= m_cap, SHK_ERR_DIM, "coo_to_csr: the row pointer holds fewer than m + 1 entries");
    if (nnz == 0) {
        hipLaunchKernelGGL(k_fill_int, dim3(red_blocks(m + 1)), dim3(kBlock), 0, st, (int64_t)m + 1, 0, ptr);
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }
    SHK_CHECK(nnz < ((int64_t)1 << 32), SHK_ERR_ARG, "coo_to_csr: more than 2^32 - 1 entries");
    const int colbits = bits_for(n - 1), rowbits = std::max(bits_for(m - 1), colbits == 0 ? 1 : 0);
    // workspace: keys and indices, in and out, then rocPRIM's temporary
    const size_t kb = ((size_t)nnz * 8 + 255) & ~(size_t)255, ib = ((size_t)nnz * 4 + 255) & ~(size_t)255;
    size_t tmp_bytes = 0;
    SHK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                      (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)nnz, 0,
                                      colbits + rowbits, st));
    SHK_TRY(ensure_ws2(ctx, 2 * kb + 2 * ib + tmp_bytes + 256));
    char *w = (char *)ctx->ws2;
    uint64_t *k0 = (uint64_t *)w, *k1 = (uint64_t *)(w + kb);
    uint32_t *i0 = (uint32_t *)(w + 2 * kb), *i1 = (uint32_t *)(w + 2 * kb + ib);
    void *tmp = w + 2 * kb + 2 * ib;
    hipLaunchKernelGGL(k_coo_keys, dim3(nb), dim3(kBlock), 0, st, nnz, row, col, index, colbits, k0, i0);
    SHK_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, (const uint64_t *)k0, k1, (const uint32_t *)i0, i1,
                                      (size_t)nnz, 0, colbits + rowbits, st));
    hipLaunchKernelGGL(k_coo_fill<T>, dim3(nb), dim3(kBlock), 0, st, nnz, m, colbits, (const uint64_t *)k1,
                       (const uint32_t *)i1, val, ptr, ind, val_out);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

extern "C" {

int shk_coo_to_csr(shk_ctx ctx, shk_dtype dt, int64_t nnz, const int *drow, const int *dcol, const void *dval,
                   int index, int m_min, int n_min, int m_cap, int *m, int *n, int *dptr, int *dind,
                   void *dval_out)
{
    SHK_CHECK(ctx && m && n, SHK_ERR_ARG, "coo_to_csr: null argument");
    SHK_CHECK(nnz >= 0, SHK_ERR_ARG, "coo_to_csr: negative entry count");
    SHK_CHECK(dt == SHK_F64 || dt == SHK_F32, SHK_ERR_ARG, "coo_to_csr: bad dtype");
    if (nnz > 0) {
        SHK_TRY(check_device_ptr(drow, "row"));
        SHK_TRY(check_device_ptr(dcol, "col"));
        SHK_TRY(check_device_ptr(dval, "val"));
        if (dptr) {
            SHK_TRY(check_device_ptr(dind, "ind out"));
            SHK_TRY(check_device_ptr(dval_out, "val out"));
        }
    }
    if (dt == SHK_F64)
        return coo_to_csr_typed<double>(ctx, nnz, drow, dcol, (const double *)dval, index, m_min, n_min, m_cap, m,
                                        n, dptr, dind, (double *)dval_out);
    return coo_to_csr_typed<float>(ctx, nnz, drow, dcol, (const float *)dval, index, m_min, n_min, m_cap, m, n,
                                   dptr, dind, (float *)dval_out);
}

}  // extern "C"

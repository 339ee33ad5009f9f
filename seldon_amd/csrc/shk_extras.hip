// shk_extras.hip -- the adjoint SDDMM gradient kernel (no Seldon counterpart).
#include "shk_kernels.hpp"

namespace shk {

// g_pos = -(lam_row * x_col) for every device entry of the row blocks.  The
// block's row pointers and lam values are staged in LDS; entries are taken
// by consecutive threads (coalesced), each finding its row by binary search
// over the staged pointers.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sddmm(int nblk, int xcd_q, const int4 *blk_info, const int *ptr,
                                                  const uint32_t *ind, const T *lam, const T *x, T *g)
{
    __shared__ int sptr[kBlock + 1];
    __shared__ T slam[kBlock];
    const int lb = xcd_logical_block(xcd_q);
    if (lb >= nblk) return;
    const int4 bi = blk_info[lb];
    const int rs = bi.x, re = bi.y, js = bi.z, je = bi.w;
    const int nr = re - rs;  // <= kBlock (the planner's row cap)
    for (int i = threadIdx.x; i <= nr; i += kBlock) sptr[i] = ptr[rs + i];
    for (int i = threadIdx.x; i < nr; i += kBlock) slam[i] = lam[rs + i];
    __syncthreads();
    for (int j = js + (int)threadIdx.x; j < je; j += kBlock) {
        int lo = 0, hi = nr - 1;  // last row with sptr[row] <= j
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sptr[mid] <= j) lo = mid;
            else hi = mid - 1;
        }
        g[j] = -(slam[lo] * x[ind[j] & ~kLowerFlag]);
    }
}

// Panel layout: one work-group per panel walks its chunks in order with a
// per-row count of the entries already seen, so each entry's gradient lands
// at the caller's stored position ptr[row] + k.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sddmm_panel(const int *pchunk, const int4 *chunk,
                                                        const uint32_t *word, const unsigned short *slot_row,
                                                        const int *ptr, const T *lam, const T *x, T *g)
{
    __shared__ int cnt[kPanelRows], segs[kPanelRows];
    const int p = blockIdx.x, r0 = p * kPanelRows, t = threadIdx.x;
    for (int i = t; i < kPanelRows; i += kBlock) cnt[i] = 0;
    for (int cc = pchunk[p]; cc < pchunk[p + 1]; cc++) {
        const int4 ci = chunk[cc];
        __syncthreads();
        for (int e = ci.z + t; e < ci.w; e += kBlock) {
            const uint32_t w = word[e];
            if (w & kPanFirst) segs[(w >> kPanRowShift) & kPanRowMask] = e;
        }
        __syncthreads();
        for (int e = ci.z + t; e < ci.w; e += kBlock) {
            const uint32_t w = word[e];
            const int lr = (w >> kPanRowShift) & kPanRowMask;
            const int row = r0 + slot_row[(size_t)p * kPanelRows + lr];
            const int col = chunk_c0(ci) + (int)(w & kPanColMask);
            g[ptr[row] + cnt[lr] + (e - segs[lr])] = -(lam[row] * x[col]);
        }
        __syncthreads();
        for (int e = ci.z + t; e < ci.w; e += kBlock) {
            const uint32_t w = word[e];
            if (w & kPanLast) {
                const int lr = (w >> kPanRowShift) & kPanRowMask;
                cnt[lr] += e - segs[lr] + 1;
            }
        }
    }
}

// Panel layout, RowSymSparse: the per-slot terms G for the fold below, at
// the panel positions the maps point to.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sddmm_panel_slots(const int *pchunk, const int4 *chunk,
                                                              const uint32_t *word, const unsigned short *slot_row,
                                                              const T *lam, const T *x, T *G)
{
    const int p = blockIdx.x, r0 = p * kPanelRows, t = threadIdx.x;
    for (int cc = pchunk[p]; cc < pchunk[p + 1]; cc++) {
        const int4 ci = chunk[cc];
        for (int e = ci.z + t; e < ci.w; e += kBlock) {
            const uint32_t w = word[e];
            const int row = r0 + slot_row[(size_t)p * kPanelRows + ((w >> kPanRowShift) & kPanRowMask)];
            G[e] = -(lam[row] * x[chunk_c0(ci) + (int)(w & kPanColMask)]);
        }
    }
}

// RowSymSparse: an upper entry (i, c) stands for a_ic and a_ci:
// g = -(lam_i x_c) + -(lam_c x_i); the diagonal has only the first term.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sddmm_fold(int64_t nnz, const T *G, const int64_t *map_up,
                                                       const int64_t *map_lo, T *g)
{
    for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * kBlock) {
        T v = G[map_up[j]];
        if (map_lo[j] >= 0) v += G[map_lo[j]];
        g[j] = v;
    }
}

template <typename T>
static int sddmm_typed(shk_ctx ctx, shk_csr A, const T *lam, const T *x, T *g)
{
    T *G = g;
    if (A->storage == SHK_ROWSYMSPARSE) {
        SHK_TRY(ensure_ws2(ctx, sizeof(T) * (size_t)std::max<int64_t>(A->nslots, 1)));
        G = (T *)ctx->ws2;
    }
    if (A->layout == SHK_LAYOUT_PANELS && A->storage == SHK_ROWSYMSPARSE) {
        if (A->npanel)
            hipLaunchKernelGGL(k_sddmm_panel_slots<T>, dim3(A->npanel), dim3(kBlock), 0, ctx->stream,
                               (const int *)A->pchunk, (const int4 *)A->chunk, (const uint32_t *)A->ind,
                               (const unsigned short *)A->slot_row, lam, x, G);
    } else if (A->layout == SHK_LAYOUT_PANELS) {
        if (A->npanel)
            hipLaunchKernelGGL(k_sddmm_panel<T>, dim3(A->npanel), dim3(kBlock), 0, ctx->stream,
                               (const int *)A->pchunk, (const int4 *)A->chunk, (const uint32_t *)A->ind,
                               (const unsigned short *)A->slot_row, (const int *)A->ptr, lam, x, G);
    } else if (A->nblk > 0)
        hipLaunchKernelGGL(k_sddmm<T>, dim3(A->nblk_launch), dim3(kBlock), 0, ctx->stream, A->nblk, A->xcd_q,
                           (const int4 *)A->blk_info, (const int *)A->ptr, (const uint32_t *)A->ind, lam, x, G);
    if (A->storage == SHK_ROWSYMSPARSE && A->nnz_stored > 0)
        hipLaunchKernelGGL(k_sddmm_fold<T>, dim3(red_blocks(A->nnz_stored)), dim3(kBlock), 0, ctx->stream,
                           A->nnz_stored, (const T *)G, (const int64_t *)A->map_up, (const int64_t *)A->map_lo, g);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

extern "C" {

int shk_sddmm(shk_ctx ctx, shk_csr A, const void *dlam, const void *dx, void *dg)
{
    SHK_CHECK(ctx && A, SHK_ERR_ARG, "shk_sddmm: null argument");
    SHK_CHECK(A->ctx == ctx, SHK_ERR_ARG, "shk_sddmm: matrix of another context");
    if (A->nnz_stored == 0 || A->m == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dlam, "lam"));
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dg, "g"));
    if (A->dtype == SHK_F64)
        return sddmm_typed<double>(ctx, A, (const double *)dlam, (const double *)dx, (double *)dg);
    return sddmm_typed<float>(ctx, A, (const float *)dlam, (const float *)dx, (float *)dg);
}

}  // extern "C"

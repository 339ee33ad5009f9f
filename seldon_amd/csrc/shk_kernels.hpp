// shk_kernels.hpp -- device kernels of the sparse Krylov path (gfx950).
//
// SpMV design (memory-bound; no MFMA: SpMV is not a dense contraction):
//   * rows are cut into row blocks holding at most `budget` stored entries
//     (and at most 256 rows); block boundaries are computed once at upload;
//   * phase 1: the 256 threads of a work-group stream the block's contiguous
//     ind/val range with coalesced loads, gather x[col] and write the products
//     a_j * x_col(j) to LDS (one per entry, balanced over lanes whatever the
//     row lengths are);
//   * phase 2: one thread per row sums its products from LDS sequentially in
//     stored order -- exactly Seldon's `temp += data[j] * X(ind[j])` loop
//     (computation/basic_functions/Functions_MatVect.cxx:424-430) -- so the
//     row results are bitwise equal to the CPU's (no FMA contraction);
//   * rows longer than the budget get a block of their own, processed in
//     LDS-sized chunks with the same sequential order;
//   * the launch is XCD-aware: consecutive row blocks run on one XCD so the
//     x entries they gather are shared through that XCD's L2;
//   * fused epilogues: Jacobi scaling, up to two dot products against the
//     result, and a solver finalize run by the last work-group (deterministic
//     fixed-order grid reduction, no float atomics).
#pragma once

#include "shk_internal.hpp"

namespace shk {

enum SpmvMode { kCsr = 0, kSym = 1, kArray = 2 };
enum DotKind { kDotNone = 0, kDotVec = 1, kDotSelf = 2 };

template <typename T>
struct SpmvArgs {
    int m;
    int nblk, nblk_launch, xcd_q;
    const int *blk_rows;
    const int4 *blk_info;  // per block {first row, end row, first entry, end entry}
    const int *ptr;
    const int *mid;
    const uint32_t *ind;
    const T *val;
    const T *x;
    const T *y_in;     // beta * y_in (ignored when mlt)
    T *y_out;
    T alpha, beta;
    int mlt;           // Seldon Mlt(): Y.Fill(0) then MltAdd(1,..,0,..)
    const T *diag;     // Jacobi epilogue: out = y / diag (null = none)
    int dot0, dot1;    // DotKind for the two partial sums
    const T *u0, *u1;  // dot vectors
    T *partials;
    unsigned *ticket;
    const int *skip;   // device flag: nonzero = return at once (solver done)
};

template <typename T, int BUDGET>
struct alignas(16) SpmvSmem {
    T prod[BUDGET];
};

// Row result from the staged products [p0, p1) with the temp/direct split pm.
template <typename T, int MODE>
__device__ __forceinline__ T spmv_row_result(const SpmvArgs<T> &a, int r, const T *prod, int p0,
                                             int pm, int p1)
{
    T y;
    if (a.mlt) {
        y = (T)0;
    } else {
        T yin = a.y_in[r];
        if (MODE == kArray) y = (a.beta == (T)0) ? (T)0 : yin * a.beta;
        else y = yin * a.beta;
    }
    if (MODE != kArray) {
        T temp = (T)0;
        for (int q = p0; q < pm; q++) temp += prod[q];
        y += a.alpha * temp;
    }
    for (int q = pm; q < p1; q++) y += prod[q];
    return y;
}

template <typename T, int MODE>
__device__ __forceinline__ void spmv_epilogue(const SpmvArgs<T> &a, int r, T y, T (&c)[2])
{
    T out = a.diag ? y / a.diag[r] : y;
    a.y_out[r] = out;
    if (a.dot0 == kDotVec) c[0] += a.u0[r] * out;
    else if (a.dot0 == kDotSelf) c[0] += out * out;
    if (a.dot1 == kDotVec) c[1] += a.u1[r] * out;
    else if (a.dot1 == kDotSelf) c[1] += out * out;
}

template <typename T>
__device__ __forceinline__ T spmv_product(const SpmvArgs<T> &a, int j)
{
    uint32_t c = a.ind[j];
    T v = a.val[j];
    T xv = a.x[c & ~kLowerFlag];
    // direct terms (transposed symmetric part, ArrayRowSparse) multiply
    // alpha into the value first: `alpha * data[j] * X(i)`
    // (Functions_MatVect.cxx:563, Functions_MatrixArray.cxx:571)
    return (c & kLowerFlag) ? (a.alpha * v) * xv : v * xv;
}

// Vector widths of the coalesced streaming loads: lane t of load k reads
// elements [k*256*V + t*V, +V) of the block's range -- 512 B (ind) and 1 KiB
// (values) contiguous per wave instruction.
template <typename T>
struct VecOf;
template <>
struct VecOf<double> {
    static constexpr int V = 2;
    typedef unsigned int IV __attribute__((ext_vector_type(2)));
    typedef double FV __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<float> {
    static constexpr int V = 4;
    typedef unsigned int IV __attribute__((ext_vector_type(4)));
    typedef float FV __attribute__((ext_vector_type(4)));
};

// Phase 1, branch-free so that every load of the block is in flight at once:
// the streaming loads of ind/val (non-temporal, so they do not evict x from
// L2) read the whole budget window [js0, js0 + BUDGET) -- the arrays are padded
// by BUDGET entries -- then every x gather is issued with its index clamped to
// 0 outside [js, je), then the products go to LDS.  prod[e - js0] holds the
// product of entry e (0 outside the block).
template <typename T, int BUDGET>
__device__ __forceinline__ void spmv_stage_products(const SpmvArgs<T> &a, int js0, int js, int je, T *prod)
{
    typedef VecOf<T> VT;
    constexpr int V = VT::V;
    constexpr int K = BUDGET / (kBlock * V);
    typename VT::IV iv[K];
    typename VT::FV fv[K];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = js0 + (k * kBlock + t) * V;
        iv[k] = __builtin_nontemporal_load((const typename VT::IV *)(a.ind + e));
        fv[k] = __builtin_nontemporal_load((const typename VT::FV *)(a.val + e));
    }
    T xv[K][V];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = js0 + (k * kBlock + t) * V;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const bool in = (e + v < je) & (e + v >= js);
            const unsigned col = in ? (iv[k][v] & ~kLowerFlag) : 0u;
            xv[k][v] = a.x[col];
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = js0 + (k * kBlock + t) * V;
        typename VT::FV p;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const T val = fv[k][v];
            // direct terms (transposed symmetric part, ArrayRowSparse)
            // multiply alpha into the value first: `alpha * data[j] * X(i)`
            // (Functions_MatVect.cxx:563, Functions_MatrixArray.cxx:571)
            const T pr = (iv[k][v] & kLowerFlag) ? (a.alpha * val) * xv[k][v] : val * xv[k][v];
            const bool in = (e + v < je) & (e + v >= js);
            p[v] = in ? pr : (T)0;
        }
        *(typename VT::FV *)(prod + (e - js0)) = p;
    }
}

// Sequential sum of prod[q0, q1) onto acc, four products per step so the LDS
// reads overlap; the additions stay strictly left to right.
template <typename T>
__device__ __forceinline__ T seq_sum(T acc, const T *prod, int q0, int q1)
{
    int q = q0;
    for (; q + 4 <= q1; q += 4) {
        const T b0 = prod[q], b1 = prod[q + 1], b2 = prod[q + 2], b3 = prod[q + 3];
        acc += b0;
        acc += b1;
        acc += b2;
        acc += b3;
    }
    for (; q < q1; q++) acc += prod[q];
    return acc;
}

struct FinNone {
    template <typename T>
    __device__ void operator()(const T (&)[2]) const {}
};

// Store the two sums to device memory (the plain fused C-ABI calls).
template <typename T>
struct FinStore {
    T *dst;
    __device__ void operator()(const T (&s)[2]) const
    {
        dst[0] = s[0];
        dst[1] = s[1];
    }
};

template <typename T, int MODE, int BUDGET, typename Fin>
__global__ __launch_bounds__(kBlock) void k_spmv(SpmvArgs<T> a, Fin fin)
{
    if (a.skip && *a.skip) return;
    __shared__ SpmvSmem<T, BUDGET> sm;
    const int lb = xcd_logical_block(a.xcd_q);
    T c[2] = {(T)0, (T)0};
    if (lb < a.nblk) {
        const int4 bi = a.blk_info[lb];
        const int rs = bi.x, re = bi.y, js = bi.z, je = bi.w;
        constexpr int V = VecOf<T>::V;
        const int js0 = js & ~(V - 1);  // the planner keeps je - js0 <= BUDGET
        if (je - js0 <= BUDGET) {
            // the planner puts at most kBlock rows in a block: thread t owns
            // row rs + t; its metadata and y_in are fetched before phase 1
            const int r = rs + (int)threadIdx.x;
            const bool own = r < re;
            const int rr = own ? r : rs;
            int p0 = a.ptr[rr], p1 = a.ptr[rr + 1];
            int pm = MODE == kSym ? a.mid[rr] : (MODE == kCsr ? p1 : p0);
            T yin = (T)0;
            if (!a.mlt) yin = a.y_in[rr];
            spmv_stage_products<T, BUDGET>(a, js0, js, je, sm.prod);
            __syncthreads();
            if (own) {
                p0 -= js0; p1 -= js0; pm -= js0;
                T y;
                if (a.mlt) y = (T)0;
                else if (MODE == kArray) y = (a.beta == (T)0) ? (T)0 : yin * a.beta;
                else y = yin * a.beta;
                if (MODE != kArray) {
                    const T temp = seq_sum((T)0, sm.prod, p0, pm);
                    y += a.alpha * temp;
                }
                y = seq_sum(y, sm.prod, pm, p1);
                spmv_epilogue<T, MODE>(a, r, y, c);
            }
        } else {
            // one long row, staged chunk by chunk; thread 0 keeps the
            // sequential accumulators across chunks
            const int r = rs;
            const int jm = MODE == kSym ? a.mid[r] : (MODE == kCsr ? je : js);
            T y = (T)0, temp = (T)0;
            if (threadIdx.x == 0) {
                if (!a.mlt) {
                    T yin = a.y_in[r];
                    if (MODE == kArray) y = (a.beta == (T)0) ? (T)0 : yin * a.beta;
                    else y = yin * a.beta;
                }
            }
            for (int cs = js; cs < je; cs += BUDGET) {
                const int ce = min(je, cs + BUDGET);
                for (int j = cs + (int)threadIdx.x; j < ce; j += kBlock) sm.prod[j - cs] = spmv_product(a, j);
                __syncthreads();
                if (threadIdx.x == 0) {
                    for (int j = cs; j < ce; j++) {
                        if (j < jm) {
                            temp += sm.prod[j - cs];
                        } else {
                            if (MODE != kArray && j == jm) y += a.alpha * temp;
                            y += sm.prod[j - cs];
                        }
                    }
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                if (MODE != kArray && jm >= je) y += a.alpha * temp;
                spmv_epilogue<T, MODE>(a, r, y, c);
            }
        }
    }
    if (a.dot0 != kDotNone || a.dot1 != kDotNone) {
        T out[2];
        if (grid_sum<T, 2>(c, a.partials, a.ticket, a.nblk_launch, blockIdx.x, out)) {
            if (threadIdx.x == 0) fin(out);
        }
    }
}

template <typename T>
constexpr int budget_of() { return sizeof(T) == 8 ? 2048 : 4096; }

// ------------------------------------------------------------ launch ----
template <typename T, typename Fin>
inline int launch_spmv(shk_csr A, SpmvArgs<T> &a, Fin fin)
{
    shk_ctx ctx = A->ctx;
    a.m = A->m;
    a.nblk = A->nblk;
    a.nblk_launch = A->nblk_launch;
    a.xcd_q = A->xcd_q;
    a.blk_rows = A->blk_rows;
    a.blk_info = (const int4 *)A->blk_info;
    a.ptr = A->ptr;
    a.mid = A->mid;
    a.ind = A->ind;
    a.val = (const T *)A->val;
    if (a.dot0 != kDotNone || a.dot1 != kDotNone) {
        SHK_TRY(ensure_red(ctx, sizeof(T) * 2 * (size_t)A->nblk_launch));
        a.partials = (T *)ctx->red.partials;
        a.ticket = ctx->red.ticket;
    }
    if (A->m == 0) return SHK_OK;
    constexpr int B = budget_of<T>();
    dim3 grid(A->nblk_launch), block(kBlock);
    switch (A->storage) {
    case SHK_ROWSPARSE:
        hipLaunchKernelGGL((k_spmv<T, kCsr, B, Fin>), grid, block, 0, ctx->stream, a, fin);
        break;
    case SHK_ROWSYMSPARSE:
        hipLaunchKernelGGL((k_spmv<T, kSym, B, Fin>), grid, block, 0, ctx->stream, a, fin);
        break;
    default:
        hipLaunchKernelGGL((k_spmv<T, kArray, B, Fin>), grid, block, 0, ctx->stream, a, fin);
        break;
    }
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}



// ------------------------------------------------------------ BLAS-1 ----
// Grid-stride elementwise kernels; the grid (red_blocks(n)) is a function of
// n only, so every reduction is reproducible run to run.

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, T v, T *x, const int *skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const T *x, T *y, const int *skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        y[i] = x[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scal(int64_t n, T alpha, T *x)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] *= alpha;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(int64_t n, T alpha, const T *x, T *y)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        y[i] += alpha * x[i];
}

template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const S *x, D *y)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        y[i] = (D)x[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi(int64_t n, const T *r, const T *d, T *z)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        z[i] = r[i] / d[i];
}

// Generic fused reduction: optional y += alpha x first (alpha == 0 skips, as
// Add does), then sum of u0*v and (optional) u1*v over v = y.
template <typename T, typename Fin>
__global__ __launch_bounds__(kBlock) void k_axpy_dot(int64_t n, T alpha, const T *x, T *y,
                                                     const T *u0, const T *u1, T *partials,
                                                     unsigned *ticket, const int *skip, Fin fin)
{
    if (skip && *skip) return;
    T c[2] = {(T)0, (T)0};
    const bool do_axpy = x && alpha != (T)0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        T v = y[i];
        if (do_axpy) {
            v += alpha * x[i];
            y[i] = v;
        }
        c[0] += u0[i] * v;
        if (u1) c[1] += u1[i] * v;
    }
    T out[2];
    if (grid_sum<T, 2>(c, partials, ticket, gridDim.x, blockIdx.x, out)) {
        if (threadIdx.x == 0) fin(out);
    }
}

}  // namespace shk

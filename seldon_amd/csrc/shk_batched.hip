// shk_batched.hip -- many independent SPD systems in one persistent kernel.
//
// The systems are stacked block-diagonally in one CSR (system s owns rows
// [row_off[s], row_off[s+1]) and its columns point into the same range).  A
// persistent grid of work-groups pulls systems from a device queue, largest
// first; one work-group runs a whole Cg for its system with work-group-level
// reductions only (no grid-wide synchronisation), then takes the next one.
//
// Per system the recurrences are Seldon's Cg with the identity
// preconditioner (computation/solver/iterative/Cg.cxx:45-116): Init(b),
// x = 0, r = b, Finished(r) with tolerance relative to |b_s|, rho = (r, r),
// breakdown codes 1 and 2, ErrorCode -2 at max_iter.  Row sums of the SpMV are
// sequential in stored order, as in Functions_MatVect.cxx:424-430.
//
// The mixed variant runs, per system, FP32 Cg inner solves on the scaled
// correction equation A d = r / |r| and FP64 refinement x += |r| d with the
// true residual r = b - A x in FP64.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "shk_kernels.hpp"

namespace shk {

constexpr int kMaxGroup = 1024;

// Work-group all-reduce of one value, fixed order, result in every thread.
// `buf` alternates between two 16-slot halves so consecutive calls need a
// single barrier each.
template <typename T>
__device__ __forceinline__ T group_allreduce(T v, T *buf, int &phase)
{
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    T s = wave_sum_lane0(v);
    T *b = buf + 16 * phase;
    if (lane == 0) b[wid] = s;
    __syncthreads();
    T tot = b[0];
    for (int w = 1; w < nw; w++) tot += b[w];
    phase ^= 1;
    return tot;
}

template <typename T>
__device__ __forceinline__ T sqrt_b(T v) { return sqrt(v); }
template <>
__device__ __forceinline__ float sqrt_b<float>(float v) { return sqrtf(v); }

// Row access to a matrix in the row-block layout (the caller's CSR order).
struct RowsView {
    const int *ptr;
    const uint32_t *ind;
    const void *val;
};

inline RowsView rows_view(shk_csr A)
{
    RowsView v;
    v.ptr = A->ptr;
    v.ind = A->ind;
    v.val = A->val;
    return v;
}

// sum_j a_ij p_j of row i in stored order
template <typename T, typename P>
__device__ __forceinline__ T row_sum(const RowsView &M, const P *p, int i)
{
    const T *val = (const T *)M.val;
    T temp = (T)0;
    for (int j = M.ptr[i]; j < M.ptr[i + 1]; j++) temp += val[j] * (T)p[M.ind[j]];
    return temp;
}

// The work-group's rows are visited as in a plain grid-stride loop (thread
// tid takes rows r0 + tid, r0 + tid + G, ...), U rows of a thread at a time:
// their loads are all issued before any is used, so a thread keeps U rows'
// memory accesses in flight instead of one (a large system's phases are
// latency-bound otherwise).  Every row and every per-thread partial sum is
// still formed in the same order.
//
// Row sums of U rows: step k adds entry k of each row that has one, so each
// row's additions stay in stored order (Functions_MatVect.cxx:424-430).
template <typename T, typename P, int U>
__device__ __forceinline__ void row_sums(const RowsView &M, const P *p, int i0, int stride, int r1, T (&temp)[U])
{
    const T *val = (const T *)M.val;
    int jb[U], je[U];
    int L = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int i = i0 + u * stride;
        const bool in = i < r1;
        jb[u] = in ? M.ptr[i] : 0;
        je[u] = in ? M.ptr[i + 1] : 0;
        temp[u] = (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) L = max(L, je[u] - jb[u]);
    for (int k = 0; k < L; k++) {
        T av[U], pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = jb[u] + k < je[u] ? jb[u] + k : 0;
            av[u] = val[j];
            pv[u] = (T)p[M.ind[j]];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (jb[u] + k < je[u]) temp[u] += av[u] * pv[u];
    }
}

// One Cg (identity) on rows [r0, r1): Seldon's recurrences.  x must hold the
// initial guess when !zero_x.  Returns the iteration count; *code gets
// ErrorCode().  facteur is 1/|b| (Init), already checked nonzero.
template <typename T, int U>
__device__ int group_cg(int r0, int r1, const RowsView &M, const T *b, T *x,
                        T *r, T *p, T *q, double facteur, double tol, int max_iter, T *buf, int &phase,
                        int *code)
{
    const int tid = threadIdx.x, G = blockDim.x;
    // r = b; x = 0 (Cg.cxx:63-67)
    T part = (T)0;
    for (int i0 = r0 + tid; i0 < r1; i0 += G * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * G;
            v[u] = i < r1 ? b[i] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * G;
            if (i < r1) {
                r[i] = v[u];
                x[i] = (T)0;
                part += v[u] * v[u];
            }
        }
    }
    T rr = group_allreduce(part, buf, phase);
    T rho = (T)1, rho_1 = (T)1;
    int nb = 0;
    *code = 0;
    for (;;) {
        const double reste = facteur * (double)sqrt_b(rr);
        if (reste < tol || nb >= max_iter) break;           // Finished(r)
        rho = rr;                                            // (r, z), z = r
        if (rho == (T)0) { *code = 1; break; }
        const bool first = nb == 0;
        const T beta = first ? (T)0 : rho / rho_1;
        for (int i0 = r0 + tid; i0 < r1; i0 += G * U) {
            T pv[U], rv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * G;
                const bool in = i < r1;
                rv[u] = in ? r[i] : (T)0;
                pv[u] = in && !first ? p[i] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * G;
                if (i < r1) {
                    if (first) {
                        p[i] = rv[u];
                    } else {
                        T v = pv[u] * beta;
                        v += (T)1 * rv[u];
                        p[i] = v;
                    }
                }
            }
        }
        __syncthreads();
        part = (T)0;
        for (int i0 = r0 + tid; i0 < r1; i0 += G * U) {
            T temp[U], pi[U];
            row_sums<T, T, U>(M, p, i0, G, r1, temp);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * G;
                pi[u] = i < r1 ? p[i] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * G;
                if (i < r1) {
                    T qi = (T)0;  // Mlt: 0 + 1 * temp
                    qi += (T)1 * temp[u];
                    q[i] = qi;
                    part += pi[u] * qi;
                }
            }
        }
        const T delta = group_allreduce(part, buf, phase);
        if (delta == (T)0) { *code = 2; break; }
        const T alpha = rho / delta, malpha = -alpha;
        part = (T)0;
        for (int i0 = r0 + tid; i0 < r1; i0 += G * U) {
            T xv[U], pv[U], rv[U], qv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * G;
                const bool in = i < r1;
                xv[u] = in ? x[i] : (T)0;
                pv[u] = in ? p[i] : (T)0;
                rv[u] = in ? r[i] : (T)0;
                qv[u] = in ? q[i] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * G;
                if (i < r1) {
                    if (alpha != (T)0) x[i] = xv[u] + alpha * pv[u];
                    T rn = rv[u];
                    if (malpha != (T)0) {
                        rn += malpha * qv[u];
                        r[i] = rn;
                    }
                    part += rn * rn;
                }
            }
        }
        rr = group_allreduce(part, buf, phase);
        rho_1 = rho;
        nb++;
        __syncthreads();  // p is rewritten next iteration while others may still read it
    }
    if (nb >= max_iter) *code = -2;
    return nb;
}

struct BatchArgs {
    int count;
    const int *order;    // queue position -> system
    const int *row_off;  // [count+1]
    RowsView A64, A32;
    double *x;
    const double *b;
    void *w64;           // FP64 work: r, p, q (3 * n) or r (n) for the mixed variant
    float *w32;          // FP32 work (mixed): r, p, q, d (4 * n)
    int n;
    double tol, inner_tol;
    int max_iter, max_refine;
    int *its;
    int *codes;
    unsigned *queue;
};

__global__ void k_batch_reset(unsigned *queue) { *queue = 0u; }

template <int U>
__global__ __launch_bounds__(kMaxGroup) void k_cg_batched(BatchArgs a)
{
    __shared__ double buf[32];
    __shared__ int s_q;
    int phase = 0;
    double *r = (double *)a.w64, *p = r + a.n, *q = p + a.n;
    for (;;) {
        if (threadIdx.x == 0) s_q = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int qi = s_q;
        __syncthreads();
        if (qi >= a.count) break;
        const int s = a.order[qi];
        const int r0 = a.row_off[s], r1 = a.row_off[s + 1];
        // Init(b)
        double part = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) part += a.b[i] * a.b[i];
        const double bn = sqrt(group_allreduce(part, buf, phase));
        int code = 0, its = 0;
        if (bn != 0.0)
            its = group_cg<double, U>(r0, r1, a.A64, a.b, a.x, r, p, q, 1.0 / bn,
                                      a.tol, a.max_iter, buf, phase, &code);
        if (threadIdx.x == 0) {
            a.its[s] = its;
            a.codes[s] = code;
        }
    }
}

// Mixed: FP64 refinement around FP32 inner Cg, per system.
template <int U>
__global__ __launch_bounds__(kMaxGroup) void k_cg_batched_ir(BatchArgs a)
{
    __shared__ double buf[32];
    __shared__ float buf32[32];
    __shared__ int s_q;
    int phase = 0, phase32 = 0;
    double *r64 = (double *)a.w64;
    float *r32 = a.w32, *p32 = r32 + a.n, *q32 = p32 + a.n, *d32 = q32 + a.n;
    for (;;) {
        if (threadIdx.x == 0) s_q = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int qi = s_q;
        __syncthreads();
        if (qi >= a.count) break;
        const int s = a.order[qi];
        const int r0 = a.row_off[s], r1 = a.row_off[s + 1];
        double part = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            part += a.b[i] * a.b[i];
            a.x[i] = 0.0;
        }
        const double bn = sqrt(group_allreduce(part, buf, phase));
        int code = 0, total = 0;
        if (bn != 0.0) {
            for (int k = 0;; k++) {
                // true residual r = b - A x in FP64
                part = 0;
                for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += blockDim.x * U) {
                    double temp[U], bv[U];
                    row_sums<double, double, U>(a.A64, a.x, i0, blockDim.x, r1, temp);
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int i = i0 + u * blockDim.x;
                        bv[u] = i < r1 ? a.b[i] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int i = i0 + u * blockDim.x;
                        if (i < r1) {
                            double w = bv[u];
                            w += -1.0 * temp[u];
                            r64[i] = w;
                            part += w * w;
                        }
                    }
                }
                const double rn = sqrt(group_allreduce(part, buf, phase));
                if (rn / bn < a.tol) break;
                if (k >= a.max_refine) { code = -2; break; }
                const double inv = 1.0 / rn;
                float part32 = 0.0f;
                for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
                    const float v = (float)(r64[i] * inv);
                    r32[i] = v;
                    part32 += v * v;
                }
                const float rh = sqrt_b(group_allreduce(part32, buf32, phase32));
                int icode = 0;
                // right-hand side and residual share r32: Cg copies b into r
                // first (r = b) and reads b nowhere else
                total += group_cg<float, U>(r0, r1, a.A32, r32, d32, r32, p32, q32,
                                            1.0 / (double)rh, a.inner_tol, a.max_iter, buf32, phase32, &icode);
                for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) a.x[i] += rn * (double)d32[i];
                __syncthreads();
                if (icode > 0) { code = icode; break; }
            }
        }
        if (threadIdx.x == 0) {
            a.its[s] = total;
            a.codes[s] = code;
        }
    }
}

struct BatchPlanHost {
    std::vector<int> order;
};

static int batched_common(shk_ctx ctx, shk_csr A64, shk_csr A32, int count, const int *row_off_host, void *dx,
                          const void *db, double tol, int max_iter, double inner_tol, int group, int rpp,
                          int use_graph, int *its, int *codes, bool mixed)
{
    SHK_CHECK(ctx && A64 && row_off_host && its && codes, SHK_ERR_ARG, "cg_batched: null argument");
    SHK_CHECK(A64->ctx == ctx, SHK_ERR_ARG, "cg_batched: matrix of another context");
    SHK_CHECK(A64->storage == SHK_ROWSPARSE && A64->dtype == SHK_F64, SHK_ERR_ARG,
              "cg_batched: the stacked matrix must be FP64 RowSparse");
    SHK_CHECK(A64->layout == SHK_LAYOUT_ROWS && (!mixed || !A32 || A32->layout == SHK_LAYOUT_ROWS), SHK_ERR_ARG,
              "cg_batched: create the stacked matrices with SHK_LAYOUT_ROWS");
    SHK_CHECK(count >= 0, SHK_ERR_ARG, "cg_batched: negative count");
    if (mixed)
        SHK_CHECK(A32 && A32->ctx == ctx && A32->dtype == SHK_F32 && A32->storage == SHK_ROWSPARSE &&
                      A32->m == A64->m && A32->nnz_stored == A64->nnz_stored,
                  SHK_ERR_ARG, "cg_batched_ir: A32 must be the FP32 copy of A64");
    if (group == 0) group = 512;
    SHK_CHECK(group >= 64 && group <= kMaxGroup && (group & (group - 1)) == 0, SHK_ERR_ARG,
              "cg_batched: group must be a power of two in [64, 1024]");
    if (rpp == 0) rpp = 4;
    SHK_CHECK(rpp == 1 || rpp == 2 || rpp == 4 || rpp == 8, SHK_ERR_ARG, "cg_batched: rows_per_pass must be 1, 2, 4 or 8");
    SHK_CHECK(max_iter >= 0, SHK_ERR_ARG, "cg_batched: negative max_iter");
    if (count == 0) return SHK_OK;
    SHK_CHECK(row_off_host[0] == 0 && row_off_host[count] == A64->m, SHK_ERR_DIM,
              "cg_batched: row_off must cover the matrix rows");
    std::vector<int> sizes(count);
    for (int s = 0; s < count; s++) {
        SHK_CHECK(row_off_host[s + 1] >= row_off_host[s], SHK_ERR_ARG, "cg_batched: row_off not sorted");
        sizes[s] = row_off_host[s + 1] - row_off_host[s];
    }
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(db, "b"));
    // largest systems first: the long ones start early, small ones fill in
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int i, int j) { return sizes[i] > sizes[j]; });
    const int n = A64->m;
    const size_t off_bytes = sizeof(int) * (size_t)(count + 1), ord_bytes = sizeof(int) * (size_t)count;
    const size_t w64_bytes = sizeof(double) * (size_t)n * (mixed ? 1 : 3);
    const size_t w32_bytes = mixed ? sizeof(float) * (size_t)n * 4 : 0;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t need = al(off_bytes) + al(ord_bytes) + al(w64_bytes) + al(w32_bytes) + 2 * al(ord_bytes) + 256;
    SHK_TRY(ensure_ws2(ctx, need));
    char *w = (char *)ctx->ws2;
    int *d_off = (int *)w;
    int *d_order = (int *)(w + al(off_bytes));
    double *d_w64 = (double *)(w + al(off_bytes) + al(ord_bytes));
    float *d_w32 = (float *)((char *)d_w64 + al(w64_bytes));
    int *d_its = (int *)((char *)d_w32 + al(w32_bytes));
    int *d_codes = (int *)((char *)d_its + al(ord_bytes));
    hipStream_t st = ctx->stream;
    SHK_HIP(hipMemcpyAsync(d_off, row_off_host, off_bytes, hipMemcpyHostToDevice, st));
    SHK_HIP(hipMemcpyAsync(d_order, order.data(), ord_bytes, hipMemcpyHostToDevice, st));
    BatchArgs a{};
    a.count = count;
    a.order = d_order;
    a.row_off = d_off;
    a.A64 = rows_view(A64);
    if (mixed) a.A32 = rows_view(A32);
    a.x = (double *)dx;
    a.b = (const double *)db;
    a.w64 = d_w64;
    a.w32 = d_w32;
    a.n = n;
    a.tol = tol;
    a.inner_tol = inner_tol;
    a.max_iter = max_iter;
    a.max_refine = 50;
    a.its = d_its;
    a.codes = d_codes;
    a.queue = ctx->red.ticket + 32;  // a ticket slot of its own
    const int per_cu = std::max(1, 2048 / group);
    const int grid = std::max(1, std::min(count, ctx->num_cus * per_cu));
    auto enqueue = [&]() -> int {
        hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, st, a.queue);
        const dim3 gd(grid), bd(group);
        if (mixed) {
            if (rpp == 1) hipLaunchKernelGGL(k_cg_batched_ir<1>, gd, bd, 0, st, a);
            else if (rpp == 2) hipLaunchKernelGGL(k_cg_batched_ir<2>, gd, bd, 0, st, a);
            else if (rpp == 4) hipLaunchKernelGGL(k_cg_batched_ir<4>, gd, bd, 0, st, a);
            else hipLaunchKernelGGL(k_cg_batched_ir<8>, gd, bd, 0, st, a);
        } else {
            if (rpp == 1) hipLaunchKernelGGL(k_cg_batched<1>, gd, bd, 0, st, a);
            else if (rpp == 2) hipLaunchKernelGGL(k_cg_batched<2>, gd, bd, 0, st, a);
            else if (rpp == 4) hipLaunchKernelGGL(k_cg_batched<4>, gd, bd, 0, st, a);
            else hipLaunchKernelGGL(k_cg_batched<8>, gd, bd, 0, st, a);
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    };
    if (use_graph) {
        // Recorded on the context's own non-blocking stream (the legacy
        // default stream cannot be captured), launched on the caller's stream
        // so it stays ordered after the copies above.  The queue reset is a
        // kernel node, not a memset node.
        hipGraph_t graph;
        hipGraphExec_t exec;
        hipStream_t cs = ctx->own;
        hipStream_t keep = st;
        st = cs;
        SHK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int e = enqueue();
        hipError_t ce = hipStreamEndCapture(cs, &graph);
        st = keep;
        SHK_TRY(e);
        SHK_HIP(ce);
        SHK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        SHK_HIP(hipGraphLaunch(exec, st));
        SHK_HIP(hipStreamSynchronize(st));
        SHK_HIP(hipGraphExecDestroy(exec));
        SHK_HIP(hipGraphDestroy(graph));
    } else {
        SHK_TRY(enqueue());
    }
    SHK_HIP(hipMemcpyAsync(its, d_its, ord_bytes, hipMemcpyDeviceToHost, st));
    SHK_HIP(hipMemcpyAsync(codes, d_codes, ord_bytes, hipMemcpyDeviceToHost, st));
    SHK_HIP(hipStreamSynchronize(st));
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

extern "C" {

int shk_cg_batched(shk_ctx ctx, shk_csr A, int count, const int *row_off_host, void *dx, const void *db,
                   double tol, int max_iter, int group, int rows_per_pass, int use_graph, int *its, int *codes)
{
    return batched_common(ctx, A, nullptr, count, row_off_host, dx, db, tol, max_iter, 0.0, group, rows_per_pass,
                          use_graph, its, codes, false);
}

int shk_cg_batched_ir(shk_ctx ctx, shk_csr A64, shk_csr A32, int count, const int *row_off_host, void *dx,
                      const void *db, double tol, int max_iter, double inner_tol, int group, int rows_per_pass,
                      int use_graph, int *its, int *codes)
{
    return batched_common(ctx, A64, A32, count, row_off_host, dx, db, tol, max_iter, inner_tol, group,
                          rows_per_pass, use_graph, its, codes, true);
}

}  // extern "C"

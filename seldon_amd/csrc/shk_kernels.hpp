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

#include <algorithm>
#include <type_traits>

#include "shk_internal.hpp"

namespace shk {

// Entries staged per row block (LDS products): 16 KiB per work-group.
#ifndef SHK_BUDGET_F64
#define SHK_BUDGET_F64 2048
#endif
#ifndef SHK_BUDGET_F32
#define SHK_BUDGET_F32 4096
#endif
template <typename T>
constexpr int budget_of() { return sizeof(T) == 8 ? SHK_BUDGET_F64 : SHK_BUDGET_F32; }
#ifndef SHK_SPMV_WG_PER_CU
#define SHK_SPMV_WG_PER_CU 4
#endif
// Work-groups of the persistent SpMV per CU; the kernel is compiled for
// that occupancy (4 waves per SIMD: at most 128 VGPRs).
constexpr int kSpmvWgPerCu = SHK_SPMV_WG_PER_CU;

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
    T *partials;       // [2][grid] block sums of the two dots (null: the context's buffer)
    int grid_used;     // set by launch_spmv: work-groups of the launch (partials per dot)
    const int *skip;   // device flag: nonzero = return at once (solver done)
    // panel layout
    int npanel;
    const int *pchunk;   // [npanel+1] first chunk of each panel
    const unsigned short *slot_row;  // panel-local row of each row slot
};

// Entries seq_sum may read past the end of a row (its look-ahead window);
// their values are never used.
constexpr int kSeqPad = 8;

template <typename T, int BUDGET>
struct alignas(16) SpmvSmem {
    T prod[BUDGET + kSeqPad];
};

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

// Phase 1 is split in two so that the streaming loads of the *next* row block
// are in flight while the current one is processed:
//   spmv_load   -- non-temporal, coalesced loads of the whole budget window
//                  [js0, js0 + BUDGET) of ind/val (arrays padded by BUDGET);
//   spmv_gather -- every x gather issued at once (index clamped to 0 outside
//                  [js, je));
//   spmv_products -- the products written to LDS: prod[e - js0] holds the
//                  product of entry e (0 outside the block).
// The kernel issues block i's gathers, then block i+W's streaming loads, and
// only then consumes the gathers: vector loads retire in order, so this way
// waiting on the gathers never waits on the next block's loads.
template <typename T, int BUDGET, int NT = kBlock>
struct SpmvRegs {
    static constexpr int V = VecOf<T>::V;
    static constexpr int K = BUDGET / (NT * V);
    typename VecOf<T>::IV iv[K];
    typename VecOf<T>::FV fv[K];
};

// live = false loads entry 0 in every lane instead (one cache line): the
// loads are never predicated, so the wait counters stay exact.
template <typename T, int BUDGET, int NT = kBlock>
__device__ __forceinline__ void spmv_load(const SpmvArgs<T> &a, int js0, SpmvRegs<T, BUDGET, NT> &R,
                                          bool live = true)
{
    typedef VecOf<T> VT;
    constexpr int V = VT::V, K = SpmvRegs<T, BUDGET, NT>::K;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = live ? js0 + (k * NT + t) * V : 0;
        R.iv[k] = __builtin_nontemporal_load((const typename VT::IV *)(a.ind + e));
        R.fv[k] = __builtin_nontemporal_load((const typename VT::FV *)(a.val + e));
    }
}

template <typename T, int BUDGET>
struct SpmvGather {
    T xv[SpmvRegs<T, BUDGET>::K][VecOf<T>::V];
};

template <typename T, int BUDGET>
__device__ __forceinline__ void spmv_gather(const SpmvArgs<T> &a, int js0, int js, int je,
                                            const SpmvRegs<T, BUDGET> &R, SpmvGather<T, BUDGET> &G)
{
    constexpr int V = VecOf<T>::V, K = SpmvRegs<T, BUDGET>::K;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = js0 + (k * kBlock + t) * V;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const bool in = (e + v < je) & (e + v >= js);
            const unsigned col = in ? (R.iv[k][v] & ~kLowerFlag) : 0u;
            G.xv[k][v] = a.x[col];
        }
    }
}

template <typename T, int BUDGET>
__device__ __forceinline__ void spmv_products(const SpmvArgs<T> &a, int js0, int js, int je,
                                              const SpmvRegs<T, BUDGET> &R, const SpmvGather<T, BUDGET> &G,
                                              T *prod)
{
    typedef VecOf<T> VT;
    constexpr int V = VT::V, K = SpmvRegs<T, BUDGET>::K;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = js0 + (k * kBlock + t) * V;
        typename VT::FV p;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const T val = R.fv[k][v];
            // direct terms (transposed symmetric part, ArrayRowSparse)
            // multiply alpha into the value first: `alpha * data[j] * X(i)`
            // (Functions_MatVect.cxx:563, Functions_MatrixArray.cxx:571)
            const T pr = (R.iv[k][v] & kLowerFlag) ? (a.alpha * val) * G.xv[k][v] : val * G.xv[k][v];
            const bool in = (e + v < je) & (e + v >= js);
            p[v] = in ? pr : (T)0;
        }
        *(typename VT::FV *)(prod + (e - js0)) = p;
    }
}

// Per-thread row metadata of a block (thread t owns row rs + t: the planner
// puts at most kBlock rows in a block).
template <typename T>
struct RowMeta {
    int p0, p1, pm;
    T yin;
};

template <typename T, int MODE>
__device__ __forceinline__ void row_meta_load(const SpmvArgs<T> &a, int4 bi, RowMeta<T> &M)
{
    const int r = bi.x + (int)threadIdx.x;
    const int rr = r < bi.y ? r : bi.x;
    M.p0 = a.ptr[rr];
    M.p1 = a.ptr[rr + 1];
    M.pm = MODE == kSym ? a.mid[rr] : (MODE == kCsr ? M.p1 : M.p0);
    M.yin = a.mlt ? (T)0 : a.y_in[rr];
}

// Sequential sum of prod[q0, q1) onto acc, strictly left to right.  Two
// register groups of eight alternate: while one group is added, the LDS reads
// of the other are in flight, so a long row costs about one dependent add per
// product instead of one LDS round trip per group.  Reads reach at most
// kSeqPad entries past q1.
template <typename T>
__device__ __forceinline__ T seq_sum(T acc, const T *prod, int q0, int q1)
{
    constexpr int W = kSeqPad / 2;
    T A[W], B[W];
#pragma unroll
    for (int k = 0; k < W; k++) A[k] = prod[q0 + k];
#pragma unroll
    for (int k = 0; k < W; k++) B[k] = prod[q0 + W + k];
    int q = q0;
    for (; q + 2 * W <= q1; q += 2 * W) {
#pragma unroll
        for (int k = 0; k < W; k++) acc += A[k];
#pragma unroll
        for (int k = 0; k < W; k++) A[k] = prod[q + 2 * W + k];
#pragma unroll
        for (int k = 0; k < W; k++) acc += B[k];
#pragma unroll
        for (int k = 0; k < W; k++) B[k] = prod[q + 3 * W + k];
    }
#pragma unroll
    for (int k = 0; k < W; k++)
        if (q + k < q1) acc += A[k];
#pragma unroll
    for (int k = 0; k < W; k++)
        if (q + W + k < q1) acc += B[k];
    return acc;
}

struct FinNone {
    template <typename T>
    __device__ void operator()(const T (&)[2]) const {}
};

// The SpMV's dot partials are left for the next kernel, which folds the
// finish into its own start (no finish launch).
struct FinDefer {
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

// One row block whose products are staged: sequential row sums and epilogue.
template <typename T, int MODE, int BUDGET>
__device__ __forceinline__ void spmv_rows(const SpmvArgs<T> &a, int4 bi, int js0, const RowMeta<T> &M,
                                          const T *prod, T (&c)[2])
{
    const int r = bi.x + (int)threadIdx.x;
    if (r >= bi.y) return;
    const int p0 = M.p0 - js0, p1 = M.p1 - js0, pm = M.pm - js0;
    T y;
    if (a.mlt) y = (T)0;
    else if (MODE == kArray) y = (a.beta == (T)0) ? (T)0 : M.yin * a.beta;
    else y = M.yin * a.beta;
    if (MODE != kArray) {
        const T temp = seq_sum((T)0, prod, p0, pm);
        y += a.alpha * temp;
    }
    y = seq_sum(y, prod, pm, p1);
    spmv_epilogue<T, MODE>(a, r, y, c);
}

// A row longer than the budget: staged chunk by chunk, thread 0 keeps the
// sequential accumulators across chunks.
template <typename T, int MODE, int BUDGET>
__device__ __forceinline__ void spmv_long_row(const SpmvArgs<T> &a, int4 bi, T *prod, T (&c)[2])
{
    const int r = bi.x, js = bi.z, je = bi.w;
    const int jm = MODE == kSym ? a.mid[r] : (MODE == kCsr ? je : js);
    T y = (T)0, temp = (T)0;
    if (threadIdx.x == 0 && !a.mlt) {
        const T yin = a.y_in[r];
        if (MODE == kArray) y = (a.beta == (T)0) ? (T)0 : yin * a.beta;
        else y = yin * a.beta;
    }
    for (int cs = js; cs < je; cs += BUDGET) {
        const int ce = min(je, cs + BUDGET);
        for (int j = cs + (int)threadIdx.x; j < ce; j += kBlock) prod[j - cs] = spmv_product(a, j);
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = cs; j < ce; j++) {
                if (j < jm) {
                    temp += prod[j - cs];
                } else {
                    if (MODE != kArray && j == jm) y += a.alpha * temp;
                    y += prod[j - cs];
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

// Persistent, software-pipelined SpMV.  The grid is 8 * W work-groups; the
// ones with blockIdx % 8 == k run on XCD k and sweep that XCD's contiguous
// range of row blocks [k*q, (k+1)*q) together, W apart, so the x entries
// they gather are shared through one L2.  While a work-group stages and sums
// block i, the streaming loads and row metadata of block i+W are already in
// flight, and the block table entry of block i+2W is being fetched.
template <typename T, int MODE, int BUDGET>
__global__ __launch_bounds__(kBlock, SHK_SPMV_WG_PER_CU) void k_spmv(SpmvArgs<T> a, const int4 *__restrict__ blk_info)
{
    if (a.skip && *a.skip) return;
    __shared__ SpmvSmem<T, BUDGET> sm;
    constexpr int V = VecOf<T>::V;
    const int xcd = blockIdx.x % kNumXcd;
    const int stride = gridDim.x / kNumXcd;
    const int lb_end = min(a.nblk, (xcd + 1) * a.xcd_q);
    int lb = xcd * a.xcd_q + (int)(blockIdx.x / kNumXcd);
    T c[2] = {(T)0, (T)0};

    int4 bi = make_int4(0, 0, 0, 0), bi_next = make_int4(0, 0, 0, 0);
    SpmvRegs<T, BUDGET> R;
    RowMeta<T> M;
    if (lb < lb_end) {
        bi = blk_info[lb];
        if (lb + stride < lb_end) bi_next = blk_info[lb + stride];
        if (bi.w - (bi.z & ~(V - 1)) <= BUDGET) spmv_load<T, BUDGET>(a, bi.z & ~(V - 1), R);
        row_meta_load<T, MODE>(a, bi, M);
    }
    while (lb < lb_end) {
        const int lbn = lb + stride;
        const bool haven = lbn < lb_end;
        int4 bi_nn = make_int4(0, 0, 0, 0);
        if (lbn + stride < lb_end) bi_nn = blk_info[lbn + stride];
        const int js0 = bi.z & ~(V - 1);
        const bool short_blk = bi.w - js0 <= BUDGET;
        SpmvGather<T, BUDGET> G;
        if (short_blk) spmv_gather<T, BUDGET>(a, js0, bi.z, bi.w, R, G);
        SpmvRegs<T, BUDGET> Rn;
        RowMeta<T> Mn;
        if (haven) {
            if (bi_next.w - (bi_next.z & ~(V - 1)) <= BUDGET) spmv_load<T, BUDGET>(a, bi_next.z & ~(V - 1), Rn);
            row_meta_load<T, MODE>(a, bi_next, Mn);
        }
        if (short_blk) {
            spmv_products<T, BUDGET>(a, js0, bi.z, bi.w, R, G, sm.prod);
            __syncthreads();
            spmv_rows<T, MODE, BUDGET>(a, bi, js0, M, sm.prod, c);
            __syncthreads();
        } else {
            spmv_long_row<T, MODE, BUDGET>(a, bi, sm.prod, c);
        }
        bi = bi_next;
        bi_next = bi_nn;
        R = Rn;
        M = Mn;
        lb = lbn;
    }
    if (a.dot0 != kDotNone || a.dot1 != kDotNone) block_partial<T, 2>(c, a.partials, gridDim.x, blockIdx.x);
}

// ------------------------------------------------------- panel layout ----
// Rows are grouped in panels of kPanelRows consecutive rows, one work-group
// of NT threads per panel, each thread owning kPanelRows / NT rows (two, in
// the default launch) and their running sums in registers.  A panel's
// entries are stored as a sequence of chunks, each chunk holding, row after
// row, the entries of the panel whose columns fall in one window [c0, c1) of
// at most panel_win<T>() columns, at most `budget` entries in all.  Seldon's
// rows have sorted columns, so walking the chunks in order visits every row's
// entries in stored order: the row sums stay sequential
// (Functions_MatVect.cxx:424-430) and bitwise equal to the CPU's.  For
// RowSymSparse the chunks of the upper parts come first (the temp sums),
// then, from the chunk flagged kChunkDirect on, those of the transposed
// parts; at that chunk each row's y = beta y + alpha temp is formed and the
// direct terms `(alpha * data[j]) * X(i)` follow (Functions_MatVect.cxx:
// 553-563).
//
// Per chunk the x window is staged in LDS with coalesced loads, and the
// products read x from LDS: the row-block kernel's per-entry gathers each
// cost a 128-byte L2-to-L1 line for 8 useful bytes, which bounded it far
// below the HBM roofline on matrices whose columns spread over a wide band.
// Row slot s of a panel holds the panel's s-th longest row, so the long rows
// share wavefronts and the waves of short rows finish their sums early.
// Each entry's 32-bit word packs its window-local column (13 bits), its row
// slot (11 bits), "first / last entry of the row in this chunk" flags, and
// kLowerFlag (ArrayRowSparse: a direct term).  The entries of the next
// panel_sets() - 1 chunks and the next panel_xlead() x windows are in flight
// while a chunk is summed.
constexpr int kPanelRows = 2048;
constexpr int kPanelThreads = 1024;  // work-group of the panel kernel: two rows per thread
constexpr uint32_t kPanColMask = 0x1FFFu;
constexpr int kPanRowShift = 13;
constexpr uint32_t kPanRowMask = 0x7FFu;
constexpr uint32_t kPanFirst = 1u << 24;
constexpr uint32_t kPanLast = 1u << 25;
// chunk descriptor {c0 | flag, c1, e0, e1}: the flag marks a RowSymSparse
// panel's first chunk of transposed parts
constexpr uint32_t kChunkDirect = 1u << 31;
__device__ __forceinline__ int chunk_c0(int4 ci) { return (int)((uint32_t)ci.x & ~kChunkDirect); }
template <typename T>
constexpr int panel_win() { return sizeof(T) == 8 ? 3072 : 6144; }  // 24 KiB of x
#ifndef SHK_PANEL_BUDGET_F64
#define SHK_PANEL_BUDGET_F64 4096
#endif
#ifndef SHK_PANEL_BUDGET_F32
#define SHK_PANEL_BUDGET_F32 4096
#endif
// register sets of chunk entries in the panel kernel's pipeline: loads run
// panel_sets() - 1 chunks ahead of the sums
#ifndef SHK_PANEL_SETS_F64
#define SHK_PANEL_SETS_F64 2
#endif
#ifndef SHK_PANEL_SETS_F32
#define SHK_PANEL_SETS_F32 3
#endif
// waves per SIMD the panel kernel is compiled for: 8 = two 1024-thread
// groups per CU, which caps it at 64 VGPRs
#ifndef SHK_PANEL_MINW
#define SHK_PANEL_MINW 8
#endif
// products read per step of a row's sum in the panel kernel's phase 2: a
// wave steps as long as its longest segment, so a larger step reads more
// products that its shorter rows discard (measured on the power-law matrix:
// one per step beats four by ~1% FP64 and ~4% FP32, six is slower)
#ifndef SHK_PANEL_P2_UNROLL
#define SHK_PANEL_P2_UNROLL 1
#endif
template <typename T>
constexpr int panel_sets() { return sizeof(T) == 8 ? SHK_PANEL_SETS_F64 : SHK_PANEL_SETS_F32; }
// x windows in flight (register sets of the window loads)
#ifndef SHK_PANEL_XLEAD_F64
#define SHK_PANEL_XLEAD_F64 1
#endif
#ifndef SHK_PANEL_XLEAD_F32
#define SHK_PANEL_XLEAD_F32 1
#endif
template <typename T>
constexpr int panel_xlead() { return sizeof(T) == 8 ? SHK_PANEL_XLEAD_F64 : SHK_PANEL_XLEAD_F32; }
// entries per chunk (LDS products: 32 KiB FP64, 16 KiB FP32)
template <typename T>
constexpr int panel_budget() { return sizeof(T) == 8 ? SHK_PANEL_BUDGET_F64 : SHK_PANEL_BUDGET_F32; }

// The staged x window (32-bit halves in two planes were tried for FP64: the
// scattered 64-bit reads are faster).
template <typename T>
struct PanelXs {
    T v[panel_win<T>()];
    __device__ __forceinline__ void put(int i, T x) { v[i] = x; }
    __device__ __forceinline__ T get(uint32_t i) const { return v[i]; }
};

// seg[s]: the row in slot s has its entries of the current chunk at
// [seg & 0xffff, seg >> 16) of prod (both halves 0: none)
template <typename T, int BUDGET>
struct alignas(16) PanelSmem {
    PanelXs<T> xs;
    T prod[BUDGET];
    unsigned seg[kPanelRows + 4];  // [kPanelRows]: spare word for masked-off stores
};

template <typename T, int NT>
struct PanelX {
    static constexpr int XK = panel_win<T>() / NT;
    T v[XK];
};

// x[c0 + i] for i < w (an index past w re-reads x[c0]: in bounds, unused);
// live = false reads x[0] in every lane (never predicated, as spmv_load)
template <typename T, int NT>
__device__ __forceinline__ void panel_xload(const T *x, int4 ci, PanelX<T, NT> &X, bool live = true)
{
    const int t = threadIdx.x, c0 = chunk_c0(ci), w = ci.y - c0;
#pragma unroll
    for (int j = 0; j < PanelX<T, NT>::XK; j++) {
        const int i = j * NT + t;
        X.v[j] = x[live ? c0 + (i < w ? i : 0) : 0];
    }
}

// beta * y_in for one row, the value y starts from: Seldon's Mlt(beta, Y)
// before the sums (Functions_MatVect.cxx:414); ArrayRowSparse zero-fills Y
// when beta is 0 (Functions_MatrixArray.cxx:548-551).  Read where it is
// used, so the panel kernel holds no registers for it meanwhile.
template <typename T, int MODE>
__device__ __forceinline__ T panel_y0(const SpmvArgs<T> &a, int row)
{
    if (a.mlt || row < 0 || (MODE == kArray && a.beta == (T)0)) return (T)0;
    return a.y_in[row] * a.beta;
}

// One chunk.  On entry R holds the chunk's entries and X its x window.  The
// loads of a later chunk's window (into X) and of a later chunk's entries
// (into R2, a free register set) are issued first, together: every load then
// waits behind whole chunks of work, and the wait counters, which retire
// loads in issue order, never hold this chunk's reads behind newer ones.
// The caller rotates the register sets.
template <typename T, int MODE, int BUDGET, int NT>
__device__ __forceinline__ void panel_step(const SpmvArgs<T> &a, PanelSmem<T, BUDGET> &sm, int4 ci,
                                           const SpmvRegs<T, BUDGET, NT> &R, SpmvRegs<T, BUDGET, NT> &R2,
                                           PanelX<T, NT> &X, bool xnext, int4 cx, bool rnext, int4 c2,
                                           T (&acc)[kPanelRows / NT], const int (&row)[kPanelRows / NT])
{
    constexpr int RPT = kPanelRows / NT;
    constexpr int V = VecOf<T>::V, K = BUDGET / (NT * V);
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < PanelX<T, NT>::XK; j++) sm.xs.put(j * NT + t, X.v[j]);
    __builtin_amdgcn_sched_barrier(0);
    panel_xload<T, NT>(a.x, cx, X, xnext);
    spmv_load<T, BUDGET, NT>(a, c2.z, R2, rnext);
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();  // x window in; the previous chunk's sums are done
    // products a_j * x_col(j) into LDS, row segments marked; the x reads are
    // all issued before any LDS write
    T xv[K][V];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            const bool in = ci.z + (k * NT + t) * V + v < ci.w;
            xv[k][v] = sm.xs.get(in ? (R.iv[k][v] & kPanColMask) : 0u);
        }
#pragma unroll
    for (int k = 0; k < K; k++) {
        typename VecOf<T>::FV pv;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const bool in = ci.z + (k * NT + t) * V + v < ci.w;
            const T val = R.fv[k][v];
            // direct terms multiply alpha into the value first:
            // `alpha * data[j] * X(i)` (Functions_MatrixArray.cxx:571,
            // Functions_MatVect.cxx:563)
            const bool direct = MODE == kArray || (MODE == kSym && (R.iv[k][v] & kLowerFlag));
            const T pr = direct ? (a.alpha * val) * xv[k][v] : val * xv[k][v];
            pv[v] = in ? pr : (T)0;
        }
        *(typename VecOf<T>::FV *)(sm.prod + (k * NT + t) * V) = pv;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            const int el = (k * NT + t) * V + v;
            const bool in = ci.z + el < ci.w;
            const uint32_t w = R.iv[k][v];
            const int lr = (w >> kPanRowShift) & kPanRowMask;
            // unconditional stores (entries that are not a segment's first or
            // last write to a spare word): no per-store exec masking
            unsigned short *half = (unsigned short *)sm.seg;
            const bool f = in && (w & kPanFirst), l = in && (w & kPanLast);
            half[f ? 2 * lr : 2 * kPanelRows] = (unsigned short)el;
            half[l ? 2 * lr + 1 : 2 * kPanelRows + 1] = (unsigned short)(el + 1);
        }
    __syncthreads();
    if (MODE == kSym && ((uint32_t)ci.x & kChunkDirect)) {
        // the upper parts are done: y = beta y + alpha temp, then the direct
        // terms (Functions_MatVect.cxx:553-563)
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            T y = panel_y0<T, MODE>(a, row[i]);
            y += a.alpha * acc[i];
            acc[i] = y;
        }
    }
    // each thread adds its rows' segments in order, slot by slot (a slot's
    // loop runs as long as its longest segment in the wave), U products per
    // step read before they are added
    constexpr int U = SHK_PANEL_P2_UNROLL;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        const int rr = t + i * NT;
        const unsigned sw = sm.seg[rr];
        const int s0 = (int)(sw & 0xffffu);
        const int n0 = (int)(sw >> 16) - s0;
        for (int j = 0; j < n0; j += U) {
            T pv[U];
#pragma unroll
            for (int h = 0; h < U; h++) pv[h] = sm.prod[(j + h < n0) ? s0 + j + h : 0];
#pragma unroll
            for (int h = 0; h < U; h++) {
                const T sum = acc[i] + pv[h];
                acc[i] = (j + h < n0) ? sum : acc[i];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < RPT; i++) sm.seg[t + i * NT] = 0u;
}

template <typename T, int MODE, int BUDGET, int NT = kPanelThreads, int MINW = SHK_PANEL_MINW>
__global__ __launch_bounds__(NT, MINW) void k_panel(SpmvArgs<T> a, const int4 *__restrict__ chunk)
{
    if (a.skip && *a.skip) return;
    __shared__ PanelSmem<T, BUDGET> sm;
    constexpr int RPT = kPanelRows / NT;
    const int p = xcd_logical_block(a.xcd_q);
    const int t = threadIdx.x;
    T c[2] = {(T)0, (T)0};
    if (p < a.npanel) {
        const int r0 = p * kPanelRows;
        const int nr = min(kPanelRows, a.m - r0);
        T acc[RPT];
        int row[RPT];
        const unsigned short *srow = a.slot_row + (size_t)p * kPanelRows;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int rr = t + i * NT;
            row[i] = rr < nr ? r0 + srow[rr] : -1;
            // ArrayRowSparse adds the direct terms onto beta * y; RowSparse
            // and RowSymSparse sum a temp first
            acc[i] = MODE == kArray ? panel_y0<T, MODE>(a, row[i]) : (T)0;
            sm.seg[rr] = 0u;
        }
        bool direct = false;  // RowSymSparse: y formed, transposed parts started
        const int cb = a.pchunk[p], ce = a.pchunk[p + 1];
        if (cb < ce) {
            // NS register sets of entries rotate: chunk cc + u is in R[u]
            // and, while it is summed, chunk cc + u + NS - 1 loads into the
            // set freed by the previous chunk.  The x windows run XL chunks
            // ahead in XL sets: a window is waited for before it is staged,
            // and the wait counters retire loads in issue order, so the
            // entry loads stay in flight across that wait only when they
            // were issued after it.  The loop has one exit (full rounds of
            // NS chunks) and the last chunks run after it: a loop left from
            // the middle of its body is rebuilt by the compiler with paths
            // that carry in-flight loads back to its head, where it then
            // waits for every load.  Chunk descriptors are read a step
            // before their loads are issued (index clamped, no branch).
            constexpr int NS = panel_sets<T>(), XL = panel_xlead<T>();
            static_assert(NS % XL == 0 && XL < NS, "x window sets must divide the round");
            SpmvRegs<T, BUDGET, NT> R[NS];
            PanelX<T, NT> X[XL];
            auto desc = [&](int k) { return chunk[k < ce ? k : ce - 1]; };
            int4 d[NS];
#pragma unroll
            for (int j = 0; j < NS; j++) d[j] = desc(cb + j);
            spmv_load<T, BUDGET, NT>(a, d[0].z, R[0]);
#pragma unroll
            for (int j = 0; j < XL; j++) panel_xload<T, NT>(a.x, d[j], X[j], cb + j < ce);
#pragma unroll
            for (int j = 1; j < NS - 1; j++) spmv_load<T, BUDGET, NT>(a, d[j].z, R[j], cb + j < ce);
            int cc = cb;
            for (; cc + NS <= ce; cc += NS) {
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    const int4 dn = desc(cc + u + NS);
                    direct |= ((uint32_t)d[u].x & kChunkDirect) != 0;
                    panel_step<T, MODE, BUDGET, NT>(a, sm, d[u], R[u], R[(u + NS - 1) % NS], X[u % XL],
                                                    cc + u + XL < ce, d[(u + XL) % NS], cc + u + NS - 1 < ce,
                                                    d[(u + NS - 1) % NS], acc, row);
                    d[u] = dn;
                }
            }
            // fewer than NS chunks left, in R[0], R[1], ...
            const int left = ce - cc;
#pragma unroll
            for (int u = 0; u < NS - 1; u++) {
                if (u >= left) break;
                direct |= ((uint32_t)d[u].x & kChunkDirect) != 0;
                panel_step<T, MODE, BUDGET, NT>(a, sm, d[u], R[u], R[(u + NS - 1) % NS], X[u % XL], u + XL < left,
                                                d[(u + XL) % NS], false, d[u], acc, row);
            }
        }
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            if (row[i] >= 0) {
                T y;
                if (MODE == kArray || (MODE == kSym && direct)) {
                    y = acc[i];
                } else {
                    y = panel_y0<T, MODE>(a, row[i]);
                    y += a.alpha * acc[i];
                }
                spmv_epilogue<T, MODE>(a, row[i], y, c);
            }
        }
    }
    if (a.dot0 != kDotNone || a.dot1 != kDotNone) block_partial<T, 2>(c, a.partials, gridDim.x, blockIdx.x);
}

}  // namespace shk

#include "shk_sliced.hpp"

namespace shk {

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
    a.npanel = A->npanel;
    a.pchunk = A->pchunk;
    a.slot_row = A->slot_row;
    const bool dots = a.dot0 != kDotNone || a.dot1 != kDotNone;
    if (dots && !a.partials) {
        SHK_TRY(ensure_red(ctx, sizeof(T) * 2 * (size_t)A->nblk_launch));
        a.partials = (T *)ctx->red.partials;
    }
    constexpr bool defer = std::is_same<Fin, FinDefer>::value;
    if (A->m == 0) return SHK_OK;
    if (A->layout == SHK_LAYOUT_SLICED) {
        SHK_CHECK(A->sl_pchunk && A->sl_desc && A->sl_wstart && A->sl_bits && A->sl_rowmap && A->sl_col &&
                      A->sl_val,
                  SHK_ERR_STATE, "spmv: sliced layout arrays missing");
        const SlicedView v{A->sl_npanel,
                           A->n,
                           A->sl_pchunk,
                           (const SlDesc *)A->sl_desc,
                           (const int2 *)A->sl_wstart,
                           (const unsigned short *)A->sl_bits,
                           A->sl_bits_bytes,
                           A->sl_rowmap,
                           A->sl_col,
                           A->sl_stream_len};
        a.val = (const T *)A->sl_val;
        a.xcd_q = A->sl_q;
        dim3 grid(A->sl_launch), block(kSlThreads);
        hipLaunchKernelGGL((k_sliced<T>), grid, block, 0, ctx->stream, a, v);
        a.grid_used = (int)grid.x;
        if (dots && !defer) grid_finish<T>(ctx->stream, a.partials, (int)grid.x, a.skip, fin);
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }
    SHK_CHECK(A->layout == SHK_LAYOUT_PANELS ? (A->chunk && A->pchunk && A->slot_row) : (A->blk_info != nullptr),
              SHK_ERR_STATE, "spmv: layout arrays missing");
    if (A->layout == SHK_LAYOUT_PANELS) {
        constexpr int PB = panel_budget<T>();
        dim3 grid(A->nblk_launch), block(kPanelThreads);
        const int4 *chunk = (const int4 *)A->chunk;
        if (A->storage == SHK_ARRAYROWSPARSE)
            hipLaunchKernelGGL((k_panel<T, kArray, PB>), grid, block, 0, ctx->stream, a, chunk);
        else if (A->storage == SHK_ROWSYMSPARSE)
            hipLaunchKernelGGL((k_panel<T, kSym, PB>), grid, block, 0, ctx->stream, a, chunk);
        else
            hipLaunchKernelGGL((k_panel<T, kCsr, PB>), grid, block, 0, ctx->stream, a, chunk);
        a.grid_used = (int)grid.x;
        if (dots && !defer) grid_finish<T>(ctx->stream, a.partials, (int)grid.x, a.skip, fin);
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }
    constexpr int B = budget_of<T>();
    // persistent grid: at most kSpmvWgPerCu work-groups per CU, a multiple of 8
    const int per_xcd = std::min(A->xcd_q, std::max(1, ctx->num_cus * kSpmvWgPerCu / kNumXcd));
    dim3 grid(per_xcd * kNumXcd), block(kBlock);
    switch (A->storage) {
    case SHK_ROWSPARSE:
        hipLaunchKernelGGL((k_spmv<T, kCsr, B>), grid, block, 0, ctx->stream, a, a.blk_info);
        break;
    case SHK_ROWSYMSPARSE:
        hipLaunchKernelGGL((k_spmv<T, kSym, B>), grid, block, 0, ctx->stream, a, a.blk_info);
        break;
    default:
        hipLaunchKernelGGL((k_spmv<T, kArray, B>), grid, block, 0, ctx->stream, a, a.blk_info);
        break;
    }
    a.grid_used = (int)grid.x;
    if (dots && !defer) grid_finish<T>(ctx->stream, a.partials, (int)grid.x, a.skip, fin);
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
// Add does), then block sums of u0*v and (optional) u1*v over v = y; the
// totals come from grid_finish on the same stream.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_dot(int64_t n, T alpha, const T *x, T *y,
                                                     const T *u0, const T *u1, T *partials, const int *skip)
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
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

}  // namespace shk

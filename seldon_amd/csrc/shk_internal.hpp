// shk_internal.hpp -- shared internals of the shk HIP library (gfx950).
//
// Contents: error plumbing, the context / matrix handle layouts, the
// deterministic reduction helpers used by every fused kernel, and the
// XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "shk.h"

namespace shk {

// ------------------------------------------------------------- errors ----
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SHK_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return ::shk::fail(SHK_ERR_HIP, std::string(#call) + ": " +        \
                                                hipGetErrorString(e_));        \
    } while (0)

#define SHK_CHECK(cond, code, msg)                                             \
    do {                                                                       \
        if (!(cond)) return ::shk::fail((code), (msg));                        \
    } while (0)

#define SHK_TRY(expr)                                                          \
    do {                                                                       \
        int s_ = (expr);                                                       \
        if (s_ != SHK_OK) return s_;                                           \
    } while (0)

// ------------------------------------------------------------- layout ----
constexpr int kBlock = 256;         // threads per work-group for most kernels
constexpr int kWave = 64;           // CDNA wavefront
constexpr int kNumXcd = 8;          // MI355X: 8 XCDs, work-groups dealt round-robin
constexpr int kMaxRedBlocks = 2048; // BLAS-1 reduction grid cap (fixed per n)
constexpr uint32_t kLowerFlag = 0x80000000u; // ind high bit: "direct" term

inline size_t dsize(int dt) { return dt == SHK_F32 ? 4 : 8; }

// Device-side scalar workspace for one in-flight reduction family.
struct RedSlot {
    void *partials = nullptr; // [nv][nblocks] of T
    unsigned *ticket = nullptr; // counters of the batched solver's work queue
    size_t cap_bytes = 0;
};

}  // namespace shk

namespace shk {
struct SorPlan;  // level-scheduled SOR sweeps (shk_sor.hip)
void sor_free(SorPlan *P);
}  // namespace shk

struct shk_ctx_s {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    char arch[64] = {0};
    // reduction workspace: block partials of the split reductions
    shk::RedSlot red;
    // small device scratch for scalar results and solver state
    void *scalars = nullptr;      // 4 KiB
    void *state = nullptr;        // solver state (device)
    void *state_host = nullptr;   // pinned mirror
    size_t state_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // grow-only work-vector arena for the solvers (Seldon allocates its
    // work vectors per call, Cg.cxx:55; here they are reused across calls)
    void *ws = nullptr;
    size_t ws_cap = 0;
    // second arena for drivers that call a solver (iterative refinement)
    void *ws2 = nullptr;
    size_t ws2_cap = 0;
};

struct shk_csr_s {
    shk_ctx ctx = nullptr;
    int storage = SHK_ROWSPARSE;
    int dtype = SHK_F64;
    int m = 0, n = 0;
    int64_t nnz_stored = 0;   // as given by the caller
    int64_t nnz_dev = 0;      // after the symmetric expansion
    int *ptr = nullptr;       // [m+1]
    int *mid = nullptr;       // [m] RowSymSparse: start of the transposed part
    uint32_t *ind = nullptr;  // [nnz_dev + pad], kLowerFlag on direct terms
    void *val = nullptr;      // [nnz_dev + pad]
    void *diag = nullptr;     // [m] a_ii for Jacobi (0 when absent)
    int *blk_rows = nullptr;  // [nblk+1] row-block boundaries
    int *blk_info = nullptr;  // [4*nblk] {rs, re, ptr[rs], ptr[re]} per block
    int nblk = 0;             // logical row blocks
    int nblk_launch = 0;      // launched blocks (padded to a multiple of 8)
    int xcd_q = 0;            // logical blocks per XCD
    int budget = 0;           // products staged per block (LDS)
    int max_row = 0;
    // RowSymSparse values: device position of each stored entry and of its
    // transposed copy (-1 for the diagonal) so new values can be scattered.
    int64_t *map_up = nullptr;
    int64_t *map_lo = nullptr;
    int64_t *map_diag = nullptr; // [m] position of a_ii on device (-1: none)
    // Row-block layout: RowSparse / ArrayRowSparse values are stored in the
    // caller's order.  Panel layout: see k_panel (the maps above point into
    // the panel arrays; ptr keeps the caller's row pointer).
    int layout = 0;           // SHK_LAYOUT_ROWS or SHK_LAYOUT_PANELS
    int64_t nslots = 0;       // length of ind/val without the load padding
    int npanel = 0, nchunk = 0;
    int *pchunk = nullptr;    // [npanel+1] first chunk of each panel
    int *chunk = nullptr;     // [4*nchunk] {c0, c1, e0, e1} per chunk
    unsigned short *slot_row = nullptr;  // [npanel*kPanelRows] panel-local row of each row slot
    // Sliced layout (RowSparse): the row layout above serves every other
    // use; the SpMV reads this copy of the entries (shk_sliced.hpp)
    int sl_npanel = 0, sl_nchunk = 0, sl_launch = 0, sl_q = 0;
    int64_t sl_ns = 0;                 // stream length
    int sl_stream_len = 0;             // with its padding
    int sl_bits_bytes = 0;
    int *sl_pchunk = nullptr;          // [sl_npanel + 1]
    void *sl_desc = nullptr;           // [sl_nchunk * 8] SlDesc
    void *sl_wstart = nullptr;         // [sl_npanel * 8] int2
    void *sl_bits = nullptr;           // [blocks * 64] activity bits
    unsigned short *sl_rowmap = nullptr;
    unsigned short *sl_col = nullptr;  // [sl_ns] window-local columns
    void *sl_val = nullptr;            // [sl_ns] values
    int *sl_src = nullptr;             // [sl_ns] caller's entry index of each stream position
    // A^T for the transposed products (RowSparse, ArrayRowSparse), built on
    // first use as an ArrayRowSparse matrix whose rows list, for each column
    // c of A, the entries (i, c) in increasing i
    shk_csr_s *trans = nullptr;
    // SOR / SSOR sweeps: the caller's CSR and the row levels of both sweep
    // directions, built on first use
    shk::SorPlan *sor = nullptr;
};

namespace shk {

// ----------------------------------------------------- device helpers ----

// Logical block index for an XCD-aware launch of `nblk_launch` = 8*q blocks:
// hardware block b runs on XCD (b % 8); logical blocks [k*q, (k+1)*q) all
// land on XCD k, so neighbouring row blocks share one L2.
__device__ __forceinline__ int xcd_logical_block(int q)
{
    int b = blockIdx.x;
    return (b % kNumXcd) * q + b / kNumXcd;
}

template <typename T>
__device__ __forceinline__ T wave_sum_lane0(T v)
{
    // xor butterfly; every lane ends with a full sum, lane 0's is the one used
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block reduction of NV values, fixed order: lanes by xor butterfly (lane 0),
// waves summed sequentially by thread 0.  Result valid in thread 0.
template <typename T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T *lds /*[NV*16]*/)
{
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        T s = wave_sum_lane0(v[k]);
        if (lane == 0) lds[k * 16 + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            T s = lds[k * 16];
            for (int w = 1; w < nw; w++) s += lds[k * 16 + w];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Grid reductions are split in two launches.  Every block of the producing
// kernel calls block_partial once: its block sum goes to
// partials[k*nblocks + b].  The stream then runs k_grid_finish, one block
// that adds the partials in a fixed order and hands the totals to the
// finalize functor.  (A last-block-done ticket would need a device-scope
// release per block, which on this multi-XCD part writes back the XCD's L2
// each time: it cost ~27 ns per block, serialised, far more than the launch
// of the second kernel.)  Results are reproducible run to run: the grid of
// every producer is a function of the problem size only.
template <typename T, int NV>
__device__ __forceinline__ void block_partial(T (&v)[NV], T *partials, int nblocks, int bidx)
{
    __shared__ T lds[NV * 16];
    block_sum<T, NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partials[(size_t)k * nblocks + bidx] = v[k];
    }
}

// The fixed-order total of a split reduction's partials, valid in thread 0
// (a kBlock-thread work-group): the finish kernel below uses it, and so do
// kernels that fold the previous reduction's finish into their own start.
template <typename T>
__device__ __forceinline__ void grid_total(const T *partials, int nblocks, T (&out)[2])
{
    __shared__ T lds[2 * 16];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        T s = 0;
        for (int i = threadIdx.x; i < nblocks; i += kBlock) s += partials[(size_t)k * nblocks + i];
        out[k] = s;
    }
    block_sum<T, 2>(out, lds);
}

template <typename T, typename Fin>
__global__ __launch_bounds__(kBlock) void k_grid_finish(const T *partials, int nblocks, const int *skip, Fin fin)
{
    if (skip && *skip) return;
    T out[2];
    grid_total<T>(partials, nblocks, out);
    if (threadIdx.x == 0) fin(out);
}

// The same for six values (at partials + k nblocks), in one pass: each
// value's per-thread sums and block sum in the same order as grid_total's.
template <typename T, typename Fin>
__global__ __launch_bounds__(kBlock) void k_grid_finish6(const T *partials, int nblocks, const int *skip, Fin fin)
{
    if (skip && *skip) return;
    __shared__ T lds[6 * 16];
    T out[6];
#pragma unroll
    for (int k = 0; k < 6; k++) out[k] = (T)0;
    for (int i = threadIdx.x; i < nblocks; i += kBlock) {
#pragma unroll
        for (int k = 0; k < 6; k++) out[k] += partials[(size_t)k * nblocks + i];
    }
    block_sum<T, 6>(out, lds);
    if (threadIdx.x == 0) fin(out);
}

// The second launch of a split grid reduction (see block_partial).
template <typename T, typename Fin>
inline void grid_finish(hipStream_t st, const T *partials, int nblocks, const int *skip, Fin fin)
{
    hipLaunchKernelGGL((k_grid_finish<T, Fin>), dim3(1), dim3(kBlock), 0, st, partials, nblocks, skip, fin);
}

// number of blocks used by the grid-stride BLAS-1 reductions (fixed per n)
inline int red_blocks(int64_t n, int items_per_thread = 8)
{
    int64_t b = (n + (int64_t)kBlock * items_per_thread - 1) / ((int64_t)kBlock * items_per_thread);
    if (b < 1) b = 1;
    if (b > kMaxRedBlocks) b = kMaxRedBlocks;
    return (int)b;
}

// ----------------------------------------------------- host helpers ----
int ensure_red(shk_ctx ctx, size_t bytes);
int ensure_ws(shk_ctx ctx, size_t bytes);
// The caller's stored arrays (RowSymSparse: upper part) read back from the
// device, in the caller's order (shk_spmv.hip; T = double, float).
template <typename T>
int stored_arrays(shk_csr A, std::vector<int> &ptr, std::vector<int> &ind, std::vector<T> &val);
int ensure_ws2(shk_ctx ctx, size_t bytes);
// device address of shk_cg_noalloc's fixed scratch array, looked up once per
// device with that device current (shk_solvers.hip)
int noalloc_scratch_init(int device);
// sliced layout (shk_sliced.hip): plan and upload the SpMV copy of a
// RowSparse matrix already uploaded in the row layout (*applies = false when
// the layout does not fit it); refresh its values from the row layout's;
// free it
int sliced_build(shk_csr A, const int *ptr, const int *ind, bool *applies);
int sliced_set_values(shk_csr A);
// the matrix behind A's transposed products (built on first use)
int transpose_of(shk_csr A, shk_csr *out);
void sliced_free(shk_csr A);
// AUTO picks the sliced layout for RowSparse matrices with at least this
// many rows (256 panels of 8192 rows)
constexpr int kSlAutoMinRows = 1 << 21;
int check_device_ptr(const void *p, const char *what);

}  // namespace shk

// shk_spmv.hip -- sparse matrix upload, row-block planning and the SpMV
// entry points of the C ABI.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "shk_kernels.hpp"

namespace shk {

// ------------------------------------------------------------ planning ----
// Greedy row blocks: at most kBlock rows and `budget` entries (counted from
// the block's first entry rounded down to the load vector width) per block;
// a row longer than that is a block by itself.
static void plan_blocks(int m, const std::vector<int> &ptr, int budget, int align, std::vector<int> &blk)
{
    blk.clear();
    blk.push_back(0);
    int r = 0;
    while (r < m) {
        int start = r;
        // the kernel stages from the vector-aligned start of the block
        int64_t base = ptr[r] & ~(align - 1);
        if (ptr[r + 1] - base > budget) {
            r++;
        } else {
            while (r < m && r - start < kBlock && ptr[r + 1] - base <= budget) r++;
        }
        blk.push_back(r);
    }
}

template <typename T>
static __global__ void k_scatter_vals(int64_t n, const T *src, const int64_t *map_up,
                                      const int64_t *map_lo, T *dst)
{
    for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
        T v = src[j];
        dst[map_up[j]] = v;
        if (map_lo[j] >= 0) dst[map_lo[j]] = v;
    }
}

template <typename T>
static __global__ void k_gather_diag(int m, const int64_t *pos, const T *val, T *diag)
{
    for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < m; r += (int64_t)gridDim.x * kBlock)
        diag[r] = pos[r] >= 0 ? val[pos[r]] : (T)0;
}

struct HostPlan {
    std::vector<int> ptr, mid, blk;
    std::vector<uint32_t> ind;
    std::vector<int64_t> src_of;   // expanded position -> stored entry index
    std::vector<int64_t> map_up, map_lo, diag_pos;
    int max_row = 0;
};

// Build the device layout (expanded for RowSymSparse) from the caller's CSR.
static int build_plan(int storage, int m, int n, int64_t nnz, const int *ptr, const int *ind,
                      int budget, int align, HostPlan &P)
{
    SHK_CHECK(ptr[0] == 0 && ptr[m] == nnz, SHK_ERR_ARG, "csr: ptr[0] must be 0 and ptr[m] == nnz");
    for (int i = 0; i < m; i++)
        SHK_CHECK(ptr[i + 1] >= ptr[i], SHK_ERR_ARG, "csr: ptr must be non-decreasing");
    for (int64_t j = 0; j < nnz; j++)
        SHK_CHECK(ind[j] >= 0 && ind[j] < n, SHK_ERR_ARG, "csr: column index out of range");
    SHK_CHECK(nnz < (int64_t)INT32_MAX, SHK_ERR_ARG, "csr: nnz must fit in int (Seldon's int indices)");
    P.diag_pos.assign(m, -1);
    if (storage == SHK_ROWSPARSE || storage == SHK_ARRAYROWSPARSE) {
        P.ptr.assign(ptr, ptr + m + 1);
        P.ind.resize(nnz);
        const uint32_t flag = storage == SHK_ARRAYROWSPARSE ? kLowerFlag : 0u;
        for (int64_t j = 0; j < nnz; j++) P.ind[j] = (uint32_t)ind[j] | flag;
        for (int i = 0; i < m; i++)
            for (int j = ptr[i]; j < ptr[i + 1]; j++)
                if (ind[j] == i) P.diag_pos[i] = j;  // last match, as the oracle
        P.src_of.clear();
    } else {
        SHK_CHECK(m == n, SHK_ERR_DIM, "RowSymSparse needs a square matrix");
        // transposed part: the scatter of Functions_MatVect.cxx:560-563 runs
        // i = 0..m-2, entries j of row i with ind[j] != i, adding to row ind[j]
        std::vector<int64_t> cnt(m + 1, 0);
        for (int i = 0; i < m - 1; i++)
            for (int j = ptr[i]; j < ptr[i + 1]; j++)
                if (ind[j] != i) cnt[ind[j]]++;
        int64_t total = nnz;
        for (int i = 0; i < m; i++) total += cnt[i];
        SHK_CHECK(total < (int64_t)INT32_MAX, SHK_ERR_ARG, "csr: expanded nnz must fit in int");
        P.ptr.assign(m + 1, 0);
        P.mid.assign(m, 0);
        for (int i = 0; i < m; i++) {
            int64_t up = ptr[i + 1] - ptr[i];
            P.ptr[i + 1] = (int)(P.ptr[i] + up + cnt[i]);
            P.mid[i] = (int)(P.ptr[i] + up);
        }
        P.ind.resize(total);
        P.src_of.resize(total);
        P.map_up.resize(nnz);
        P.map_lo.assign(nnz, -1);
        std::vector<int64_t> fillpos(m);
        for (int i = 0; i < m; i++) {
            int64_t pos = P.ptr[i];
            for (int j = ptr[i]; j < ptr[i + 1]; j++, pos++) {
                P.ind[pos] = (uint32_t)ind[j];
                P.src_of[pos] = j;
                P.map_up[j] = pos;
                if (ind[j] == i) P.diag_pos[i] = pos;
            }
            fillpos[i] = P.mid[i];
        }
        for (int i = 0; i < m - 1; i++)
            for (int j = ptr[i]; j < ptr[i + 1]; j++)
                if (ind[j] != i) {
                    int c = ind[j];
                    int64_t pos = fillpos[c]++;
                    P.ind[pos] = (uint32_t)i | kLowerFlag;
                    P.src_of[pos] = j;
                    P.map_lo[j] = pos;
                }
    }
    for (int i = 0; i < m; i++) P.max_row = std::max(P.max_row, P.ptr[i + 1] - P.ptr[i]);
    plan_blocks(m, P.ptr, budget, align, P.blk);
    return SHK_OK;
}

// ------------------------------------------------------- panel layout ----
// Host side of the layout read by k_panel: per panel of kPanelRows rows, the
// entries are cut into chunks by column windows.  The cut walks the panel's
// columns in increasing order and closes a chunk before it would exceed
// `budget` entries or panel_win columns; every chunk then lists, row after
// row, that row's entries inside the window (a contiguous run, the row's
// columns being sorted).  Chunk starts are aligned to the load width.
struct PanelPlan {
    int npanel = 0;
    std::vector<unsigned short> slot_row;  // [npanel*kPanelRows] panel-local row of each slot
    std::vector<int> pchunk;        // [npanel+1]
    std::vector<int> chunk;         // 4 per chunk: c0, c1, e0, e1
    std::vector<uint32_t> word;     // packed entry words
    std::vector<int64_t> src;       // panel position -> expanded CSR index (-1: padding)
    std::vector<int64_t> diag_pos;  // [m] panel position of a_ii
    int64_t nslots = 0;
    double x_bytes = 0;             // x window bytes per SpMV
};

// Build the panel layout, or return false when it does not apply: a row
// with unsorted columns, or x windows that would cost more traffic than the
// matrix itself (columns spread far from the diagonal).
static bool plan_panels(int m, int n, const HostPlan &P, int budget, int win, int align, size_t vsize,
                        PanelPlan &Q)
{
    if (m == 0) return false;
    // RowSymSparse rows are [upper part ; transposed part] (mid[r] splits
    // them): the chunks of a panel walk the upper parts first, then -- from
    // the chunk flagged kChunkDirect on -- the transposed parts, each part
    // column-sorted, so every expanded row is still visited in stored order.
    const bool sym = !P.mid.empty();
    auto part_lo = [&](int r, int pass) { return pass == 0 ? P.ptr[r] : P.mid[r]; };
    auto part_hi = [&](int r, int pass) { return pass == 0 ? (sym ? P.mid[r] : P.ptr[r + 1]) : P.ptr[r + 1]; };
    const int npass = sym ? 2 : 1;
    for (int pass = 0; pass < npass; pass++)
        for (int r = 0; r < m; r++)
            for (int j = part_lo(r, pass) + 1; j < part_hi(r, pass); j++)
                if ((P.ind[j] & ~kLowerFlag) < (P.ind[j - 1] & ~kLowerFlag)) return false;
    const int64_t nnz = P.ptr[m];
    Q.npanel = (m + kPanelRows - 1) / kPanelRows;
    Q.pchunk.assign(Q.npanel + 1, 0);
    Q.chunk.clear();
    Q.word.clear();
    Q.src.clear();
    Q.diag_pos.assign(m, -1);
    Q.slot_row.assign((size_t)Q.npanel * kPanelRows, 0);
    // Panels are planned in parallel, each into its own arrays with positions
    // counted from the panel's start, then laid end to end.  A panel's arrays
    // start at a multiple of `align` (the previous panel is padded at its
    // end), so the padding before each chunk is the same as in one
    // sequential pass and the result is identical.
    struct Local {
        std::vector<int> chunk;  // c0 | flag, c1, e0, e1 (e0, e1 from the panel's start)
        std::vector<uint32_t> word;
        std::vector<int64_t> src;
        std::vector<std::pair<int, int64_t>> diag;  // (row, position from the panel's start)
        double x_bytes = 0;
        bool ok = true;
    };
    std::vector<Local> L(Q.npanel);
    auto plan_one = [&](int p, Local &Lp) {
        std::vector<int> cols, cur(kPanelRows), slot_of(kPanelRows), order(kPanelRows);
        const int r0 = p * kPanelRows, r1 = std::min(m, r0 + kPanelRows);
        // Row slots by length, longest first: the panel kernel gives slot s
        // to one thread, so the panel's long rows land in the same wavefronts
        // and the waves of short rows finish their sums early.
        const int nr = r1 - r0;
        for (int k = 0; k < nr; k++) order[k] = k;
        std::stable_sort(order.begin(), order.begin() + nr, [&](int a, int b) {
            return P.ptr[r0 + a + 1] - P.ptr[r0 + a] > P.ptr[r0 + b + 1] - P.ptr[r0 + b];
        });
        for (int k = 0; k < nr; k++) {
            slot_of[order[k]] = k;
            Q.slot_row[(size_t)p * kPanelRows + k] = (unsigned short)order[k];
        }
        for (int pass = 0; pass < npass; pass++) {
            cols.clear();
            for (int r = r0; r < r1; r++)
                for (int j = part_lo(r, pass); j < part_hi(r, pass); j++) cols.push_back((int)(P.ind[j] & ~kLowerFlag));
            std::sort(cols.begin(), cols.end());
            for (int r = r0; r < r1; r++) cur[r - r0] = part_lo(r, pass);
            bool first_chunk = true;
            size_t i = 0;
            while (i < cols.size()) {
                const int c0 = cols[i];
                size_t k = i;
                // whole columns only, at most `budget` entries and `win` columns
                while (k < cols.size() && cols[k] - c0 < win) {
                    size_t k2 = k;
                    while (k2 < cols.size() && cols[k2] == cols[k]) k2++;
                    if ((int64_t)(k2 - i) > budget) break;
                    k = k2;
                }
                if (k == i) {  // one column with more entries than a chunk holds
                    Lp.ok = false;
                    return;
                }
                const int c1 = cols[k - 1] + 1;
                while ((int64_t)Lp.word.size() % align) {
                    Lp.word.push_back(0u);
                    Lp.src.push_back(-1);
                }
                const int64_t e0 = (int64_t)Lp.word.size();
                for (int r = r0; r < r1; r++) {
                    int j = cur[r - r0];
                    const int jb = j;
                    while (j < part_hi(r, pass) && (int)(P.ind[j] & ~kLowerFlag) < c1) j++;
                    for (int q = jb; q < j; q++) {
                        const uint32_t col = P.ind[q] & ~kLowerFlag;
                        uint32_t w = (col - (uint32_t)c0) | ((uint32_t)slot_of[r - r0] << kPanRowShift) |
                                     (P.ind[q] & kLowerFlag);
                        if (q == jb) w |= kPanFirst;
                        if (q == j - 1) w |= kPanLast;
                        if (P.diag_pos[r] == q) Lp.diag.emplace_back(r, (int64_t)Lp.word.size());
                        Lp.word.push_back(w);
                        Lp.src.push_back(q);
                    }
                    cur[r - r0] = j;
                }
                const int64_t e1 = (int64_t)Lp.word.size();
                Lp.chunk.push_back((int)((uint32_t)c0 | (pass == 1 && first_chunk ? kChunkDirect : 0u)));
                Lp.chunk.push_back(c1);
                Lp.chunk.push_back((int)std::min<int64_t>(e0, INT32_MAX));
                Lp.chunk.push_back((int)std::min<int64_t>(e1, INT32_MAX));
                Lp.x_bytes += (double)(c1 - c0) * (double)vsize;
                first_chunk = false;
                i = k;
            }
            for (int r = r0; r < r1; r++)
                if (cur[r - r0] != part_hi(r, pass)) {  // not reached with sorted parts
                    Lp.ok = false;
                    return;
                }
        }
    };
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++)
        th.emplace_back([&, t]() {
            for (int p = t; p < Q.npanel; p += nth) plan_one(p, L[p]);
        });
    for (auto &t : th) t.join();
    size_t total = 0;
    for (const Local &Lp : L) {
        if (!Lp.ok) return false;
        total += Lp.word.size() + (size_t)align;
    }
    Q.word.reserve(total);
    Q.src.reserve(total);
    for (int p = 0; p < Q.npanel; p++) {
        Local &Lp = L[p];
        while ((int64_t)Q.word.size() % align) {
            Q.word.push_back(0u);
            Q.src.push_back(-1);
        }
        const int64_t base = (int64_t)Q.word.size();
        if (base + (int64_t)Lp.word.size() >= (int64_t)INT32_MAX) return false;
        for (size_t c = 0; c < Lp.chunk.size(); c += 4) {
            Q.chunk.push_back(Lp.chunk[c]);
            Q.chunk.push_back(Lp.chunk[c + 1]);
            Q.chunk.push_back((int)(base + Lp.chunk[c + 2]));
            Q.chunk.push_back((int)(base + Lp.chunk[c + 3]));
        }
        for (const auto &d : Lp.diag) Q.diag_pos[d.first] = base + d.second;
        Q.word.insert(Q.word.end(), Lp.word.begin(), Lp.word.end());
        Q.src.insert(Q.src.end(), Lp.src.begin(), Lp.src.end());
        Q.x_bytes += Lp.x_bytes;
        Q.pchunk[p + 1] = (int)(Q.chunk.size() / 4);
        Local().word.swap(Lp.word);
        std::vector<int64_t>().swap(Lp.src);
    }
    Q.nslots = (int64_t)Q.word.size();
    (void)n;
    // the windows replace per-entry gathers; they must not cost more bytes
    // than the entries they serve
    return Q.x_bytes <= 1.5 * (double)nnz * (double)(4 + vsize);
}

template <typename T>
static __global__ void k_panel_scatter(const int *pchunk, const int4 *chunk, const uint32_t *word,
                                       const unsigned short *slot_row, const int *ptr, const T *src, T *dst)
{
    // Positions of the caller's stored entries in the panel arrays: one
    // work-group per panel walks its chunks in order, counting per row the
    // entries already placed (a row's k-th entry lands in the chunk whose
    // window holds its column).
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
            const int lr = (word[e] >> kPanRowShift) & kPanRowMask;
            dst[e] = src[ptr[r0 + slot_row[(size_t)p * kPanelRows + lr]] + cnt[lr] + (e - segs[lr])];
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

template <typename T>
static int upload_panels(shk_csr A, const HostPlan &P, const PanelPlan &Q, const T *val_host)
{
    const int m = A->m;
    const int64_t ns = Q.nslots;
    const int64_t pad = A->budget + 16;  // k_panel streams whole budget windows
    const int nchunk = (int)(Q.chunk.size() / 4);
    A->nslots = ns;
    A->npanel = Q.npanel;
    A->nchunk = nchunk;
    std::vector<uint32_t> word(ns + pad, 0u);
    std::copy(Q.word.begin(), Q.word.end(), word.begin());
    hipStream_t s = A->ctx->stream;
    SHK_HIP(hipMalloc(&A->ptr, sizeof(int) * (m + 1)));
    SHK_HIP(hipMemcpyAsync(A->ptr, P.ptr.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
    SHK_HIP(hipMalloc(&A->ind, sizeof(uint32_t) * (ns + pad)));
    SHK_HIP(hipMalloc(&A->val, sizeof(T) * (ns + pad)));
    SHK_HIP(hipMalloc(&A->diag, sizeof(T) * std::max(m, 1)));
    SHK_HIP(hipMalloc(&A->pchunk, sizeof(int) * (Q.npanel + 1)));
    SHK_HIP(hipMalloc(&A->slot_row, sizeof(unsigned short) * Q.slot_row.size()));
    SHK_HIP(hipMemcpyAsync(A->slot_row, Q.slot_row.data(), sizeof(unsigned short) * Q.slot_row.size(),
                           hipMemcpyHostToDevice, s));
    SHK_HIP(hipMalloc(&A->chunk, sizeof(int) * std::max<size_t>(Q.chunk.size(), 4)));
    SHK_HIP(hipMalloc(&A->map_diag, sizeof(int64_t) * std::max(m, 1)));
    SHK_HIP(hipMemcpyAsync(A->ind, word.data(), sizeof(uint32_t) * (ns + pad), hipMemcpyHostToDevice, s));
    SHK_HIP(hipMemcpyAsync(A->pchunk, Q.pchunk.data(), sizeof(int) * (Q.npanel + 1), hipMemcpyHostToDevice, s));
    if (nchunk)
        SHK_HIP(hipMemcpyAsync(A->chunk, Q.chunk.data(), sizeof(int) * Q.chunk.size(), hipMemcpyHostToDevice, s));
    SHK_HIP(hipMemcpyAsync(A->map_diag, Q.diag_pos.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, s));
    std::vector<T> val(ns + pad, (T)0);
    const bool sym = !P.mid.empty();
    if (val_host)
        for (int64_t q = 0; q < ns; q++)
            if (Q.src[q] >= 0) val[q] = val_host[sym ? P.src_of[Q.src[q]] : Q.src[q]];
    if (sym) {
        // stored entry -> panel position of its upper copy and of its
        // transposed copy, for value updates and the gradient fold
        std::vector<int64_t> up(std::max<int64_t>(A->nnz_stored, 1), -1), lo(std::max<int64_t>(A->nnz_stored, 1), -1);
        std::vector<int> row_of(P.ind.size());
        for (int r = 0; r < m; r++)
            for (int q = P.ptr[r]; q < P.ptr[r + 1]; q++) row_of[q] = r;
        for (int64_t pos = 0; pos < ns; pos++) {
            const int64_t q = Q.src[pos];
            if (q < 0) continue;
            if (q < P.mid[row_of[q]]) up[P.src_of[q]] = pos;
            else lo[P.src_of[q]] = pos;
        }
        SHK_HIP(hipMalloc(&A->mid, sizeof(int) * std::max(m, 1)));
        if (m) SHK_HIP(hipMemcpyAsync(A->mid, P.mid.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
        SHK_HIP(hipMalloc(&A->map_up, sizeof(int64_t) * up.size()));
        SHK_HIP(hipMalloc(&A->map_lo, sizeof(int64_t) * lo.size()));
        SHK_HIP(hipMemcpyAsync(A->map_up, up.data(), sizeof(int64_t) * up.size(), hipMemcpyHostToDevice, s));
        SHK_HIP(hipMemcpyAsync(A->map_lo, lo.data(), sizeof(int64_t) * lo.size(), hipMemcpyHostToDevice, s));
        SHK_HIP(hipStreamSynchronize(s));
    }
    SHK_HIP(hipMemcpyAsync(A->val, val.data(), sizeof(T) * (ns + pad), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gather_diag<T>, dim3(red_blocks(m)), dim3(kBlock), 0, s, m,
                       (const int64_t *)A->map_diag, (const T *)A->val, (T *)A->diag);
    SHK_HIP(hipGetLastError());
    SHK_HIP(hipStreamSynchronize(s));
    return SHK_OK;
}

template <typename T>
static int upload(shk_csr A, const HostPlan &P, const T *val_host)
{
    const int64_t nd = (int64_t)P.ind.size();
    // the SpMV streams whole budget windows: pad so the last one stays in bounds
    const int64_t pad = A->budget + 16;
    SHK_HIP(hipMalloc(&A->ptr, sizeof(int) * (A->m + 1)));
    SHK_HIP(hipMalloc(&A->ind, sizeof(uint32_t) * (nd + pad)));
    SHK_HIP(hipMalloc(&A->val, sizeof(T) * (nd + pad)));
    SHK_HIP(hipMalloc(&A->diag, sizeof(T) * std::max(A->m, 1)));
    SHK_HIP(hipMalloc(&A->blk_rows, sizeof(int) * P.blk.size()));
    hipStream_t s = A->ctx->stream;
    SHK_HIP(hipMemcpyAsync(A->ptr, P.ptr.data(), sizeof(int) * (A->m + 1), hipMemcpyHostToDevice, s));
    SHK_HIP(hipMemsetAsync(A->ind, 0, sizeof(uint32_t) * (nd + pad), s));
    SHK_HIP(hipMemsetAsync(A->val, 0, sizeof(T) * (nd + pad), s));
    if (nd) SHK_HIP(hipMemcpyAsync(A->ind, P.ind.data(), sizeof(uint32_t) * nd, hipMemcpyHostToDevice, s));
    SHK_HIP(hipMemcpyAsync(A->blk_rows, P.blk.data(), sizeof(int) * P.blk.size(), hipMemcpyHostToDevice, s));
    {
        const size_t nb = P.blk.size() - 1;
        std::vector<int> info(4 * std::max<size_t>(nb, 1), 0);
        for (size_t b = 0; b < nb; b++) {
            info[4 * b + 0] = P.blk[b];
            info[4 * b + 1] = P.blk[b + 1];
            info[4 * b + 2] = P.ptr[P.blk[b]];
            info[4 * b + 3] = P.ptr[P.blk[b + 1]];
        }
        SHK_HIP(hipMalloc(&A->blk_info, sizeof(int) * info.size()));
        SHK_HIP(hipMemcpyAsync(A->blk_info, info.data(), sizeof(int) * info.size(), hipMemcpyHostToDevice, s));
        SHK_HIP(hipStreamSynchronize(s));
    }
    if (!P.mid.empty()) {
        SHK_HIP(hipMalloc(&A->mid, sizeof(int) * A->m));
        SHK_HIP(hipMemcpyAsync(A->mid, P.mid.data(), sizeof(int) * A->m, hipMemcpyHostToDevice, s));
        SHK_HIP(hipMalloc(&A->map_up, sizeof(int64_t) * std::max<int64_t>(A->nnz_stored, 1)));
        SHK_HIP(hipMalloc(&A->map_lo, sizeof(int64_t) * std::max<int64_t>(A->nnz_stored, 1)));
        if (A->nnz_stored) {
            SHK_HIP(hipMemcpyAsync(A->map_up, P.map_up.data(), sizeof(int64_t) * A->nnz_stored,
                                   hipMemcpyHostToDevice, s));
            SHK_HIP(hipMemcpyAsync(A->map_lo, P.map_lo.data(), sizeof(int64_t) * A->nnz_stored,
                                   hipMemcpyHostToDevice, s));
        }
    }
    // diag positions stay on device for value updates
    int64_t *dpos = nullptr;
    SHK_HIP(hipMalloc(&dpos, sizeof(int64_t) * std::max(A->m, 1)));
    SHK_HIP(hipMemcpyAsync(dpos, P.diag_pos.data(), sizeof(int64_t) * A->m, hipMemcpyHostToDevice, s));
    SHK_HIP(hipStreamSynchronize(s));
    A->map_diag = dpos;
    if (val_host) {
        std::vector<T> v(nd);
        if (P.src_of.empty()) std::memcpy(v.data(), val_host, sizeof(T) * nd);
        else
            for (int64_t q = 0; q < nd; q++) v[q] = val_host[P.src_of[q]];
        if (nd) SHK_HIP(hipMemcpyAsync(A->val, v.data(), sizeof(T) * nd, hipMemcpyHostToDevice, s));
        if (A->m)
            hipLaunchKernelGGL(k_gather_diag<T>, dim3(red_blocks(A->m)), dim3(kBlock), 0, s, A->m,
                               (const int64_t *)A->map_diag, (const T *)A->val, (T *)A->diag);
        SHK_HIP(hipGetLastError());
        SHK_HIP(hipStreamSynchronize(s));
    }
    return SHK_OK;
}

template <typename T>
static int set_values_dev(shk_csr A, const T *dval)
{
    hipStream_t s = A->ctx->stream;
    if (A->layout == SHK_LAYOUT_PANELS && A->storage != SHK_ROWSYMSPARSE) {
        if (A->npanel)
            hipLaunchKernelGGL(k_panel_scatter<T>, dim3(A->npanel), dim3(kBlock), 0, s, (const int *)A->pchunk,
                               (const int4 *)A->chunk, (const uint32_t *)A->ind,
                               (const unsigned short *)A->slot_row, (const int *)A->ptr, dval, (T *)A->val);
    } else if (A->storage == SHK_ROWSYMSPARSE) {
        if (A->nnz_stored)
            hipLaunchKernelGGL(k_scatter_vals<T>, dim3(red_blocks(A->nnz_stored)), dim3(kBlock), 0, s,
                               A->nnz_stored, dval, (const int64_t *)A->map_up,
                               (const int64_t *)A->map_lo, (T *)A->val);
    } else if (A->nnz_stored) {
        SHK_HIP(hipMemcpyAsync(A->val, dval, sizeof(T) * A->nnz_stored, hipMemcpyDeviceToDevice, s));
    }
    if (A->m)
        hipLaunchKernelGGL(k_gather_diag<T>, dim3(red_blocks(A->m)), dim3(kBlock), 0, s, A->m,
                           (const int64_t *)A->map_diag, (const T *)A->val, (T *)A->diag);
    SHK_HIP(hipGetLastError());
    if (A->layout == SHK_LAYOUT_SLICED) SHK_TRY(sliced_set_values(A));
    return SHK_OK;
}

template <typename T>
static int spmv_typed(shk_csr A, int mlt, T alpha, const void *dx, T beta, const void *dy_in,
                      void *dy_out)
{
    SpmvArgs<T> a{};
    a.x = (const T *)dx;
    a.y_in = (const T *)dy_in;
    a.y_out = (T *)dy_out;
    a.alpha = alpha;
    a.beta = beta;
    a.mlt = mlt;
    return launch_spmv<T>(A, a, FinNone{});
}

// The device rows as coordinates (row, col, val): RowSymSparse rows are the
// expanded ones, so this is the full symmetric matrix.  Row layout: the
// device ptr walks ind/val; panel layout: every chunk entry knows its row
// slot and window-local column.
template <typename T>
static __global__ void k_panel_coo(const int *pchunk, const int4 *chunk, const uint32_t *word,
                                   const unsigned short *slot_row, const T *val, int *row_out, int *col_out,
                                   T *val_out)
{
    const int p = blockIdx.x, r0 = p * kPanelRows;
    for (int cc = pchunk[p]; cc < pchunk[p + 1]; cc++) {
        const int4 ci = chunk[cc];
        for (int e = ci.z + (int)threadIdx.x; e < ci.w; e += kBlock) {
            const uint32_t w = word[e];
            row_out[e] = r0 + slot_row[(size_t)p * kPanelRows + ((w >> kPanRowShift) & kPanRowMask)];
            col_out[e] = chunk_c0(ci) + (int)(w & kPanColMask);
            val_out[e] = val[e];
        }
    }
}

template <typename T>
static int coo_download(shk_csr A, std::vector<int> &row, std::vector<int> &col, std::vector<T> &val)
{
    shk_ctx ctx = A->ctx;
    if (A->layout == SHK_LAYOUT_PANELS) {
        const int64_t ns = A->nslots;
        std::vector<int> r((size_t)ns), c((size_t)ns);
        std::vector<T> v((size_t)ns);
        int *drow = nullptr, *dcol = nullptr;
        T *dval = nullptr;
        hipError_t e0 = hipMalloc(&drow, sizeof(int) * (size_t)std::max<int64_t>(ns, 1));
        hipError_t e1 = hipMalloc(&dcol, sizeof(int) * (size_t)std::max<int64_t>(ns, 1));
        hipError_t e2 = hipMalloc(&dval, sizeof(T) * (size_t)std::max<int64_t>(ns, 1));
        hipError_t e3 = hipSuccess, e4 = hipSuccess, e5 = hipSuccess, e6 = hipSuccess, e7 = hipSuccess;
        if (e0 == hipSuccess && e1 == hipSuccess && e2 == hipSuccess && ns > 0) {
            // positions outside every chunk (alignment padding) keep row -1
            e3 = hipMemsetAsync(drow, 0xff, sizeof(int) * (size_t)ns, ctx->stream);
            if (A->npanel)
                hipLaunchKernelGGL(k_panel_coo<T>, dim3(A->npanel), dim3(kBlock), 0, ctx->stream,
                                   (const int *)A->pchunk, (const int4 *)A->chunk, (const uint32_t *)A->ind,
                                   (const unsigned short *)A->slot_row, (const T *)A->val, drow, dcol, dval);
            e4 = hipMemcpyAsync(r.data(), drow, sizeof(int) * (size_t)ns, hipMemcpyDeviceToHost, ctx->stream);
            e5 = hipMemcpyAsync(c.data(), dcol, sizeof(int) * (size_t)ns, hipMemcpyDeviceToHost, ctx->stream);
            e6 = hipMemcpyAsync(v.data(), dval, sizeof(T) * (size_t)ns, hipMemcpyDeviceToHost, ctx->stream);
            e7 = hipStreamSynchronize(ctx->stream);
        }
        hipFree(drow);
        hipFree(dcol);
        hipFree(dval);
        SHK_CHECK(e0 == hipSuccess && e1 == hipSuccess && e2 == hipSuccess, SHK_ERR_NOMEM,
                  "transpose: device memory");
        SHK_HIP(e3);
        SHK_HIP(e4);
        SHK_HIP(e5);
        SHK_HIP(e6);
        SHK_HIP(e7);
        for (int64_t e = 0; e < ns; e++)
            if (r[e] >= 0) {
                row.push_back(r[e]);
                col.push_back(c[e]);
                val.push_back(v[e]);
            }
        return SHK_OK;
    }
    const int64_t nd = A->nnz_dev;
    std::vector<int> ptr((size_t)A->m + 1);
    std::vector<uint32_t> ind((size_t)nd);
    val.resize((size_t)nd);
    SHK_HIP(hipMemcpyAsync(ptr.data(), A->ptr, sizeof(int) * ((size_t)A->m + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (nd > 0) {
        SHK_HIP(hipMemcpyAsync(ind.data(), A->ind, sizeof(uint32_t) * (size_t)nd, hipMemcpyDeviceToHost,
                               ctx->stream));
        SHK_HIP(hipMemcpyAsync(val.data(), A->val, sizeof(T) * (size_t)nd, hipMemcpyDeviceToHost, ctx->stream));
    }
    SHK_HIP(hipStreamSynchronize(ctx->stream));
    row.resize((size_t)nd);
    col.resize((size_t)nd);
    for (int i = 0; i < A->m; i++)
        for (int j = ptr[i]; j < ptr[(size_t)i + 1]; j++) {
            row[j] = i;
            col[j] = (int)(ind[j] & ~kLowerFlag);
        }
    return SHK_OK;
}

// CSR with `nmajor` rows from coordinates keyed (major, minor): two stable
// counting sorts, by minor then by major, so each row lists its entries in
// increasing minor index.
template <typename T>
static void coo_to_csr(int nmajor, int nminor, const std::vector<int> &major, const std::vector<int> &minor,
                       const std::vector<T> &val, std::vector<int> &ptr, std::vector<int> &ind,
                       std::vector<T> &out)
{
    const size_t nz = major.size();
    std::vector<int64_t> order(nz), tmp(nz);
    std::vector<int64_t> cnt((size_t)std::max(nmajor, nminor) + 1);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t k = 0; k < nz; k++) cnt[(size_t)minor[k] + 1]++;
    for (int i = 0; i < nminor; i++) cnt[(size_t)i + 1] += cnt[i];
    for (size_t k = 0; k < nz; k++) tmp[cnt[minor[k]]++] = (int64_t)k;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t k = 0; k < nz; k++) cnt[(size_t)major[k] + 1]++;
    for (int i = 0; i < nmajor; i++) cnt[(size_t)i + 1] += cnt[i];
    ptr.assign((size_t)nmajor + 1, 0);
    for (int i = 0; i <= nmajor; i++) ptr[i] = (int)cnt[i];
    for (size_t q = 0; q < nz; q++) {
        const int64_t k = tmp[q];
        order[cnt[major[k]]++] = k;
    }
    ind.resize(nz);
    out.resize(nz);
    for (size_t q = 0; q < nz; q++) {
        ind[q] = minor[order[q]];
        out[q] = val[order[q]];
    }
}

// The matrix behind the transposed products, built on first use (on the
// device for the row and sliced layouts, below; the panel layout's entries
// are read back and sorted on the host):
// RowSparse / ArrayRowSparse: A^T as an ArrayRowSparse matrix (every term
// "direct": alpha * a * x added onto beta * y) whose row c lists the entries
// (i, c) of A in increasing i, the order of Seldon's transposed loops
// (Functions_MatVect.cxx:753-755, Functions_MatrixArray.cxx:607-613).
// RowSymSparse: the full symmetric matrix as a RowSparse one with sorted
// columns, the order of the generic loop behind Mlt(SeldonTrans, ...)
// (Functions_MatVect.cxx:1307-1313: temp = sum over j of M(j, i) X(j)).
// coordinates of the row layout's entries (the transposed halves of a
// RowSymSparse matrix and ArrayRowSparse terms unflagged)
__global__ void k_rows_coo(int m, const int *ptr, const uint32_t *ind, int *row, int *col)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock)
        for (int j = ptr[i]; j < ptr[i + 1]; j++) {
            row[j] = i;
            col[j] = (int)(ind[j] & ~kLowerFlag);
        }
}

// The same matrix built on the device from the row layout's arrays (row and
// sliced layouts): the coordinates are sorted by (major, minor) with the
// library's stable radix sort (shk_coo_to_csr: ties keep their input order,
// as the host's two stable counting sorts do), so the result is the host
// path's entry for entry; only the layout plan is made on the host.
static int build_trans_dev(shk_csr A)
{
    shk_ctx ctx = A->ctx;
    const int64_t nd = A->nnz_dev;
    const bool sym = A->storage == SHK_ROWSYMSPARSE;
    const int nmajor = sym ? A->m : A->n, nminor = sym ? A->n : A->m;
    const size_t vs = dsize(A->dtype);
    int *row = nullptr, *col = nullptr, *tptr = nullptr, *tind = nullptr;
    void *tval = nullptr;
    auto release = [&]() {
        hipFree(row);
        hipFree(col);
        hipFree(tptr);
        hipFree(tind);
        hipFree(tval);
    };
    const size_t ne = (size_t)std::max<int64_t>(nd, 1);
    if (hipMalloc(&row, sizeof(int) * ne) != hipSuccess || hipMalloc(&col, sizeof(int) * ne) != hipSuccess ||
        hipMalloc(&tptr, sizeof(int) * ((size_t)nmajor + 1)) != hipSuccess ||
        hipMalloc(&tind, sizeof(int) * ne) != hipSuccess || hipMalloc(&tval, vs * ne) != hipSuccess) {
        release();
        SHK_CHECK(false, SHK_ERR_NOMEM, "transpose: device memory");
    }
    int st = SHK_OK;
    if (A->m > 0) {
        hipLaunchKernelGGL(k_rows_coo, dim3(red_blocks(A->m)), dim3(kBlock), 0, ctx->stream, A->m, (const int *)A->ptr,
                           (const uint32_t *)A->ind, row, col);
        if (hipGetLastError() != hipSuccess) st = SHK_ERR_HIP;
    }
    int mm = 0, nn = 0;
    if (st == SHK_OK)
        st = shk_coo_to_csr(ctx, (shk_dtype)A->dtype, nd, sym ? row : col, sym ? col : row, A->val, 0, nmajor,
                            nminor, nmajor, &mm, &nn, tptr, tind, tval);
    if (st == SHK_OK && (mm != nmajor || nn != nminor)) st = SHK_ERR_STATE;  // an entry out of range
    shk_csr T_ = nullptr;
    if (st == SHK_OK)
        st = shk_csr_create_layout(ctx, sym ? SHK_ROWSPARSE : SHK_ARRAYROWSPARSE, (shk_dtype)A->dtype, nmajor, nminor,
                                   nd, tptr, tind, tval, 1, SHK_LAYOUT_AUTO, &T_);
    release();
    SHK_CHECK(st != SHK_ERR_STATE, SHK_ERR_STATE, "transpose: entry out of range");
    SHK_TRY(st);
    A->trans = T_;
    return SHK_OK;
}

template <typename T>
static int build_trans(shk_csr A)
{
    if (A->layout != SHK_LAYOUT_PANELS) return build_trans_dev(A);
    std::vector<int> row, col;
    std::vector<T> val;
    SHK_TRY(coo_download<T>(A, row, col, val));
    for (size_t k = 0; k < row.size(); k++)
        SHK_CHECK(row[k] >= 0 && row[k] < A->m && col[k] >= 0 && col[k] < A->n, SHK_ERR_STATE,
                  "transpose: entry out of range");
    std::vector<int> ptr, ind;
    std::vector<T> out;
    const bool sym = A->storage == SHK_ROWSYMSPARSE;
    if (sym) coo_to_csr<T>(A->m, A->n, row, col, val, ptr, ind, out);
    else coo_to_csr<T>(A->n, A->m, col, row, val, ptr, ind, out);
    const int64_t nz = (int64_t)ind.size();
    shk_csr T_ = nullptr;
    SHK_TRY(shk_csr_create_layout(A->ctx, sym ? SHK_ROWSPARSE : SHK_ARRAYROWSPARSE, (shk_dtype)A->dtype,
                                  sym ? A->m : A->n, sym ? A->n : A->m, nz, ptr.data(), nz ? ind.data() : nullptr,
                                  nz ? (const void *)out.data() : nullptr, 0, SHK_LAYOUT_AUTO, &T_));
    A->trans = T_;
    return SHK_OK;
}

// The caller's stored arrays (RowSymSparse: the upper part) read back from
// the device.  Row layout: RowSparse / ArrayRowSparse entries are stored in
// the caller's order; RowSymSparse rows are [stored part ; transposed part]
// with mid[r] the split.  Panel layout (rows with sorted columns): the
// coordinates of the device entries ordered by (row, column), the transposed
// copies of a symmetric matrix left out.
template <typename T>
static int stored_download(shk_csr A, std::vector<int> &ptr, std::vector<int> &ind, std::vector<T> &val)
{
    shk_ctx ctx = A->ctx;
    const bool sym = A->storage == SHK_ROWSYMSPARSE;
    if (A->layout == SHK_LAYOUT_PANELS) {
        std::vector<int> row, col;
        std::vector<T> v;
        SHK_TRY(coo_download<T>(A, row, col, v));
        if (sym) {
            size_t q = 0;
            for (size_t k = 0; k < row.size(); k++)
                if (col[k] >= row[k]) {
                    row[q] = row[k];
                    col[q] = col[k];
                    v[q] = v[k];
                    q++;
                }
            row.resize(q);
            col.resize(q);
            v.resize(q);
        }
        coo_to_csr<T>(A->m, A->n, row, col, v, ptr, ind, val);
        return SHK_OK;
    }
    const int m = A->m;
    const int64_t nd = A->nnz_dev;
    std::vector<int> dptr((size_t)m + 1), dmid(sym ? (size_t)m : 0);
    std::vector<uint32_t> dind((size_t)nd);
    std::vector<T> dval((size_t)nd);
    SHK_HIP(hipMemcpyAsync(dptr.data(), A->ptr, sizeof(int) * ((size_t)m + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (sym && m) SHK_HIP(hipMemcpyAsync(dmid.data(), A->mid, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost, ctx->stream));
    if (nd) {
        SHK_HIP(hipMemcpyAsync(dind.data(), A->ind, sizeof(uint32_t) * (size_t)nd, hipMemcpyDeviceToHost,
                               ctx->stream));
        SHK_HIP(hipMemcpyAsync(dval.data(), A->val, sizeof(T) * (size_t)nd, hipMemcpyDeviceToHost, ctx->stream));
    }
    SHK_HIP(hipStreamSynchronize(ctx->stream));
    ptr.assign((size_t)m + 1, 0);
    ind.clear();
    val.clear();
    for (int i = 0; i < m; i++) {
        const int e = sym ? dmid[i] : dptr[(size_t)i + 1];
        for (int j = dptr[i]; j < e; j++) {
            ind.push_back((int)(dind[j] & ~kLowerFlag));
            val.push_back(dval[j]);
        }
        ptr[(size_t)i + 1] = (int)ind.size();
    }
    return SHK_OK;
}

template <typename T>
int stored_arrays(shk_csr A, std::vector<int> &ptr, std::vector<int> &ind, std::vector<T> &val)
{
    return stored_download<T>(A, ptr, ind, val);
}
template int stored_arrays<double>(shk_csr, std::vector<int> &, std::vector<int> &, std::vector<double> &);
template int stored_arrays<float>(shk_csr, std::vector<int> &, std::vector<int> &, std::vector<float> &);

static int ensure_trans(shk_csr A)
{
    if (A->trans) return SHK_OK;
    return A->dtype == SHK_F64 ? build_trans<double>(A) : build_trans<float>(A);
}

// The matrix behind A's transposed products, built on first use (the fused
// BiCg runs its A^T p~ products on it).
int transpose_of(shk_csr A, shk_csr *out)
{
    SHK_TRY(ensure_trans(A));
    *out = A->trans;
    return SHK_OK;
}

template <typename T>
static int csr_write_typed(shk_csr A, const char *path)
{
    std::vector<int> ptr, ind;
    std::vector<T> val;
    SHK_TRY(stored_download<T>(A, ptr, ind, val));
    FILE *f = std::fopen(path, "wb");
    SHK_CHECK(f, SHK_ERR_ARG, std::string("shk_csr_write: cannot open ") + path);
    const int head[3] = {A->m, A->n, (int)ind.size()};
    bool ok = std::fwrite(head, sizeof(int), 3, f) == 3;
    ok = ok && std::fwrite(ptr.data(), sizeof(int), ptr.size(), f) == ptr.size();
    if (!ind.empty()) {
        ok = ok && std::fwrite(ind.data(), sizeof(int), ind.size(), f) == ind.size();
        ok = ok && std::fwrite(val.data(), sizeof(T), val.size(), f) == val.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    SHK_CHECK(ok, SHK_ERR_ARG, std::string("shk_csr_write: write failed: ") + path);
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

static int check_csr(shk_ctx ctx, shk_csr A)
{
    SHK_CHECK(ctx && A, SHK_ERR_ARG, "null ctx or matrix");
    SHK_CHECK(A->ctx == ctx, SHK_ERR_ARG, "matrix belongs to another context");
    return SHK_OK;
}

extern "C" {

int shk_csr_create(shk_ctx ctx, shk_storage storage, shk_dtype dtype, int m, int n, int64_t nnz,
                   const int *ptr, const int *ind, const void *val, int src_on_device, shk_csr *out)
{
    return shk_csr_create_layout(ctx, storage, dtype, m, n, nnz, ptr, ind, val, src_on_device, SHK_LAYOUT_AUTO,
                                 out);
}

int shk_csr_create_layout(shk_ctx ctx, shk_storage storage, shk_dtype dtype, int m, int n, int64_t nnz,
                          const int *ptr, const int *ind, const void *val, int src_on_device, int want,
                          shk_csr *out)
{
    SHK_CHECK(ctx && out, SHK_ERR_ARG, "shk_csr_create: null argument");
    SHK_CHECK(want == SHK_LAYOUT_AUTO || want == SHK_LAYOUT_ROWS || want == SHK_LAYOUT_PANELS ||
                  want == SHK_LAYOUT_SLICED,
              SHK_ERR_ARG, "shk_csr_create: bad layout");
    SHK_CHECK(want != SHK_LAYOUT_SLICED || storage == SHK_ROWSPARSE, SHK_ERR_ARG,
              "shk_csr_create: the sliced layout is for RowSparse matrices");
    *out = nullptr;
    SHK_CHECK(m >= 0 && n >= 0 && nnz >= 0, SHK_ERR_ARG, "shk_csr_create: negative size");
    SHK_CHECK(storage >= 0 && storage <= 2, SHK_ERR_ARG, "shk_csr_create: bad storage");
    SHK_CHECK(dtype == SHK_F64 || dtype == SHK_F32, SHK_ERR_ARG, "shk_csr_create: bad dtype");
    SHK_CHECK(ptr && (nnz == 0 || (ind && val)), SHK_ERR_ARG, "shk_csr_create: null array");
    SHK_HIP(hipSetDevice(ctx->device));
    const size_t vs = dsize(dtype);
    std::vector<int> hptr(m + 1), hind(nnz);
    std::vector<char> hval(vs * (size_t)nnz);
    if (src_on_device) {
        // the arrays may come from work queued on the context's stream
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        SHK_HIP(hipMemcpy(hptr.data(), ptr, sizeof(int) * (m + 1), hipMemcpyDeviceToHost));
        if (nnz) {
            SHK_HIP(hipMemcpy(hind.data(), ind, sizeof(int) * nnz, hipMemcpyDeviceToHost));
            SHK_HIP(hipMemcpy(hval.data(), val, vs * nnz, hipMemcpyDeviceToHost));
        }
    } else {
        std::memcpy(hptr.data(), ptr, sizeof(int) * (m + 1));
        if (nnz) {
            std::memcpy(hind.data(), ind, sizeof(int) * nnz);
            std::memcpy(hval.data(), val, vs * nnz);
        }
    }
    HostPlan P;
    const int budget = dtype == SHK_F64 ? budget_of<double>() : budget_of<float>();
    const int align = dtype == SHK_F64 ? VecOf<double>::V : VecOf<float>::V;
    SHK_TRY(build_plan(storage, m, n, nnz, hptr.data(), hind.data(), budget, align, P));
    PanelPlan Q;
    int layout = SHK_LAYOUT_ROWS;
    // the sliced layout: asked for, or AUTO on a RowSparse matrix with enough
    // rows to give every CU two of its 8192-row panels' work-groups
    const bool try_sliced = want == SHK_LAYOUT_SLICED ||
                            (want == SHK_LAYOUT_AUTO && storage == SHK_ROWSPARSE && m >= kSlAutoMinRows);
    if (try_sliced) layout = SHK_LAYOUT_SLICED;  // confirmed after the plan below, else reverted
    else if (want != SHK_LAYOUT_ROWS) {
        const int win = dtype == SHK_F64 ? panel_win<double>() : panel_win<float>();
        const int pbudget = dtype == SHK_F64 ? panel_budget<double>() : panel_budget<float>();
        if (plan_panels(m, n, P, pbudget, win, align, vs, Q)) layout = SHK_LAYOUT_PANELS;
        else SHK_CHECK(want != SHK_LAYOUT_PANELS, SHK_ERR_ARG,
                       "shk_csr_create: the panel layout needs rows with sorted columns near the diagonal");
    }
    shk_csr A = new shk_csr_s();
    A->layout = layout;
    A->ctx = ctx;
    A->storage = storage;
    A->dtype = dtype;
    A->m = m;
    A->n = n;
    A->nnz_stored = nnz;
    A->nnz_dev = (int64_t)P.ind.size();
    A->budget = layout == SHK_LAYOUT_PANELS ? (dtype == SHK_F64 ? panel_budget<double>() : panel_budget<float>())
                                            : budget;
    A->max_row = P.max_row;
    int st;
    if (layout == SHK_LAYOUT_PANELS) {
        A->nblk = Q.npanel;  // one work-group per panel
        A->xcd_q = (A->nblk + kNumXcd - 1) / kNumXcd;
        A->nblk_launch = A->xcd_q * kNumXcd;
        st = dtype == SHK_F64 ? upload_panels<double>(A, P, Q, (const double *)hval.data())
                              : upload_panels<float>(A, P, Q, (const float *)hval.data());
    } else {
        A->nblk = (int)P.blk.size() - 1;
        A->xcd_q = (A->nblk + kNumXcd - 1) / kNumXcd;
        A->nblk_launch = A->xcd_q * kNumXcd;
        A->nslots = A->nnz_dev;
        st = dtype == SHK_F64 ? upload<double>(A, P, (const double *)hval.data())
                              : upload<float>(A, P, (const float *)hval.data());
    }
    if (st == SHK_OK && layout == SHK_LAYOUT_SLICED) {
        bool applies = false;
        st = sliced_build(A, hptr.data(), hind.data(), &applies);
        if (st == SHK_OK && applies) {
            st = sliced_set_values(A);
            if (st == SHK_OK) st = hipStreamSynchronize(ctx->stream) == hipSuccess ? SHK_OK : SHK_ERR_HIP;
        } else if (st == SHK_OK && want == SHK_LAYOUT_SLICED) {
            st = fail(SHK_ERR_ARG, "shk_csr_create: the sliced layout needs rows with sorted columns near "
                                   "the diagonal");
        } else if (st == SHK_OK) {
            // AUTO: the plain row layout already uploaded serves as is
            sliced_free(A);
            A->layout = SHK_LAYOUT_ROWS;
        }
    }
    if (st != SHK_OK) {
        shk_csr_destroy(A);
        return st;
    }
    *out = A;
    return SHK_OK;
}

int shk_csr_destroy(shk_csr A)
{
    if (!A) return SHK_OK;
    if (A->trans) shk_csr_destroy(A->trans);
    if (A->ctx) hipStreamSynchronize(A->ctx->stream);
    sor_free(A->sor);
    hipFree(A->ptr);
    hipFree(A->mid);
    hipFree(A->ind);
    hipFree(A->val);
    hipFree(A->diag);
    hipFree(A->blk_rows);
    hipFree(A->blk_info);
    hipFree(A->map_up);
    hipFree(A->map_lo);
    hipFree(A->map_diag);
    hipFree(A->pchunk);
    hipFree(A->chunk);
    hipFree(A->slot_row);
    sliced_free(A);
    delete A;
    return SHK_OK;
}

int shk_csr_set_values(shk_ctx ctx, shk_csr A, const void *val, int src_on_device)
{
    SHK_TRY(check_csr(ctx, A));
    const size_t vs = dsize(A->dtype);
    if (A->trans) {  // rebuilt from the new values when next needed
        shk_csr_destroy(A->trans);
        A->trans = nullptr;
    }
    if (A->sor) {
        hipStreamSynchronize(ctx->stream);
        sor_free(A->sor);
        A->sor = nullptr;
    }
    if (A->nnz_stored == 0) return SHK_OK;
    const void *dsrc = val;
    void *tmp = nullptr;
    if (!src_on_device) {
        SHK_HIP(hipMalloc(&tmp, vs * A->nnz_stored));
        SHK_HIP(hipMemcpyAsync(tmp, val, vs * A->nnz_stored, hipMemcpyHostToDevice, ctx->stream));
        dsrc = tmp;
    }
    int st = A->dtype == SHK_F64 ? set_values_dev<double>(A, (const double *)dsrc)
                                 : set_values_dev<float>(A, (const float *)dsrc);
    if (tmp) {
        hipStreamSynchronize(ctx->stream);
        hipFree(tmp);
    }
    return st;
}

int shk_csr_layout(shk_csr A, int *layout)
{
    SHK_CHECK(A && layout, SHK_ERR_ARG, "null argument");
    *layout = A->layout;
    return SHK_OK;
}

int shk_csr_write(shk_ctx ctx, shk_csr A, const char *path)
{
    SHK_TRY(check_csr(ctx, A));
    SHK_CHECK(path, SHK_ERR_ARG, "shk_csr_write: null path");
    return A->dtype == SHK_F64 ? csr_write_typed<double>(A, path) : csr_write_typed<float>(A, path);
}

int shk_csr_read(shk_ctx ctx, const char *path, shk_storage storage, shk_dtype dtype, int layout, shk_csr *out)
{
    SHK_CHECK(ctx && path && out, SHK_ERR_ARG, "shk_csr_read: null argument");
    SHK_CHECK(dtype == SHK_F64 || dtype == SHK_F32, SHK_ERR_ARG, "shk_csr_read: bad dtype");
    FILE *f = std::fopen(path, "rb");
    SHK_CHECK(f, SHK_ERR_ARG, std::string("shk_csr_read: cannot open ") + path);
    int head[3] = {0, 0, 0};
    bool ok = std::fread(head, sizeof(int), 3, f) == 3 && head[0] >= 0 && head[1] >= 0 && head[2] >= 0;
    const size_t vs = dsize(dtype);
    std::vector<int> ptr(ok ? (size_t)head[0] + 1 : 0), ind(ok ? (size_t)head[2] : 0);
    std::vector<char> val(ok ? vs * (size_t)head[2] : 0);
    ok = ok && std::fread(ptr.data(), sizeof(int), ptr.size(), f) == ptr.size();
    ok = ok && std::fread(ind.data(), sizeof(int), ind.size(), f) == ind.size();
    ok = ok && std::fread(val.data(), vs, (size_t)head[2], f) == (size_t)head[2];
    std::fclose(f);
    SHK_CHECK(ok, SHK_ERR_ARG, std::string("shk_csr_read: not a Seldon binary CSR file of this value type: ") + path);
    SHK_CHECK(ptr[0] == 0 && ptr[(size_t)head[0]] == head[2], SHK_ERR_ARG, "shk_csr_read: inconsistent row pointer");
    return shk_csr_create_layout(ctx, storage, dtype, head[0], head[1], head[2], ptr.data(),
                                 ind.empty() ? nullptr : ind.data(), val.empty() ? nullptr : val.data(), 0, layout,
                                 out);
}

int shk_csr_info(shk_csr A, int *m, int *n, int64_t *nnz_stored, int64_t *nnz_dev, int *dtype,
                 int *storage)
{
    SHK_CHECK(A, SHK_ERR_ARG, "null matrix");
    if (m) *m = A->m;
    if (n) *n = A->n;
    if (nnz_stored) *nnz_stored = A->nnz_stored;
    if (nnz_dev) *nnz_dev = A->nnz_dev;
    if (dtype) *dtype = A->dtype;
    if (storage) *storage = A->storage;
    return SHK_OK;
}

double shk_csr_spmv_bytes(shk_csr A, int beta_nonzero)
{
    if (!A) return 0.0;
    const double V = (double)dsize(A->dtype);
    return (4.0 + V) * (double)A->nnz_stored + 4.0 * (A->m + 1) + V * A->n + V * A->m +
           (beta_nonzero ? V * A->m : 0.0);
}

int shk_spmv(shk_ctx ctx, shk_csr A, double alpha, const void *dx, double beta, void *dy)
{
    return shk_spmv_out(ctx, A, alpha, dx, beta, dy, dy);
}

int shk_spmv_out(shk_ctx ctx, shk_csr A, double alpha, const void *dx, double beta,
                 const void *dy_in, void *dy_out)
{
    SHK_TRY(check_csr(ctx, A));
    if (A->m == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy_in, "y_in"));
    SHK_TRY(check_device_ptr(dy_out, "y_out"));
    if (A->dtype == SHK_F64) return spmv_typed<double>(A, 0, alpha, dx, beta, dy_in, dy_out);
    return spmv_typed<float>(A, 0, (float)alpha, dx, (float)beta, dy_in, dy_out);
}

int shk_spmv_trans(shk_ctx ctx, shk_csr A, double alpha, const void *dx, double beta, void *dy)
{
    SHK_TRY(check_csr(ctx, A));
    if (A->storage == SHK_ROWSYMSPARSE) return shk_spmv(ctx, A, alpha, dx, beta, dy);
    if (A->n == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy, "y"));
    SHK_TRY(ensure_trans(A));
    if (A->storage == SHK_ROWSPARSE && beta == 0.0) {
        // RowSparse starts from Mlt(beta, Y), a product even for beta = 0
        // (Inf and NaN stay NaN); A^T's ArrayRowSparse form would zero-fill
        SHK_TRY(shk_scal(ctx, (shk_dtype)A->dtype, A->n, 0.0, dy));
        beta = 1.0;
    }
    return shk_spmv(ctx, A->trans, alpha, dx, beta, dy);
}

int shk_mlt_trans(shk_ctx ctx, shk_csr A, const void *dx, void *dy)
{
    SHK_TRY(check_csr(ctx, A));
    if (A->n == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy, "y"));
    SHK_TRY(ensure_trans(A));
    return shk_mlt(ctx, A->trans, dx, dy);
}

int shk_mlt(shk_ctx ctx, shk_csr A, const void *dx, void *dy)
{
    SHK_TRY(check_csr(ctx, A));
    if (A->m == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy, "y"));
    if (A->dtype == SHK_F64) return spmv_typed<double>(A, 1, 1.0, dx, 0.0, nullptr, dy);
    return spmv_typed<float>(A, 1, 1.0f, dx, 0.0f, nullptr, dy);
}

int shk_mlt_dot(shk_ctx ctx, shk_csr A, const void *dx, void *dy, double *result)
{
    SHK_TRY(check_csr(ctx, A));
    SHK_CHECK(result, SHK_ERR_ARG, "null result");
    if (A->m == 0) { *result = 0.0; return SHK_OK; }
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy, "y"));
    if (A->dtype == SHK_F64) {
        SpmvArgs<double> a{};
        a.x = (const double *)dx; a.y_out = (double *)dy; a.alpha = 1.0; a.mlt = 1;
        a.dot0 = kDotVec; a.u0 = (const double *)dx;
        SHK_TRY(launch_spmv<double>(A, a, FinStore<double>{(double *)ctx->scalars}));
        double h[2];
        SHK_HIP(hipMemcpyAsync(h, ctx->scalars, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        *result = h[0];
    } else {
        SpmvArgs<float> a{};
        a.x = (const float *)dx; a.y_out = (float *)dy; a.alpha = 1.0f; a.mlt = 1;
        a.dot0 = kDotVec; a.u0 = (const float *)dx;
        SHK_TRY(launch_spmv<float>(A, a, FinStore<float>{(float *)ctx->scalars}));
        float h[2];
        SHK_HIP(hipMemcpyAsync(h, ctx->scalars, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        *result = h[0];
    }
    return SHK_OK;
}

int shk_jacobi(shk_ctx ctx, shk_csr A, const void *dr, void *dz)
{
    SHK_TRY(check_csr(ctx, A));
    if (A->m == 0) return SHK_OK;
    if (A->dtype == SHK_F64)
        hipLaunchKernelGGL(k_jacobi<double>, dim3(red_blocks(A->m)), dim3(kBlock), 0, ctx->stream,
                           (int64_t)A->m, (const double *)dr, (const double *)A->diag, (double *)dz);
    else
        hipLaunchKernelGGL(k_jacobi<float>, dim3(red_blocks(A->m)), dim3(kBlock), 0, ctx->stream,
                           (int64_t)A->m, (const float *)dr, (const float *)A->diag, (float *)dz);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

int shk_time_spmv(shk_ctx ctx, shk_csr A, int mode, const void *dx, void *dy, int reps, double *avg_ms)
{
    SHK_TRY(check_csr(ctx, A));
    SHK_CHECK(reps > 0 && avg_ms, SHK_ERR_ARG, "shk_time_spmv: bad reps/out");
    SHK_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < reps; i++) {
        int st = mode == 0 ? shk_mlt(ctx, A, dx, dy) : shk_spmv(ctx, A, 1.0, dx, 1.0, dy);
        if (st != SHK_OK) return st;
    }
    SHK_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    SHK_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    SHK_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *avg_ms = ms / reps;
    return SHK_OK;
}

}  // extern "C"

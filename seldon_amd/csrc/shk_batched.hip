// shk_batched.hip -- many independent SPD systems in one persistent kernel.
//
// The systems are stacked block-diagonally in one CSR (system s owns rows
// [row_off[s], row_off[s+1]) and its columns point into the same range).  A
// persistent grid of work-groups pulls systems from a device queue, largest
// first.  A system is solved by a team of k work-groups (k = 1 for most):
// member j owns a contiguous share of the system's rows, the team's dots are
// each member's work-group sum added in member order (the same value in
// every member), and the members meet at a counter in device memory where
// the recurrences need another member's rows (p before the product).  k is
// planned on the host from the system's share of the batch's estimated work
// (rows^1.5: Cg's iteration count grows like the square root of the
// condition number, the grid side for the Laplacian-like systems this is
// for), so a few large systems no longer hold the whole batch up on one CU.
// Slots (system, member) are queued largest share first, the k slots of a
// system next to each other; only the last system taken can be short of
// members, and the members it waits for are freed by systems that have all
// of theirs, so every wait ends.  A wait that outlives a wall-clock limit
// (a GPU shared with other work) sets a flag the host reports as an error.
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

constexpr int kMaxTeam = 16;
constexpr long long kTeamWaitSeconds = 20;  // a member's wait at a meeting, at most

// One system's team.  `cnt` counts arrivals at the system's meeting points
// (it only grows: meeting n ends when it reaches (n+1) k); `slot` holds the
// members' partial sums of the last two reductions (a member can be at most
// one meeting ahead of another, so two halves suffice).
struct Team {
    int k, j;
    unsigned *cnt;
    double *slot;         // [2][kMaxTeam]
    unsigned *stall;      // set by a wait that timed out; every wait then ends
    long long limit;      // wall-clock ticks a wait may take
    unsigned n;           // meetings passed
};

// All members' writes before the call are visible to every member after it.
__device__ __forceinline__ void team_wait(Team &t)
{
    // every thread's stores have reached L2 before thread 0's device-scope
    // release writes the L2 back (one release per work-group, not per wave)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release this work-group's writes
        atomicAdd(t.cnt, 1u);
        const unsigned target = (t.n + 1u) * (unsigned)t.k;
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(t.cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(t.stall, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (wall_clock64() - t0 > t.limit) {
                atomicExch(t.stall, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();  // acquire the other members' writes
    }
    __syncthreads();
    t.n++;
}

__device__ __forceinline__ void team_barrier(Team &t)
{
    if (t.k == 1) __syncthreads();
    else team_wait(t);
}

// Team all-reduce: the members' work-group sums added in member order.
template <typename T>
__device__ __forceinline__ T team_allreduce(T v, T *buf, int &phase, Team &t)
{
    const T g = group_allreduce(v, buf, phase);
    if (t.k == 1) return g;
    const double *sl = t.slot + (t.n & 1u) * kMaxTeam;
    if (threadIdx.x == 0) t.slot[(t.n & 1u) * kMaxTeam + t.j] = (double)g;  // exact for float too
    team_wait(t);
    T tot = (T)sl[0];
    for (int i = 1; i < t.k; i++) tot += (T)sl[i];
    return tot;
}

template <typename T>
__device__ __forceinline__ T sqrt_b(T v) { return sqrt(v); }
template <>
__device__ __forceinline__ float sqrt_b<float>(float v) { return sqrtf(v); }

// The stacked matrix as the kernels read it: rows in slices of 64 (one per
// wavefront row step), a slice's entries stored entry-major -- entry k of its
// 64 rows at 64 consecutive positions -- so the k-th loads of a wavefront's
// rows are coalesced (the caller's CSR order has each thread's row at a
// different 40-byte stride, 20 cache lines per load of a 5-point row).  Row i's
// k-th entry is at boff[i / 64] + 64 k + i % 64; slots past a row's length are
// never read.  Built on the device per call from the row layout.
struct SellView {
    const int *ptr;        // row lengths: ptr[i+1] - ptr[i]
    const int *boff;       // [slices] first position of each slice
    const uint32_t *col;
    const void *val;
};

// Row sums of U rows of a thread: step k adds entry k of each row that has
// one, so each row's additions stay in stored order
// (Functions_MatVect.cxx:424-430).  The rows are visited as in a plain
// grid-stride loop (thread tid takes rows i0, i0 + stride, ...), U of them at
// a time: their loads are all issued before any is used, so a thread keeps U
// rows' memory accesses in flight instead of one.
template <typename T, typename P, int U>
__device__ __forceinline__ void row_sums(const SellView &M, const P *p, int i0, int stride, int r1, T (&temp)[U])
{
    const T *val = (const T *)M.val;
    int base[U], len[U];
    int L = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int i = i0 + u * stride;
        const bool in = i < r1;
        len[u] = in ? M.ptr[i + 1] - M.ptr[i] : 0;
        base[u] = in ? M.boff[i >> 6] + (i & 63) : 0;
        temp[u] = (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) L = max(L, len[u]);
    for (int k = 0; k < L; k++) {
        T av[U], pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = k < len[u] ? base[u] + 64 * k : 0;
            av[u] = val[j];
            pv[u] = (T)p[M.col[j]];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k < len[u]) temp[u] += av[u] * pv[u];
    }
}

// slice widths: the longest row of each 64-row slice
__global__ void k_sell_width(int n, const int *ptr, int *width)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b * 64 >= n) return;
    const int e = min(n, b * 64 + 64);
    int w = 0;
    for (int i = b * 64; i < e; i++) w = max(w, ptr[i + 1] - ptr[i]);
    width[b] = w;
}

// entries (and, when given, the FP32 copy's values) into the slices; every
// column must lie in its row's own system (a column in another system would
// read that system's p while it is being written): *bad is set otherwise
template <typename T>
__global__ void k_sell_fill(int n, const int *ptr, const uint32_t *ind, const T *val, const float *val32,
                            const int *boff, const int *row_off, int count, uint32_t *col, T *sval,
                            float *sval32, int *bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the system of row i: the last s with row_off[s] <= i
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t c0 = (uint32_t)row_off[lo], c1 = (uint32_t)row_off[lo + 1];
    const int j0 = ptr[i], len = ptr[i + 1] - j0;
    const int q = boff[i >> 6] + (i & 63);
    bool out = false;
    for (int k = 0; k < len; k++) {
        const uint32_t c = ind[j0 + k];
        out |= c < c0 || c >= c1;
        col[q + 64 * k] = c;
        sval[q + 64 * k] = val[j0 + k];
        if (val32) sval32[q + 64 * k] = val32[j0 + k];
    }
    if (out) atomicOr(bad, 1);
}

// One Cg (identity): Seldon's recurrences, this work-group's rows [r0, r1)
// of the system (all of them when the team has one member).  Returns the
// iteration count; *code gets ErrorCode().  facteur is 1/|b| (Init), already
// checked nonzero.
template <typename T, int U>
__device__ int group_cg(int r0, int r1, const SellView &M, const T *b, T *x,
                        T *r, T *p, T *q, double facteur, double tol, int max_iter, T *buf, int &phase,
                        Team &team, int *code)
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
    T rr = team_allreduce(part, buf, phase, team);
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
        team_barrier(team);  // the product reads p of the other members' rows
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
        const T delta = team_allreduce(part, buf, phase, team);
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
        // every member has finished its product here, so p may be rewritten
        rr = team_allreduce(part, buf, phase, team);
        rho_1 = rho;
        nb++;
        __syncthreads();  // p is rewritten next iteration while others may still read it
    }
    if (nb >= max_iter) *code = -2;
    return nb;
}

// Waves per SIMD the kernels are compiled for: two 1024-thread work-groups
// per CU with one or two rows per pass (registers then capped at 64 VGPRs and
// the SGPRs that eight waves leave), one with four or eight.
template <int U>
constexpr int kBatchWaves = U <= 2 ? 8 : 4;

struct BatchArgs {
    int count;
    int nslot;           // queue length: sum of the team sizes
    const int2 *slots;   // queue position -> {system, member}
    const int *team;     // [count] team size of each system
    unsigned *team_cnt;  // [count] meeting counters
    double *team_slot;   // [count][2][kMaxTeam] members' partial sums
    unsigned *stall;     // a wait timed out
    long long limit;     // wall-clock ticks a wait may take
    const int *row_off;  // [count+1]
    SellView A64, A32;
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

// queue head, stall flag and meeting counters to zero (a kernel, so a
// captured graph holds no memset node)
__global__ void k_batch_reset(unsigned *queue, unsigned *stall, unsigned *team_cnt, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *queue = 0u;
        *stall = 0u;
    }
    if (i < count) team_cnt[i] = 0u;
}

// The queue slot a work-group takes next: its system, its rows [m0, m1)
// and its team.
struct Taken {
    int s, r0, r1, m0, m1;
    Team team;
};

__device__ __forceinline__ bool take_slot(const BatchArgs &a, int *s_q, Taken &tk)
{
    if (threadIdx.x == 0) *s_q = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int qi = *s_q;
    __syncthreads();
    if (qi >= a.nslot) return false;
    const int2 sl = a.slots[qi];
    tk.s = sl.x;
    tk.r0 = a.row_off[tk.s];
    tk.r1 = a.row_off[tk.s + 1];
    const int k = a.team[tk.s];
    const int share = (tk.r1 - tk.r0 + k - 1) / k;
    tk.m0 = min(tk.r1, tk.r0 + sl.y * share);
    tk.m1 = min(tk.r1, tk.m0 + share);
    tk.team.k = k;
    tk.team.j = sl.y;
    tk.team.cnt = a.team_cnt + tk.s;
    tk.team.slot = a.team_slot + (size_t)tk.s * 2 * kMaxTeam;
    tk.team.stall = a.stall;
    tk.team.limit = a.limit;
    tk.team.n = 0u;
    return true;
}

template <int U>
__global__ __launch_bounds__(kMaxGroup, kBatchWaves<U>) void k_cg_batched(BatchArgs a)
{
    __shared__ double buf[32];
    __shared__ int s_q;
    int phase = 0;
    double *r = (double *)a.w64, *p = r + a.n, *q = p + a.n;
    Taken tk;
    while (take_slot(a, &s_q, tk)) {
        // Init(b)
        double part = 0;
        for (int i = tk.m0 + threadIdx.x; i < tk.m1; i += blockDim.x) part += a.b[i] * a.b[i];
        const double bn = sqrt(team_allreduce(part, buf, phase, tk.team));
        int code = 0, its = 0;
        if (bn != 0.0)
            its = group_cg<double, U>(tk.m0, tk.m1, a.A64, a.b, a.x, r, p, q, 1.0 / bn,
                                      a.tol, a.max_iter, buf, phase, tk.team, &code);
        if (threadIdx.x == 0 && tk.team.j == 0) {
            a.its[tk.s] = its;
            a.codes[tk.s] = code;
        }
    }
}

// Mixed: FP64 refinement around FP32 inner Cg, per system.
template <int U>
__global__ __launch_bounds__(kMaxGroup, kBatchWaves<U>) void k_cg_batched_ir(BatchArgs a)
{
    __shared__ double buf[32];
    __shared__ float buf32[32];
    __shared__ int s_q;
    int phase = 0, phase32 = 0;
    double *r64 = (double *)a.w64;
    float *r32 = a.w32, *p32 = r32 + a.n, *q32 = p32 + a.n, *d32 = q32 + a.n;
    Taken tk;
    while (take_slot(a, &s_q, tk)) {
        const int s = tk.s, r0 = tk.m0, r1 = tk.m1;  // this member's rows
        double part = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            part += a.b[i] * a.b[i];
            a.x[i] = 0.0;
        }
        const double bn = sqrt(team_allreduce(part, buf, phase, tk.team));  // also: x = 0 everywhere
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
                const double rn = sqrt(team_allreduce(part, buf, phase, tk.team));
                if (rn / bn < a.tol) break;
                if (k >= a.max_refine) { code = -2; break; }
                const double inv = 1.0 / rn;
                float part32 = 0.0f;
                for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
                    const float v = (float)(r64[i] * inv);
                    r32[i] = v;
                    part32 += v * v;
                }
                const float rh = sqrt_b(team_allreduce(part32, buf32, phase32, tk.team));
                int icode = 0;
                // right-hand side and residual share r32: Cg copies b into r
                // first (r = b) and reads b nowhere else
                total += group_cg<float, U>(r0, r1, a.A32, r32, d32, r32, p32, q32,
                                            1.0 / (double)rh, a.inner_tol, a.max_iter, buf32, phase32, tk.team,
                                            &icode);
                for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) a.x[i] += rn * (double)d32[i];
                team_barrier(tk.team);  // the next residual reads x of every member's rows
                if (icode > 0) { code = icode; break; }
            }
        }
        if (threadIdx.x == 0 && tk.team.j == 0) {
            a.its[s] = total;
            a.codes[s] = code;
        }
    }
}

// Team sizes and the queue.  A system's estimated work is rows^1.5; it gets
// k = ceil(work / (the batch's work per resident work-group)) members,
// 1 <= k <= kMaxTeam (and at most its rows).  Slots are queued by their share
// of work, largest first, a system's k slots next to each other.
static void plan_teams(const std::vector<int> &sizes, int capacity, std::vector<int> &team,
                       std::vector<int2> &slots)
{
    const int count = (int)sizes.size();
    std::vector<double> work(count);
    double total = 0.0;
    for (int s = 0; s < count; s++) {
        work[s] = std::pow((double)sizes[s], 1.5);
        total += work[s];
    }
    const double per_group = total / capacity;
    team.assign(count, 1);
    for (int s = 0; s < count; s++) {
        int k = per_group > 0.0 ? (int)std::ceil(work[s] / per_group - 1e-9) : 1;
        k = std::max(1, std::min({k, kMaxTeam, capacity, std::max(sizes[s], 1)}));
        team[s] = k;
    }
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int i, int j) { return work[i] / team[i] > work[j] / team[j]; });
    slots.clear();
    for (int s : order)
        for (int j = 0; j < team[s]; j++) slots.push_back(make_int2(s, j));
}

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
    // work-groups drawing on the queue at a time
    const void *fn = mixed ? (rpp == 1   ? (const void *)k_cg_batched_ir<1>
                              : rpp == 2 ? (const void *)k_cg_batched_ir<2>
                              : rpp == 4 ? (const void *)k_cg_batched_ir<4>
                                         : (const void *)k_cg_batched_ir<8>)
                           : (rpp == 1   ? (const void *)k_cg_batched<1>
                              : rpp == 2 ? (const void *)k_cg_batched<2>
                              : rpp == 4 ? (const void *)k_cg_batched<4>
                                         : (const void *)k_cg_batched<8>);
    // The plan depends on the work-group size only (2048 threads per CU), not
    // on the kernel's register use: the teams, and with them the order of
    // every sum, are the same for every rows-per-pass setting.
    const int capacity = std::max(1, ctx->num_cus * std::max(1, 2048 / group));
    std::vector<int> team;
    std::vector<int2> slots;
    plan_teams(sizes, capacity, team, slots);
    const int nslot = (int)slots.size();
    const int grid = std::min(nslot, capacity);
    const int n = A64->m;
    const size_t off_bytes = sizeof(int) * (size_t)(count + 1), ord_bytes = sizeof(int) * (size_t)count;
    const size_t slot_bytes = sizeof(int2) * (size_t)nslot;
    const size_t tslot_bytes = sizeof(double) * 2 * kMaxTeam * (size_t)count;
    const size_t w64_bytes = sizeof(double) * (size_t)n * (mixed ? 1 : 3);
    const size_t w32_bytes = mixed ? sizeof(float) * (size_t)n * 4 : 0;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    hipStream_t st = ctx->stream;
    // the sliced copy of the matrix: slice widths on the device, their
    // offsets on the host (a few hundred KB each way), then the entries
    const int nsl = (n + 63) / 64;
    const size_t sl_bytes = sizeof(int) * (size_t)std::max(nsl, 1);
    SHK_TRY(ensure_ws2(ctx, sl_bytes));
    std::vector<int> boff(std::max(nsl, 1), 0);
    if (nsl > 0) {
        hipLaunchKernelGGL(k_sell_width, dim3((nsl + 255) / 256), dim3(256), 0, st, n, (const int *)A64->ptr,
                           (int *)ctx->ws2);
        SHK_HIP(hipGetLastError());
        SHK_HIP(hipMemcpyAsync(boff.data(), ctx->ws2, sizeof(int) * nsl, hipMemcpyDeviceToHost, st));
        SHK_HIP(hipStreamSynchronize(st));
    }
    int64_t padded = 0;
    for (int b = 0; b < nsl; b++) {
        const int wdt = boff[b];
        boff[b] = (int)padded;
        padded += 64 * (int64_t)wdt;
    }
    SHK_CHECK(padded < INT32_MAX, SHK_ERR_ARG, "cg_batched: the stacked matrix is too large for 32-bit positions");
    const size_t pad = (size_t)std::max<int64_t>(padded, 1);
    const size_t col_bytes = sizeof(uint32_t) * pad, sv_bytes = sizeof(double) * pad,
                 sv32_bytes = mixed ? sizeof(float) * pad : 0;
    const size_t need = al(off_bytes) + al(slot_bytes) + al(w64_bytes) + al(w32_bytes) + 4 * al(ord_bytes) +
                        al(tslot_bytes) + al(sl_bytes) + al(col_bytes) + al(sv_bytes) + al(sv32_bytes) + 256;
    SHK_TRY(ensure_ws2(ctx, need));
    char *w = (char *)ctx->ws2;
    int *d_boff = (int *)w;
    w += al(sl_bytes);
    uint32_t *d_scol = (uint32_t *)w;
    w += al(col_bytes);
    double *d_sval = (double *)w;
    w += al(sv_bytes);
    float *d_sval32 = (float *)w;
    w += al(sv32_bytes);
    int *d_off = (int *)w;
    w += al(off_bytes);
    int2 *d_slots = (int2 *)w;
    w += al(slot_bytes);
    double *d_w64 = (double *)w;
    w += al(w64_bytes);
    float *d_w32 = (float *)w;
    w += al(w32_bytes);
    int *d_its = (int *)w;
    w += al(ord_bytes);
    int *d_codes = (int *)w;
    w += al(ord_bytes);
    int *d_team = (int *)w;
    w += al(ord_bytes);
    unsigned *d_tcnt = (unsigned *)w;
    w += al(ord_bytes);
    double *d_tslot = (double *)w;
    w += al(tslot_bytes);
    unsigned *d_stall = (unsigned *)w;
    int *d_bad = (int *)ctx->scalars;  // one int
    SHK_HIP(hipMemcpyAsync(d_boff, boff.data(), sizeof(int) * nsl, hipMemcpyHostToDevice, st));
    SHK_HIP(hipMemcpyAsync(d_off, row_off_host, off_bytes, hipMemcpyHostToDevice, st));
    SHK_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), st));
    if (n > 0)
        hipLaunchKernelGGL(k_sell_fill<double>, dim3((n + 255) / 256), dim3(256), 0, st, n, (const int *)A64->ptr,
                           (const uint32_t *)A64->ind, (const double *)A64->val,
                           mixed ? (const float *)A32->val : (const float *)nullptr, (const int *)d_boff,
                           (const int *)d_off, count, d_scol, d_sval, mixed ? d_sval32 : (float *)nullptr, d_bad);
    SHK_HIP(hipGetLastError());
    int bad = 0;
    SHK_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st));
    SHK_HIP(hipStreamSynchronize(st));
    SHK_CHECK(bad == 0, SHK_ERR_ARG, "cg_batched: a system has a column outside its own rows");
    SHK_HIP(hipMemcpyAsync(d_slots, slots.data(), slot_bytes, hipMemcpyHostToDevice, st));
    SHK_HIP(hipMemcpyAsync(d_team, team.data(), ord_bytes, hipMemcpyHostToDevice, st));
    int clock_khz = 0;
    SHK_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
    BatchArgs a{};
    a.count = count;
    a.nslot = nslot;
    a.slots = d_slots;
    a.team = d_team;
    a.team_cnt = d_tcnt;
    a.team_slot = d_tslot;
    a.stall = d_stall;
    // wall_clock64 ticks per second (the device reports kHz; 100 MHz when it
    // reports none)
    a.limit = (long long)(clock_khz > 0 ? clock_khz : 100000) * 1000LL * kTeamWaitSeconds;
    a.row_off = d_off;
    a.A64 = SellView{(const int *)A64->ptr, d_boff, d_scol, d_sval};
    if (mixed) a.A32 = SellView{(const int *)A64->ptr, d_boff, d_scol, d_sval32};
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
    auto enqueue = [&]() -> int {
        hipLaunchKernelGGL(k_batch_reset, dim3((count + 255) / 256), dim3(256), 0, st, a.queue, a.stall,
                           a.team_cnt, count);
        void *args[] = {&a};
        SHK_HIP(hipLaunchKernel(fn, dim3(grid), dim3(group), args, 0, st));
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
    unsigned stall = 0;
    SHK_HIP(hipMemcpyAsync(its, d_its, ord_bytes, hipMemcpyDeviceToHost, st));
    SHK_HIP(hipMemcpyAsync(codes, d_codes, ord_bytes, hipMemcpyDeviceToHost, st));
    SHK_HIP(hipMemcpyAsync(&stall, d_stall, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    SHK_HIP(hipStreamSynchronize(st));
    SHK_CHECK(stall == 0u, SHK_ERR_STATE,
              "cg_batched: a team of work-groups waited too long for its members (is the GPU shared?)");
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

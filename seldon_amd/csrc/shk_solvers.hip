// shk_solvers.hip -- Cg, BiCgStab and Gmres(m) driven from the device.
//
// Each solver is the reference template's loop body unrolled into fused
// kernels; the scalar recurrences and Iteration semantics run in the
// finalize steps of shk_solver_state.hpp.  The host enqueues iterations in
// chunks and reads the state back once per chunk; kernels queued past the end
// of the solve see `skip` and return at once.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "shk_solver_state.hpp"

namespace shk {

// ------------------------------------------------- allocation-free Cg ----
// shk_cg_noalloc keeps the solver's work vectors, its device residual
// history and its reduction partials in this array instead of the context's
// grow-only buffers, so the call makes no device allocation.  Every device
// the library runs on holds its own copy until the process ends; calls that
// overlap on one device race on it (one at a time per device).
constexpr size_t kNoAllocBytes = (size_t)32 << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) char scratch[kNoAllocBytes];
// device address of `scratch` per device (hipGetSymbolAddress with that
// device current); host code uses these, never the array's name
static void *g_scratch[kNoAllocDevices];
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kNoAllocDevices)
        return fail(SHK_ERR_ARG, "no-alloc scratch: device ordinal out of range");
    if (g_scratch[device]) return SHK_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SHK_ERR_HIP, std::string("no-alloc scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return SHK_OK;
}

static void *noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kNoAllocDevices) ? g_scratch[device] : nullptr;
}

// The vector arguments of these kernels are distinct buffers (__restrict__),
// so an unrolled loop keeps several elements' loads in flight; each thread
// still accumulates its elements in increasing index order.
// A kernel that folds the previous kernel's finish (every work-group forms
// the total of its partials before it can go on) first loads its first kPre
// grid-stride elements of each vector into registers, so those loads are in
// flight while the partials are fetched and summed; the rest follow the
// fold.  Each thread still visits its elements in GRID_STRIDE's order, so its
// sums are the same.
constexpr int kPre = 8;

// elements per thread of the solvers' grid-stride kernels, and the fewest
// work-groups a system gets while it has two elements per thread to give:
// together they fix the grid, and so the shape of every reduction, per n
#ifndef SHK_SOLVE_ITEMS
#define SHK_SOLVE_ITEMS 8
#endif
constexpr int kSolveMinBlocks = 256;

#define GRID_STRIDE(i, n) \
    _Pragma("unroll 4") for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (n); \
                             i += (int64_t)gridDim.x * kBlock)

// ---------------------------------------------------------------- Cg ----
// Init: x = 0 (x.Zero(), Cg.cxx:67) and r = b (Copy, :63); sum r^2.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init0(int64_t n, const T *__restrict__ b, T *__restrict__ x, T *__restrict__ r,
                                                     KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T v = b[i];
        x[i] = (T)0;
        r[i] = v;
        c[0] += v * v;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// z = M^{-1} r (Jacobi), sum r*z  (Cg.cxx:75-78)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_z(int64_t n, const T *__restrict__ r, const T *__restrict__ d, T *__restrict__ z,
                                                 KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T rv = r[i];
        T zv = rv / d[i];
        z[i] = zv;
        c[0] += rv * zv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// Jacobi Cg: the loop head of the previous iteration (rho_1 = rho, ++iter,
// Finished(r)) folded into z = M^-1 r.  Nothing here needs its result, so
// only the first work-group forms the |r|^2 total and runs it; when it ends
// the solve, the later kernels skip.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_z_top(int64_t n, const T *__restrict__ r, const T *__restrict__ d,
                                                     T *__restrict__ z, KState<T> *st, const T *prev, int prev_nb,
                                                     T *partials)
{
    if (st->skip) return;
    if (blockIdx.x == 0) {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) FinCgTop<T>{st, 1, 0, 0}(tot);
    }
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T rv = r[i];
        T zv = rv / d[i];
        z[i] = zv;
        c[0] += rv * zv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// Jacobi Cg: rho = (r, z) folded into the p update.  Every work-group forms
// the same total and beta = rho / rho_1 itself (rho_1 was set by an earlier
// kernel; the first work-group, which records rho and beta, changes neither
// rho_1 nor the iteration count), so beta is bitwise the state's.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_p_rho(int64_t n, const T *__restrict__ z, T *__restrict__ p,
                                                     KState<T> *st, const T *prev, int prev_nb)
{
    if (st->skip) return;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T pv[kPre], zv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            pv[u] = p[i];
            zv[u] = z[i];
        }
    }
    __shared__ T beta_sh;
    __shared__ int fail_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            fail_sh = tot[0] == (T)0;
            beta_sh = fail_sh ? (T)0 : tot[0] / st->rho_1;
            if (blockIdx.x == 0) FinCgRho<T>{st}(tot);
        }
        __syncthreads();
    }
    if (fail_sh) return;
    const bool first = st->nb_iter == 0;
    const T beta = beta_sh;
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            if (first) {
                p[i] = zv[u];
            } else {
                T v = pv[u] * beta;
                v += (T)1 * zv[u];
                p[i] = v;
            }
        }
    }
    for (int64_t i = i0 + kPre * stride; i < n; i += stride) {
        if (first) {
            p[i] = z[i];
        } else {
            T v = p[i] * beta;
            v += (T)1 * z[i];
            p[i] = v;
        }
    }
}

// First ? p = z : { p *= beta; p += 1 * z }   (Cg.cxx:86-94)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_p(int64_t n, const T *__restrict__ z, T *__restrict__ p, const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T beta = st->beta;
    GRID_STRIDE(i, n) {
        if (first) {
            p[i] = z[i];
        } else {
            T v = p[i] * beta;
            v += (T)1 * z[i];
            p[i] = v;
        }
    }
}

// Identity Cg: the loop head folded into p = beta p + z.  Every work-group
// forms the same |r|^2 total from k_cg_xr's partials (the finish kernel's own
// arithmetic); the first runs the loop head (rho_1 = rho, ++iter,
// Finished(r), rho = |r|^2, beta) and records rho in rho_pp[slot]; each forms
// beta = |r|^2 / rho_old itself, rho_old read from rho_pp[slot ^ 1], which an
// earlier kernel wrote (the first work-group rewrites st->rho meanwhile).
// Same division as the head's, so beta is bitwise the same.  When the head
// ends the solve, the later kernels skip; p is then not used again.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_p_top(int64_t n, const T *__restrict__ z, T *__restrict__ p,
                                                     KState<T> *st, const T *prev, int prev_nb, int slot)
{
    if (st->skip) return;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T pv[kPre], zv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            pv[u] = p[i];
            zv[u] = z[i];
        }
    }
    __shared__ T beta_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            beta_sh = tot[0] / st->rho_pp[slot ^ 1];
            if (blockIdx.x == 0) FinCgTop<T>{st, 1, 1, slot}(tot);
        }
        __syncthreads();
    }
    const T beta = beta_sh;
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            T v = pv[u] * beta;
            v += (T)1 * zv[u];
            p[i] = v;
        }
    }
    for (int64_t i = i0 + kPre * stride; i < n; i += stride) {
        T v = p[i] * beta;
        v += (T)1 * z[i];
        p[i] = v;
    }
}

// delta = (p, q) from the SpMV's partials and alpha = rho / delta (the
// finish folded in: every work-group forms the same total, the first one
// records it, a zero delta ends the solve, Cg.cxx:98-104); then
// x += alpha p; r += (-alpha) q; sum r^2   (Cg.cxx:107-108, then Finished)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_xr(int64_t n, const T *__restrict__ p, const T *__restrict__ q, T *__restrict__ x, T *__restrict__ r,
                                                  KState<T> *st, const T *prev, int prev_nb, T *partials)
{
    if (st->skip) return;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T pv[kPre], qv[kPre], xv[kPre], rv0[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            pv[u] = p[i];
            qv[u] = q[i];
            xv[u] = x[i];
            rv0[u] = r[i];
        }
    }
    __shared__ T alpha_sh;
    __shared__ int fail_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            fail_sh = tot[0] == (T)0;
            alpha_sh = fail_sh ? (T)0 : st->rho / tot[0];
            if (blockIdx.x == 0) FinCgDelta<T>{st}(tot);
        }
        __syncthreads();
    }
    if (fail_sh) return;
    const T alpha = alpha_sh, malpha = -alpha_sh;
    const bool ax = alpha != (T)0, ar = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            if (ax) x[i] = xv[u] + alpha * pv[u];
            T rv = rv0[u];
            if (ar) {
                rv += malpha * qv[u];
                r[i] = rv;
            }
            c[0] += rv * rv;
        }
    }
    for (int64_t i = i0 + kPre * stride; i < n; i += stride) {
        if (ax) x[i] += alpha * p[i];
        T rv = r[i];
        if (ar) {
            rv += malpha * q[i];
            r[i] = rv;
        }
        c[0] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// ---------------------------------------------------------- BiCgStab ----
// [x = 0, r = b]; rtilde = r; sums r^2 and rtilde*r  (BiCgStab.cxx:61-71)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_init(int64_t n, const T *__restrict__ b, T *__restrict__ x, T *__restrict__ r, T *__restrict__ rt, int zero_x,
                                                    KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T v;
        if (zero_x) {
            x[i] = (T)0;
            v = b[i];
            r[i] = v;
        } else {
            v = r[i];
        }
        rt[i] = v;
        c[0] += v * v;
        c[1] += v * v;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// p update (:81-96) and phat = M^{-1} p (:98)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_p(int64_t n, const T *__restrict__ r, const T *__restrict__ v, T *__restrict__ p, const T *__restrict__ d,
                                                 T *__restrict__ phat, const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T momega = -st->omega, beta = st->beta;
    const bool aw = momega != (T)0;
    GRID_STRIDE(i, n) {
        T pv;
        if (first) {
            pv = r[i];
        } else {
            pv = p[i];
            if (aw) pv += momega * v[i];
            pv *= beta;
            pv += (T)1 * r[i];
        }
        p[i] = pv;
        if (d) phat[i] = pv / d[i];
    }
}

// BiCgStab: the loop head (rho_2 = rho_1, ++iter, Finished(r), rho_1 =
// (rtilde, r), breakdowns #1 and #2, beta) folded into the p update of loop
// slots after a chunk's first.  Every work-group forms the same totals and
// beta = (rho_1 / rho_2) * (alpha / omega) itself, rho_2 read from
// rho_pp[slot ^ 1] (the first work-group, which runs the head, rewrites
// rho_1 meanwhile; alpha and omega come from earlier kernels), the head's
// own expression, so beta is bitwise the state's.  When the head ends the
// solve, the later kernels skip; p is then not used again.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_p_top(int64_t n, const T *__restrict__ r, const T *__restrict__ v,
                                                     T *__restrict__ p, const T *__restrict__ d, T *__restrict__ phat,
                                                     KState<T> *st, const T *prev, int prev_nb, int slot)
{
    if (st->skip) return;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T pv0[kPre], vv[kPre], rv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            pv0[u] = p[i];
            vv[u] = v[i];
            rv[u] = r[i];
        }
    }
    __shared__ T beta_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            beta_sh = (tot[1] / st->rho_pp[slot ^ 1]) * (st->alpha / st->omega);
            if (blockIdx.x == 0) FinBiTop<T>{st, 1, slot}(tot);
        }
        __syncthreads();
    }
    const T momega = -st->omega, beta = beta_sh;
    const bool aw = momega != (T)0;
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            T pv = pv0[u];
            if (aw) pv += momega * vv[u];
            pv *= beta;
            pv += (T)1 * rv[u];
            p[i] = pv;
            if (d) phat[i] = pv / d[i];
        }
    }
    for (int64_t i = i0 + kPre * stride; i < n; i += stride) {
        T pv = p[i];
        if (aw) pv += momega * v[i];
        pv *= beta;
        pv += (T)1 * r[i];
        p[i] = pv;
        if (d) phat[i] = pv / d[i];
    }
}

// s = r; s += (-alpha) v (:111-112); shat = M^{-1} s (:124); sum s^2
// (sigma = (rtilde, v) from the SpMV's partials, breakdown #3 and
// alpha = rho_1 / sigma folded in, :104-110)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_s(int64_t n, const T *__restrict__ r, const T *__restrict__ v, T *__restrict__ s, const T *__restrict__ d,
                                                 T *__restrict__ shat, KState<T> *st, const T *prev, int prev_nb,
                                                 T *partials)
{
    if (st->skip) return;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T rp[kPre], vp[kPre], dp[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            rp[u] = r[i];
            vp[u] = v[i];
            if (d) dp[u] = d[i];
        }
    }
    __shared__ T alpha_sh;
    __shared__ int fail_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            fail_sh = tot[0] == (T)0;
            alpha_sh = fail_sh ? (T)0 : st->rho_1 / tot[0];
            if (blockIdx.x == 0) FinBiSigma<T>{st}(tot);
        }
        __syncthreads();
    }
    if (fail_sh) return;
    const T malpha = -alpha_sh;
    const bool aa = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            T sv = rp[u];
            if (aa) sv += malpha * vp[u];
            s[i] = sv;
            if (d) shat[i] = sv / dp[u];
            c[0] += sv * sv;
        }
    }
    for (int64_t i = i0 + kPre * stride; i < n; i += stride) {
        T sv = r[i];
        if (aa) sv += malpha * v[i];
        s[i] = sv;
        if (d) shat[i] = sv / d[i];
        c[0] += sv * sv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// Finished(s) ended the solve in this slot: x += alpha phat (:116-120)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_fix(int64_t n, const T *__restrict__ phat, T *__restrict__ x, const KState<T> *st,
                                                   int slot)
{
    if (st->stop_slot != slot) return;
    const T alpha = st->alpha;
    if (alpha == (T)0) return;
    GRID_STRIDE(i, n) x[i] += alpha * phat[i];
}

// x += alpha phat; x += omega shat; r = s; r += (-omega) t (:132-137);
// sums r^2, rtilde*r for the next loop head
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_xr(int64_t n, const T *__restrict__ phat, const T *__restrict__ shat, const T *__restrict__ s,
                                                  const T *__restrict__ t, const T *__restrict__ rt, T *__restrict__ x, T *__restrict__ r, KState<T> *st,
                                                  const T *prev, int prev_nb, T *partials)
{
    if (st->skip) return;
    // six vectors: half the usual number of elements ahead of the fold
    constexpr int P = kPre / 2;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T xp[P], php[P], shp[P], sp[P], tp[P], rtp[P];
#pragma unroll
    for (int u = 0; u < P; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            xp[u] = x[i];
            php[u] = phat[i];
            shp[u] = shat[i];
            sp[u] = s[i];
            tp[u] = t[i];
            rtp[u] = rt[i];
        }
    }
    // omega = (t, s) / (t, t) from the SpMV's partials (:129), folded in
    __shared__ T omega_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            omega_sh = tot[0] / tot[1];
            if (blockIdx.x == 0) FinBiOmega<T>{st}(tot);
        }
        __syncthreads();
    }
    const T alpha = st->alpha, omega = omega_sh, momega = -omega_sh;
    const bool aa = alpha != (T)0, ao = omega != (T)0, am = momega != (T)0;
    T c[2] = {(T)0, (T)0};
#pragma unroll
    for (int u = 0; u < P; u++) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            T xv = xp[u];
            if (aa) xv += alpha * php[u];
            if (ao) xv += omega * shp[u];
            x[i] = xv;
            T rv = sp[u];
            if (am) rv += momega * tp[u];
            r[i] = rv;
            c[0] += rv * rv;
            c[1] += rtp[u] * rv;
        }
    }
    for (int64_t i = i0 + P * stride; i < n; i += stride) {
        T xv = x[i];
        if (aa) xv += alpha * phat[i];
        if (ao) xv += omega * shat[i];
        x[i] = xv;
        T rv = s[i];
        if (am) rv += momega * t[i];
        r[i] = rv;
        c[0] += rv * rv;
        c[1] += rt[i] * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// --------------------------------------------------------------- Cgs ----
// [x = 0, r = b]; rtilde = r; sums |r|^2 and (rtilde, r) for the first
// loop head (Cgs.cxx:63-79).  BiCg starts the same way, its head's sum
// being (z, rtilde) with z = M^{-1} r, z~ = M^{-T} r~ (BiCg.cxx:64-80; the
// Jacobi preconditioner divides by a_ii both ways).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lz_init(int64_t n, const T *__restrict__ b, T *__restrict__ x,
                                                    T *__restrict__ r, T *__restrict__ rt, const T *__restrict__ d,
                                                    T *__restrict__ z, T *__restrict__ zt, int zero_x, KState<T> *st,
                                                    T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T v;
        if (zero_x) {
            x[i] = (T)0;
            v = b[i];
            r[i] = v;
        } else {
            v = r[i];
        }
        rt[i] = v;
        c[0] += v * v;
        if (z) {  // BiCg: (z, rtilde)
            T zv = v, ztv = v;
            if (d) {
                zv = v / d[i];
                ztv = v / d[i];
                z[i] = zv;
                zt[i] = ztv;
            }
            c[1] += zv * v;
        } else {  // Cgs: (rtilde, r)
            c[1] += v * v;
        }
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// first ? u = r, p = u : u = r + beta q, p = beta (beta p + q) + u, as
// Copy / Add / Mlt / Add / Mlt / Add (Cgs.cxx:86-102); phat = M^{-1} p
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgs_pu(int64_t n, const T *__restrict__ r, const T *__restrict__ q,
                                                   T *__restrict__ p, T *__restrict__ u, const T *__restrict__ d,
                                                   T *__restrict__ phat, const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T beta = st->beta;
    const bool ab = beta != (T)0;
    GRID_STRIDE(i, n) {
        const T rv = r[i];
        T uv = rv, pv = rv;
        if (!first) {
            const T qv = q[i];
            if (ab) uv += beta * qv;
            pv = p[i] * beta;
            pv += (T)1 * qv;
            pv *= beta;
            pv += (T)1 * uv;
        }
        u[i] = uv;
        p[i] = pv;
        if (d) phat[i] = pv / d[i];
    }
}

// ++iter; delta = (rtilde, vhat) from the product's partials, breakdown #2
// and alpha folded in; q = u - alpha vhat; u += q; uhat = M^{-1} u;
// x += alpha uhat (Cgs.cxx:106-127)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgs_qux(int64_t n, const T *__restrict__ vhat, T *__restrict__ q,
                                                    T *__restrict__ u, const T *__restrict__ d, T *__restrict__ uhat,
                                                    T *__restrict__ x, KState<T> *st, const T *prev, int prev_nb)
{
    if (st->skip) return;
    __shared__ T alpha_sh;
    __shared__ int fail_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            fail_sh = tot[0] == (T)0;
            alpha_sh = fail_sh ? (T)0 : st->rho_1 / tot[0];
            if (blockIdx.x == 0) FinLanczosDelta<T>{st}(tot);
        }
        __syncthreads();
    }
    if (fail_sh) return;
    const T alpha = alpha_sh, malpha = -alpha_sh;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    GRID_STRIDE(i, n) {
        T qv = u[i];
        if (am) qv += malpha * vhat[i];
        q[i] = qv;
        T uv = u[i];
        uv += (T)1 * qv;
        u[i] = uv;
        T uh = uv;
        if (d) {
            uh = uv / d[i];
            uhat[i] = uh;
        }
        if (aa) x[i] += alpha * uh;
    }
}

// r -= alpha qhat (Cgs.cxx:130); sums |r|^2 and (rtilde, r) for the next
// loop head
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgs_r(int64_t n, const T *__restrict__ qhat, const T *__restrict__ rt,
                                                  T *__restrict__ r, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T malpha = -st->alpha;
    const bool am = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T rv = r[i];
        if (am) {
            rv += malpha * qhat[i];
            r[i] = rv;
        }
        c[0] += rv * rv;
        c[1] += rt[i] * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// -------------------------------------------------------------- BiCg ----
// first ? p = z, p~ = z~ : p = beta p + z, p~ = beta p~ + z~ as Mlt / Add
// (BiCg.cxx:88-102)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_p(int64_t n, const T *__restrict__ z, const T *__restrict__ zt,
                                                   T *__restrict__ p, T *__restrict__ pt, const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T beta = st->beta;
    GRID_STRIDE(i, n) {
        if (first) {
            p[i] = z[i];
            pt[i] = zt[i];
        } else {
            T pv = p[i] * beta;
            pv += (T)1 * z[i];
            p[i] = pv;
            T ptv = pt[i] * beta;
            ptv += (T)1 * zt[i];
            pt[i] = ptv;
        }
    }
}

// ++iter; delta = (p~, q) from the first product's partials, breakdown #2
// and alpha folded in; x += alpha p; r -= alpha q; r~ -= alpha q~
// (BiCg.cxx:104-125); then z = M^{-1} r, z~ = M^{-T} r~ and the sums |r|^2,
// (z, r~) for the next loop head
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_xr(int64_t n, const T *__restrict__ p, const T *__restrict__ q,
                                                    const T *__restrict__ qt, T *__restrict__ x, T *__restrict__ r,
                                                    T *__restrict__ rt, const T *__restrict__ d, T *__restrict__ z,
                                                    T *__restrict__ zt, KState<T> *st, const T *prev, int prev_nb,
                                                    T *partials)
{
    if (st->skip) return;
    __shared__ T alpha_sh;
    __shared__ int fail_sh;
    {
        T tot[2];
        grid_total<T>(prev, prev_nb, tot);
        if (threadIdx.x == 0) {
            fail_sh = tot[0] == (T)0;
            alpha_sh = fail_sh ? (T)0 : st->rho_1 / tot[0];
            if (blockIdx.x == 0) FinLanczosDelta<T>{st}(tot);
        }
        __syncthreads();
    }
    if (fail_sh) return;
    const T alpha = alpha_sh, malpha = -alpha_sh;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (aa) x[i] += alpha * p[i];
        T rv = r[i], rtv = rt[i];
        if (am) {
            rv += malpha * q[i];
            rtv += malpha * qt[i];
            r[i] = rv;
            rt[i] = rtv;
        }
        c[0] += rv * rv;
        T zv = rv;
        if (d) {
            zv = rv / d[i];
            z[i] = zv;
            zt[i] = rtv / d[i];
        }
        c[1] += zv * rtv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// ------------------------------------------------------------- TfQmr ----
// r0 = M^{-1} tmp (tmp = b - A x, or b with x = 0); w = y1 = rtilde = r0;
// d = 0; sums |r0|^2 and (rtilde, r0) (TfQmr.cxx:58-110)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tq_init(int64_t n, const T *__restrict__ src, const T *__restrict__ dg,
                                                    T *__restrict__ x, T *__restrict__ w, T *__restrict__ y1,
                                                    T *__restrict__ rt, T *__restrict__ d, int zero_x, KState<T> *st,
                                                    T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        T v = src[i];
        if (dg) v = v / dg[i];
        w[i] = v;
        y1[i] = v;
        rt[i] = v;
        d[i] = (T)0;
        c[0] += v * v;
        c[1] += v * v;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// v = M^{-1} tmp; g = v (the start, :85-93) or, at a loop's end, g = M^{-1}
// tmp and v = beta^2 v + beta h + g (:227-236); both sum (rtilde, v)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tq_v(int64_t n, const T *__restrict__ tmp, const T *__restrict__ dg,
                                                 T *__restrict__ g, T *__restrict__ v, const T *__restrict__ h,
                                                 const T *__restrict__ rt, int start, const KState<T> *st,
                                                 T *partials)
{
    if (st->skip) return;
    const T beta = st->beta, bb = beta * beta;
    const bool ab = beta != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T gv = tmp[i];
        if (dg) gv = gv / dg[i];
        g[i] = gv;
        T vv = gv;
        if (!start) {
            vv = v[i] * bb;
            if (ab) vv += beta * h[i];
            vv += (T)1 * gv;
        }
        v[i] = vv;
        c[0] += rt[i] * vv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// y0 = y1 - alpha v (:123-124)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tq_y0(int64_t n, const T *__restrict__ y1, const T *__restrict__ v,
                                                  T *__restrict__ y0, const KState<T> *st)
{
    if (st->skip) return;
    const T malpha = -st->alpha;
    const bool am = malpha != (T)0;
    GRID_STRIDE(i, n) {
        T yv = y1[i];
        if (am) yv += malpha * v[i];
        y0[i] = yv;
    }
}

// first half step: h = M^{-1} tmp (tmp = A y0); w -= alpha g;
// d = (theta^2 eta / alpha) d + y1, the factor in double (:127-143);
// sums |w|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tq_h1(int64_t n, const T *__restrict__ tmp, const T *__restrict__ dg,
                                                  T *__restrict__ h, const T *__restrict__ g, T *__restrict__ w,
                                                  T *__restrict__ d, const T *__restrict__ y1, const KState<T> *st,
                                                  T *partials)
{
    if (st->skip) return;
    const T alpha = st->alpha, malpha = -alpha;
    const bool am = malpha != (T)0;
    const double s = st->tq_theta * st->tq_theta * (double)st->tq_eta / (double)alpha;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T hv = tmp[i];
        if (dg) hv = hv / dg[i];
        h[i] = hv;
        T wv = w[i];
        if (am) wv += malpha * g[i];
        w[i] = wv;
        T dv = d[i];
        dv *= s;
        dv += (T)1 * y1[i];
        d[i] = dv;
        c[0] += wv * wv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// x += eta d of the first half (:159, also when the solve ends there); then
// the second half step: g = h; w -= alpha g; d = (theta^2 eta / alpha) d +
// y0 (:170-178); sums |w|^2 and (rtilde, w)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tq_h2(int64_t n, const T *__restrict__ h, T *__restrict__ g,
                                                  T *__restrict__ w, T *__restrict__ d, const T *__restrict__ y0,
                                                  T *__restrict__ x, const T *__restrict__ rt, const KState<T> *st,
                                                  int slot, T *partials)
{
    const bool here = st->stop_slot == 2 * slot;
    if (st->skip && !here) return;
    const T eta = st->tq_eta;
    const bool ae = eta != (T)0;
    if (st->skip) {  // the solve ended at this half's Finished
        if (ae) GRID_STRIDE(i, n) x[i] += eta * d[i];
        return;
    }
    const T alpha = st->alpha, malpha = -alpha;
    const bool am = malpha != (T)0;
    const double s = st->tq_theta * st->tq_theta * (double)eta / (double)alpha;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T dv = d[i];
        if (ae) x[i] += eta * dv;
        const T hv = h[i];
        g[i] = hv;
        T wv = w[i];
        if (am) wv += malpha * hv;
        w[i] = wv;
        dv *= s;
        dv += (T)1 * y0[i];
        d[i] = dv;
        c[0] += wv * wv;
        c[1] += rt[i] * wv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// x += eta d of the second half (:198, also when the solve ends there);
// then y1 = w + beta y0 (:216-217)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tq_y1(int64_t n, const T *__restrict__ w, const T *__restrict__ y0,
                                                  const T *__restrict__ d, T *__restrict__ x, T *__restrict__ y1,
                                                  const KState<T> *st, int slot)
{
    const bool here = st->stop_slot == 2 * slot + 1;
    if (st->skip && !here) return;
    const T eta = st->tq_eta, beta = st->beta;
    const bool ae = eta != (T)0, ab = beta != (T)0;
    const bool ended = st->skip;
    GRID_STRIDE(i, n) {
        if (ae) x[i] += eta * d[i];
        if (!ended) {
            T yv = w[i];
            if (ab) yv += beta * y0[i];
            y1[i] = yv;
        }
    }
}

// --------------------------------------------------------------- Qmr ----
// [x = 0]; r = src (b, or b - A x); v = r; y = M^{-1} v; w = r; d = s = b
// (copies of b, Qmr.cxx:53-76); sums |y|^2 and |w|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_init(int64_t n, const T *__restrict__ src, const T *__restrict__ b,
                                                     const T *__restrict__ dg, T *__restrict__ x, T *__restrict__ r,
                                                     T *__restrict__ v, T *__restrict__ y, T *__restrict__ w,
                                                     T *__restrict__ d, T *__restrict__ s, int zero_x,
                                                     KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        const T rv = src[i];
        r[i] = rv;
        v[i] = rv;
        T yv = rv;
        if (dg) yv = rv / dg[i];
        y[i] = yv;
        w[i] = rv;
        const T bv = b[i];
        d[i] = bv;
        s[i] = bv;
        c[0] += yv * yv;
        c[1] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// v /= rho; y /= rho; w /= xi as Mlt(Complexe(1 / .)) (:92-97);
// z~ = M^{-T} w (:106); sums (w, y)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_scale(int64_t n, T *__restrict__ v, T *__restrict__ y,
                                                      T *__restrict__ w, const T *__restrict__ dg, T *__restrict__ z,
                                                      const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T sv = (T)(1. / st->qm_rho), sw = (T)(1. / st->qm_xi);
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        v[i] = v[i] * sv;
        const T yv = y[i] * sv;
        y[i] = yv;
        const T wv = w[i] * sw;
        w[i] = wv;
        if (dg) z[i] = wv / dg[i];
        c[0] += wv * yv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// first ? p = y, q = z~ : p = y - (xi delta / ep) p, q = z~ - (rho delta /
// ep) q as Mlt / Add (:108-121)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_pq(int64_t n, const T *__restrict__ y, const T *__restrict__ z,
                                                   T *__restrict__ p, T *__restrict__ q, const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const double delta = (double)st->delta, ep = (double)st->qm_ep;
    const T cp = (T)(-(st->qm_xi * delta / ep)), cq = (T)(-(st->qm_rho * delta / ep));
    GRID_STRIDE(i, n) {
        if (first) {
            p[i] = y[i];
            q[i] = z[i];
        } else {
            T pv = p[i] * cp;
            pv += (T)1 * y[i];
            p[i] = pv;
            T qv = q[i] * cq;
            qv += (T)1 * z[i];
            q[i] = qv;
        }
    }
}

// v = -beta v + p~; y = M^{-1} v (:139-141)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_vy(int64_t n, const T *__restrict__ pt, const T *__restrict__ dg,
                                                   T *__restrict__ v, T *__restrict__ y, const KState<T> *st)
{
    if (st->skip) return;
    const T mb = -st->beta;
    GRID_STRIDE(i, n) {
        T vv = v[i] * mb;
        vv += (T)1 * pt[i];
        v[i] = vv;
        y[i] = dg ? vv / dg[i] : vv;
    }
}

// w = -beta w + z~ (z~ = A^T q, :147-150); sums |y|^2 and |w|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_w(int64_t n, const T *__restrict__ z, const T *__restrict__ y,
                                                  T *__restrict__ w, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T mb = -st->beta;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T wv = w[i] * mb;
        wv += (T)1 * z[i];
        w[i] = wv;
        const T yv = y[i];
        c[0] += yv * yv;
        c[1] += wv * wv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// d = theta_1^2 gamma^2 d + eta p; s likewise with p~; x += d; r -= s
// (:174-187); sums |r|^2 for the next loop head
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_dsxr(int64_t n, const T *__restrict__ p, const T *__restrict__ pt,
                                                     T *__restrict__ d, T *__restrict__ s, T *__restrict__ x,
                                                     T *__restrict__ r, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T tmp = st->qm_tmp, eta = st->qm_eta;
    const bool ae = eta != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T dv = d[i] * tmp;
        if (ae) dv += eta * p[i];
        d[i] = dv;
        T sv = s[i] * tmp;
        if (ae) sv += eta * pt[i];
        s[i] = sv;
        x[i] += (T)1 * dv;
        T rv = r[i];
        rv += -(T)1 * sv;
        r[i] = rv;
        c[0] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// -------------------------------------------------------------- Cgne ----
// |q|^2 of q = A^T b (Iteration::Init and the first stopping test)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sumsq(int64_t n, const T *__restrict__ q, const KState<T> *st,
                                                  T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        const T v = q[i];
        c[0] += v * v;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// start ? r = q, x = 0 : r = -(A^T A x) + q (Cgne.cxx:64-78; r holds A^T A x)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgne_r0(int64_t n, const T *__restrict__ q, T *__restrict__ r,
                                                    T *__restrict__ x, int zero_x, const KState<T> *st)
{
    if (st->skip) return;
    GRID_STRIDE(i, n) {
        if (zero_x) {
            r[i] = q[i];
            x[i] = (T)0;
        } else {
            T rv = r[i] * (T)-1;
            rv += (T)1 * q[i];
            r[i] = rv;
        }
    }
}

// z = M^{-1} M^{-T} r (two divisions by a_ii with Jacobi, :85-86); sums
// (r, z)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgne_z(int64_t n, const T *__restrict__ r, const T *__restrict__ dg,
                                                   T *__restrict__ z, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        const T rv = r[i];
        T zv = rv;
        if (dg) {
            const T qv = rv / dg[i];
            zv = qv / dg[i];
            z[i] = zv;
        }
        c[0] += rv * zv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// first ? p = z : p = beta p + z (:94-101)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgne_p(int64_t n, const T *__restrict__ z, T *__restrict__ p,
                                                   const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T beta = st->beta;
    GRID_STRIDE(i, n) {
        if (first) {
            p[i] = z[i];
        } else {
            T pv = p[i] * beta;
            pv += (T)1 * z[i];
            p[i] = pv;
        }
    }
}

// x += alpha p; r -= alpha z (z = A^T A p, :114-115); sums |r|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgne_xr(int64_t n, const T *__restrict__ p, const T *__restrict__ z,
                                                    T *__restrict__ x, T *__restrict__ r, const KState<T> *st,
                                                    T *partials)
{
    if (st->skip) return;
    const T alpha = st->alpha, malpha = -alpha;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (aa) x[i] += alpha * p[i];
        T rv = r[i];
        if (am) {
            rv += malpha * z[i];
            r[i] = rv;
        }
        c[0] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// -------------------------------------------------------------- Lsqr ----
// sums |u|^2 (u = b - A x, or b) and |b|^2 (Lsqr.cxx:57-65)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lsqr_norms(int64_t n, const T *__restrict__ u, const T *__restrict__ b,
                                                       const KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        const T uv = u[i], bv = b[i];
        c[0] += uv * uv;
        c[1] += bv * bv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// u = b * (1 / beta); x = 0 (:66-75: the template zeroes x whatever the
// initial guess)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lsqr_u0(int64_t n, const T *__restrict__ b, T *__restrict__ u,
                                                    T *__restrict__ x, const KState<T> *st)
{
    if (st->skip) return;
    const T t = (T)(1.0 / st->ls_beta);
    GRID_STRIDE(i, n) {
        u[i] = b[i] * t;
        x[i] = (T)0;
    }
}

// v *= 1 / alpha; w = v (:71-74)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lsqr_v0(int64_t n, T *__restrict__ v, T *__restrict__ w,
                                                    const KState<T> *st)
{
    if (st->skip) return;
    const T t = (T)(1.0 / st->ls_alpha);
    GRID_STRIDE(i, n) {
        const T vv = v[i] * t;
        v[i] = vv;
        w[i] = vv;
    }
}

// y += (-c) x with c a double scale (Add(-alpha, u, u1), Add(-beta, v, v1));
// sums |y|^2 (:84-86, :95-97)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lsqr_sub(int64_t n, const T *__restrict__ xv, T *__restrict__ y,
                                                     int which, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T m = (T)(-(which ? st->ls_beta : st->ls_alpha));
    const bool am = m != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T yv = y[i];
        if (am) {
            yv += m * xv[i];
            y[i] = yv;
        }
        c[0] += yv * yv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// u1 *= 1 / beta (:92-93)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lsqr_scale(int64_t n, T *__restrict__ y, const KState<T> *st)
{
    if (st->skip) return;
    const T t = (T)(1.0 / st->ls_beta);
    GRID_STRIDE(i, n) y[i] = y[i] * t;
}

// v1 *= 1 / alpha; x += (phi / rho) w; w = (-theta / rho) w + v1
// (:103-104, :122-128); also when this loop's Finished ended the solve
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lsqr_xw(int64_t n, T *__restrict__ v1, T *__restrict__ w,
                                                    T *__restrict__ x, const KState<T> *st, int slot)
{
    if (st->skip && st->stop_slot != slot) return;
    const T t = (T)(1.0 / st->ls_alpha), xs = st->ls_xs, ws = st->ls_ws;
    const bool ax = xs != (T)0;
    GRID_STRIDE(i, n) {
        const T vv = v1[i] * t;
        v1[i] = vv;
        const T wv = w[i];
        if (ax) x[i] += xs * wv;
        T wn = wv * ws;
        wn += (T)1 * vv;
        w[i] = wn;
    }
}

// ------------------------------------------------------------ BiCgcr ----
// v = src (b, or b - A x); [x = 0]; s = M^{-1} v; p = s (BiCgcr.cxx:62-73);
// sums |v|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bcr_init(int64_t n, const T *__restrict__ src, const T *__restrict__ dg,
                                                     T *__restrict__ x, T *__restrict__ v, T *__restrict__ s,
                                                     T *__restrict__ p, int zero_x, const KState<T> *st,
                                                     T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        const T vv = src[i];
        v[i] = vv;
        T sv = vv;
        if (dg) sv = vv / dg[i];
        s[i] = sv;
        p[i] = sv;
        c[0] += vv * vv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// start ? (w = M^{-1} a) : (p = -beta p + s; a = -beta a + z; w = M^{-1} a)
// (:75, :106-111); sums (w, v) and (w, a) for the next loop
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bcr_paw(int64_t n, const T *__restrict__ sv, const T *__restrict__ z,
                                                    const T *__restrict__ dg, const T *__restrict__ v,
                                                    T *__restrict__ p, T *__restrict__ a, T *__restrict__ w,
                                                    int start, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T mb = -st->beta;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T av = a[i];
        if (!start) {
            T pv = p[i] * mb;
            pv += (T)1 * sv[i];
            p[i] = pv;
            av = av * mb;
            av += (T)1 * z[i];
            a[i] = av;
        }
        T wv = av;
        if (dg) {
            wv = av / dg[i];
            w[i] = wv;
        }
        c[0] += wv * v[i];
        c[1] += wv * av;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// x += alpha p; v -= alpha a; s -= alpha w (:91-95); sums |v|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bcr_xvs(int64_t n, const T *__restrict__ p, const T *__restrict__ a,
                                                    const T *__restrict__ w, T *__restrict__ x, T *__restrict__ v,
                                                    T *__restrict__ s, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T alpha = st->alpha, malpha = -alpha;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (aa) x[i] += alpha * p[i];
        T vv = v[i];
        if (am) {
            vv += malpha * a[i];
            v[i] = vv;
            s[i] += malpha * w[i];
        }
        c[0] += vv * vv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// ------------------------------------------------------------ QmrSym ----
// [x = 0]; r = src; v = r; y = M^{-1} v (QmrSym.cxx:62-72); sums |y|^2, |r|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qms_init(int64_t n, const T *__restrict__ src, const T *__restrict__ dg,
                                                     T *__restrict__ x, T *__restrict__ r, T *__restrict__ v,
                                                     T *__restrict__ y, int zero_x, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        const T rv = src[i];
        r[i] = rv;
        v[i] = rv;
        T yv = rv;
        if (dg) yv = rv / dg[i];
        y[i] = yv;
        c[0] += yv * yv;
        c[1] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// v /= rho; y /= rho (:88-89); sums (v, y)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qms_scale(int64_t n, T *__restrict__ v, T *__restrict__ y,
                                                      const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T sv = (T)(1. / st->qm_rho);
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        const T vv = v[i] * sv;
        v[i] = vv;
        const T yv = y[i] * sv;
        y[i] = yv;
        c[0] += vv * yv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// first ? p = y : p = y - (rho delta / ep) p (:98-104)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qms_p(int64_t n, const T *__restrict__ y, T *__restrict__ p,
                                                  const KState<T> *st)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T cp = (T)(-(st->qm_rho * (double)st->delta / (double)st->qm_ep));
    GRID_STRIDE(i, n) {
        if (first) {
            p[i] = y[i];
        } else {
            T pv = p[i] * cp;
            pv += (T)1 * y[i];
            p[i] = pv;
        }
    }
}

// v = -beta v + p~; y = M^{-1} v (:121-122); sums |y|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qms_vy(int64_t n, const T *__restrict__ pt, const T *__restrict__ dg,
                                                   T *__restrict__ v, T *__restrict__ y, const KState<T> *st,
                                                   T *partials)
{
    if (st->skip) return;
    const T mb = -st->beta;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T vv = v[i] * mb;
        vv += (T)1 * pt[i];
        v[i] = vv;
        const T yv = dg ? vv / dg[i] : vv;
        y[i] = yv;
        c[0] += yv * yv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// first ? d = eta p, s = eta p~ : d = tmp d + eta p, s = tmp s + eta p~;
// x += d; r -= s (:142-157); sums |r|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qms_dsxr(int64_t n, const T *__restrict__ p, const T *__restrict__ pt,
                                                     T *__restrict__ d, T *__restrict__ s, T *__restrict__ x,
                                                     T *__restrict__ r, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const bool first = st->nb_iter == 0;
    const T tmp = st->qm_tmp, eta = st->qm_eta;
    const bool ae = eta != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T dv, sv;
        if (first) {
            dv = p[i] * eta;
            sv = pt[i] * eta;
        } else {
            dv = d[i] * tmp;
            if (ae) dv += eta * p[i];
            sv = s[i] * tmp;
            if (ae) sv += eta * pt[i];
        }
        d[i] = dv;
        s[i] = sv;
        x[i] += (T)1 * dv;
        T rv = r[i];
        rv += -(T)1 * sv;
        r[i] = rv;
        c[0] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// ------------------------------------------------------------ MinRes ----
// [x = 0]; r = src; z = M^{-1} r (MinRes.cxx:66-79); sums (r, z), |b|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mr_init(int64_t n, const T *__restrict__ src, const T *__restrict__ b,
                                                    const T *__restrict__ dg, T *__restrict__ x, T *__restrict__ r,
                                                    T *__restrict__ z, int zero_x, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        const T rv = src[i];
        r[i] = rv;
        T zv = rv;
        if (dg) zv = rv / dg[i];
        z[i] = zv;
        c[0] += rv * zv;
        const T bv = b[i];
        c[1] += bv * bv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// v = r ibeta; u = z ibeta; u_old, v_old, w_old, w, w_oold = 0 (:73-84)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mr_start(int64_t n, const T *__restrict__ r, const T *__restrict__ z,
                                                     T *__restrict__ v, T *__restrict__ u, T *__restrict__ v_old,
                                                     T *__restrict__ u_old, T *__restrict__ w0, T *__restrict__ w1,
                                                     T *__restrict__ w2, const KState<T> *st)
{
    if (st->skip) return;
    const T ib = st->mr[kMrIbeta];
    GRID_STRIDE(i, n) {
        v[i] = r[i] * ib;
        u[i] = z[i] * ib;
        v_old[i] = (T)0;
        u_old[i] = (T)0;
        w0[i] = (T)0;
        w1[i] = (T)0;
        w2[i] = (T)0;
    }
}

// z = M^{-1} r (r = A u); r -= alpha v; z -= alpha u; r -= beta v_old;
// z -= beta u_old (:96-104); sums (r, z)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mr_rz(int64_t n, const T *__restrict__ dg, const T *__restrict__ v,
                                                  const T *__restrict__ u, const T *__restrict__ v_old,
                                                  const T *__restrict__ u_old, T *__restrict__ r, T *__restrict__ z,
                                                  const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T ma = -st->mr[kMrAlpha], mb = -st->mr[kMrBeta];
    const bool aa = ma != (T)0, ab = mb != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T rv = r[i];
        T zv = rv;
        if (dg) zv = rv / dg[i];
        if (aa) {
            rv += ma * v[i];
            zv += ma * u[i];
        }
        if (ab) {
            rv += mb * v_old[i];
            zv += mb * u_old[i];
        }
        r[i] = rv;
        z[i] = zv;
        c[0] += rv * zv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// w_new = (u - rho2 w - rho3 w_old) / rho1 into w_oold's buffer; x +=
// c eta w_new (:120-131, also when the solve ends here); then v_new =
// r ibeta, u_new = z ibeta into the old v_old / u_old buffers (:133-143)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mr_wx(int64_t n, const T *__restrict__ u, const T *__restrict__ w,
                                                  const T *__restrict__ w_old, T *__restrict__ w_new,
                                                  T *__restrict__ x, const T *__restrict__ r,
                                                  const T *__restrict__ z, T *__restrict__ v_new,
                                                  T *__restrict__ u_new, const KState<T> *st, int slot)
{
    if (st->skip && st->stop_slot != slot) return;
    const T *m = st->mr;
    const T m2 = -m[kMrRho2], m3 = -m[kMrRho3], ir = m[kMrIrho1], ce = m[kMrCeta], ib = m[kMrIbeta];
    const bool a2 = m2 != (T)0, a3 = m3 != (T)0, ac = ce != (T)0;
    const bool ended = st->skip;
    GRID_STRIDE(i, n) {
        T wv = u[i];
        if (a2) wv += m2 * w[i];
        if (a3) wv += m3 * w_old[i];
        wv *= ir;
        w_new[i] = wv;
        if (ac) x[i] += ce * wv;
        if (!ended) {
            v_new[i] = r[i] * ib;
            u_new[i] = z[i] * ib;
        }
    }
}

// ------------------------------------------------------------ Symmlq ----
// v = r ibeta; u = z ibeta; w_bar = u; v_old = u_old = 0 (Symmlq.cxx:84-89)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sy_start(int64_t n, const T *__restrict__ r, const T *__restrict__ z,
                                                     T *__restrict__ v, T *__restrict__ u, T *__restrict__ v_old,
                                                     T *__restrict__ u_old, T *__restrict__ w_bar,
                                                     const KState<T> *st)
{
    if (st->skip) return;
    const T ib = st->mr[kMrIbeta];
    GRID_STRIDE(i, n) {
        v[i] = r[i] * ib;
        const T uv = z[i] * ib;
        u[i] = uv;
        w_bar[i] = uv;
        v_old[i] = (T)0;
        u_old[i] = (T)0;
    }
}

// the update of every iteration after the first (:96-113): v = r ibeta,
// u = z ibeta (into the old v_old / u_old buffers); w = c w_bar + s u;
// w_bar = -s w_bar + c u; x += ceta w (w is not kept)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sy_upd(int64_t n, const T *__restrict__ r, const T *__restrict__ z,
                                                   T *__restrict__ v_new, T *__restrict__ u_new,
                                                   T *__restrict__ w_bar, T *__restrict__ x, const KState<T> *st)
{
    if (st->skip) return;
    const T *m = st->mr;
    const T ib = m[kMrIbeta], c = m[kMrC], s = m[kMrS], ms = -m[kMrS], ceta = m[kSyCeta];
    const bool as = s != (T)0, ac = c != (T)0, ae = ceta != (T)0;
    GRID_STRIDE(i, n) {
        v_new[i] = r[i] * ib;
        const T uv = z[i] * ib;
        u_new[i] = uv;
        const T wb = w_bar[i];
        T wv = wb * c;
        if (as) wv += s * uv;
        T wbn = wb * ms;
        if (ac) wbn += c * uv;
        w_bar[i] = wbn;
        if (ae) x[i] += ceta * wv;
    }
}

// after the loop: x += ceta_bar w_bar, ceta_bar = ceta / c (ceta 1e15 when
// c = 0) (:153-158)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sy_final(int64_t n, const T *__restrict__ w_bar, T *__restrict__ x,
                                                     const KState<T> *st)
{
    if (st->init_failed) return;
    const T c = st->mr[kMrC], ceta = st->mr[kSyCeta];
    const T cb = c == (T)0 ? (T)((double)ceta * 1e15) : ceta / c;
    if (cb == (T)0) return;
    GRID_STRIDE(i, n) x[i] += cb * w_bar[i];
}

// -------------------------------------------------------------- QCgs ----
// [x = 0]; r = src; rtilde = r_qcgs = r; x_qcgs = x; q = p = 0
// (QCgs.cxx:64-76); sums |r_qcgs|^2 and (rtilde, r)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qc_init(int64_t n, const T *__restrict__ src, T *__restrict__ x,
                                                    T *__restrict__ r, T *__restrict__ rt, T *__restrict__ rq,
                                                    T *__restrict__ xq, T *__restrict__ q, T *__restrict__ p,
                                                    int zero_x, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        const T rv = src[i];
        r[i] = rv;
        rt[i] = rv;
        rq[i] = rv;
        xq[i] = x[i];
        q[i] = (T)0;
        p[i] = (T)0;
        c[0] += rv * rv;
        c[1] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// u = r + beta q; p = beta (beta p + q) + u; phat = M^{-1} p (:92-103)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qc_up(int64_t n, const T *__restrict__ r, const T *__restrict__ q,
                                                  T *__restrict__ p, T *__restrict__ u, const T *__restrict__ dg,
                                                  T *__restrict__ phat, const KState<T> *st)
{
    if (st->skip) return;
    const T beta = st->beta;
    const bool ab = beta != (T)0;
    GRID_STRIDE(i, n) {
        const T qv = q[i];
        T uv = r[i];
        if (ab) uv += beta * qv;
        u[i] = uv;
        T pv = p[i] * beta;
        pv += (T)1 * qv;
        pv *= beta;
        pv += (T)1 * uv;
        p[i] = pv;
        if (dg) phat[i] = pv / dg[i];
    }
}

// q = u - alpha v; u += q; x += alpha u; uhat = M^{-1} u (:113-122)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qc_qu(int64_t n, const T *__restrict__ v, T *__restrict__ q,
                                                  T *__restrict__ u, T *__restrict__ x, const T *__restrict__ dg,
                                                  T *__restrict__ uhat, const KState<T> *st)
{
    if (st->skip) return;
    const T alpha = st->alpha, malpha = -alpha;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    GRID_STRIDE(i, n) {
        const T u0 = u[i];
        T qv = u0;
        if (am) qv += malpha * v[i];
        q[i] = qv;
        T uv = u0;
        uv += (T)1 * qv;
        u[i] = uv;
        if (aa) x[i] += alpha * uv;
        if (dg) uhat[i] = uv / dg[i];
    }
}

// r -= alpha A uhat; u = r_qcgs - r (:125-129); sums (u, u), (v, v),
// (u, v), (u, r), (v, r)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qc_r(int64_t n, const T *__restrict__ w, const T *__restrict__ rq,
                                                 const T *__restrict__ v, T *__restrict__ r, T *__restrict__ u,
                                                 const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T malpha = -st->alpha;
    const bool am = malpha != (T)0;
    T c[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
    GRID_STRIDE(i, n) {
        T rv = r[i];
        if (am) {
            rv += malpha * w[i];
            r[i] = rv;
        }
        T uv = rq[i];
        uv += -(T)1 * rv;
        u[i] = uv;
        const T vv = v[i];
        c[0] += uv * uv;
        c[1] += vv * vv;
        c[2] += uv * vv;
        c[3] += uv * rv;
        c[4] += vv * rv;
    }
    block_partial<T, 6>(c, partials, gridDim.x, blockIdx.x);
}

// r_qcgs = r + mu u + nu v; x_qcgs = mu x_qcgs + (1 - mu) x - nu p
// (:154-160); sums |r_qcgs|^2 and (rtilde, r) for the next loop head
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qc_smooth(int64_t n, const T *__restrict__ r, const T *__restrict__ u,
                                                      const T *__restrict__ v, const T *__restrict__ x,
                                                      const T *__restrict__ p, const T *__restrict__ rt,
                                                      T *__restrict__ rq, T *__restrict__ xq, const KState<T> *st,
                                                      T *partials)
{
    if (st->skip) return;
    const T mu = st->qc_mu, nu = st->qc_nu, omu = (T)1 - st->qc_mu, mnu = -st->qc_nu;
    const bool amu = mu != (T)0, anu = nu != (T)0, aomu = omu != (T)0, amnu = mnu != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        const T rv = r[i];
        T rqv = rv;
        if (amu) rqv += mu * u[i];
        if (anu) rqv += nu * v[i];
        rq[i] = rqv;
        T xqv = xq[i] * mu;
        if (aomu) xqv += omu * x[i];
        if (amnu) xqv += mnu * p[i];
        xq[i] = xqv;
        c[0] += rqv * rqv;
        c[1] += rt[i] * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// after the loop: x = M^{-1} x_qcgs (:162)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qc_final(int64_t n, const T *__restrict__ xq, const T *__restrict__ dg,
                                                     T *__restrict__ x, const KState<T> *st)
{
    if (st->init_failed) return;
    GRID_STRIDE(i, n) x[i] = dg ? xq[i] / dg[i] : xq[i];
}

// --------------------------------------------------------------- Gcr ----
// [x = 0]; r = M^{-1} src (src = b - A x, or b) (Gcr.cxx:71-80); sums |r|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gc_init(int64_t n, const T *__restrict__ src, const T *__restrict__ dg,
                                                    T *__restrict__ x, T *__restrict__ r, int zero_x,
                                                    const KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        T rv = src[i];
        if (dg) rv = rv / dg[i];
        r[i] = rv;
        c[0] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// p_0 = r * (1 / normr), the factor a double (:90-91)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gc_p0(int64_t n, const T *__restrict__ r, T *__restrict__ p0,
                                                  const KState<T> *st)
{
    if (st->gc_skip) return;
    const double s = 1.0 / st->gc_normr;
    GRID_STRIDE(i, n) {
        T pv = r[i];
        pv *= s;
        p0[i] = pv;
    }
}

// w_j = M^{-1} (A p_j) (:98-99); sums (w_j, w_j) and (w_j, r)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gc_w(int64_t n, const T *__restrict__ u, const T *__restrict__ dg,
                                                 const T *__restrict__ r, T *__restrict__ w, const KState<T> *st,
                                                 T *partials)
{
    if (st->gc_skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        T wv = u[i];
        if (dg) wv = wv / dg[i];
        w[i] = wv;
        c[0] += wv * wv;
        c[1] += wv * r[i];
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// x += alpha p_j; r -= alpha w_j (:109-111); sums |r|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gc_xr(int64_t n, const T *__restrict__ p, const T *__restrict__ w,
                                                  T *__restrict__ x, T *__restrict__ r, const KState<T> *st,
                                                  T *partials)
{
    if (st->gc_skip) return;
    const T alpha = st->alpha, malpha = -alpha;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (aa) x[i] += alpha * p[i];
        T rv = r[i];
        if (am) {
            rv += malpha * w[i];
            r[i] = rv;
        }
        c[0] += rv * rv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// [q = M^{-1} (A r)] (:116-117); sums (w_i, q) for up to six i from g0
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gc_q(int64_t n, const T *__restrict__ u, const T *__restrict__ dg,
                                                 T *__restrict__ q, int write_q, const T *__restrict__ wbase,
                                                 int64_t stride, int g0, int cnt, const KState<T> *st,
                                                 T *partials)
{
    if (st->gc_skip) return;
    T c[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
    GRID_STRIDE(i, n) {
        T qv;
        if (write_q) {
            qv = u[i];
            if (dg) qv = qv / dg[i];
            q[i] = qv;
        } else {
            qv = q[i];
        }
#pragma unroll
        for (int k = 0; k < 6; k++)
            if (k < cnt) c[k] += wbase[(size_t)(g0 + k) * stride + i] * qv;
    }
    block_partial<T, 6>(c, partials, gridDim.x, blockIdx.x);
}

// p_{j+1} = r + sum over i <= j of delta_i p_i, added in i order (:119-129)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gc_p(int64_t n, const T *__restrict__ r, T *__restrict__ pbase,
                                                 int64_t stride, int j, const T *__restrict__ deltas,
                                                 const KState<T> *st)
{
    if (st->gc_skip) return;
    T *pn = pbase + (size_t)(j + 1) * stride;
    GRID_STRIDE(i, n) {
        T pv = r[i];
        for (int k = 0; k <= j; k++) {
            const T d = deltas[k];
            if (d != (T)0) pv += d * pbase[(size_t)k * stride + i];
        }
        pn[i] = pv;
    }
}

// --------------------------------------------------------- BiCgStabl ----
// [x = 0]; r_0 = src (b - A x, or b); r0~ = r_0 (BiCgStabl.cxx:66-80); sums
// |r_0|^2 for the loop head's norm and for (r_0, r0~).  After a failed
// Init(b) only x.Zero(), which precedes it there, is done.  src may be r_0.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bl_init(int64_t n, const T *src, T *__restrict__ x, T *r0,
                                                    T *__restrict__ rt, int zero_x, const KState<T> *st,
                                                    T *partials)
{
    const bool dead = st->skip;
    if (dead && !zero_x) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        if (zero_x) x[i] = (T)0;
        if (dead) continue;
        const T rv = src[i];
        r0[i] = rv;
        rt[i] = rv;
        c[0] += rv * rv;
    }
    if (dead) return;
    c[1] = c[0];
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// u_i = r_i - beta u_i for i <= j (:103-106: Mlt(-beta, u_i), Add(1, r_i,
// u_i)); [q = M^{-1} u_j]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bl_u(int64_t n, const T *__restrict__ rb, T *__restrict__ ub,
                                                 int64_t stride, int j, const T *__restrict__ dg,
                                                 T *__restrict__ q, const KState<T> *st)
{
    if (st->skip) return;
    const T mb = -st->bl_beta;
    GRID_STRIDE(i, n) {
        T uv = (T)0;
        for (int k = 0; k <= j; k++) {
            uv = ub[(size_t)k * stride + i] * mb;
            uv += rb[(size_t)k * stride + i];
            ub[(size_t)k * stride + i] = uv;
        }
        if (dg) q[i] = uv / dg[i];
    }
}

// x += alpha u_0; r_i -= alpha u_{i+1} for i <= j (:117-120); [q = M^{-1} r_j]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bl_xr(int64_t n, T *__restrict__ rb, const T *__restrict__ ub,
                                                  int64_t stride, int j, const T *__restrict__ dg,
                                                  T *__restrict__ q, T *__restrict__ x, const KState<T> *st)
{
    if (st->skip) return;
    const T alpha = st->bl_alpha, malpha = -alpha;
    const bool aa = alpha != (T)0, am = malpha != (T)0;
    GRID_STRIDE(i, n) {
        if (aa) x[i] += alpha * ub[i];
        T rv = (T)0;
        for (int k = 0; k <= j; k++) {
            rv = rb[(size_t)k * stride + i];
            if (am) {
                rv += malpha * ub[(size_t)(k + 1) * stride + i];
                rb[(size_t)k * stride + i] = rv;
            }
        }
        if (dg) q[i] = rv / dg[i];
    }
}

// MR part (:131-141), step i of its modified Gram-Schmidt taken column by
// column of tau ("right-looking"): for each j of the group j0 .. j0+cnt-1
// (j >= i), r_j -= tau(i-1, j) r_{i-1} where that tau was formed (gamma(i-1)
// != 0, i > 1); then sums (r_j, r_i) for j > i, and, when the group starts
// at j0 = i, |r_i|^2 and (r_0, r_i) first.  Each r_j still takes its
// subtractions in i order, each tau(i, j) from r_j as it stands after
// i - 1 of them, as in the template's loop over j then i.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bl_mgs(int64_t n, T *__restrict__ rb, int64_t stride, int i, int j0,
                                                   int cnt, const T *__restrict__ tau, const T *__restrict__ gamma,
                                                   int ld, const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const bool formed = i > 1 && gamma[i - 1] != (T)0;
    T t[6];
#pragma unroll
    for (int k = 0; k < 6; k++) t[k] = (formed && k < cnt) ? -tau[(size_t)(i - 1) * ld + j0 + k] : (T)0;
    const bool head = j0 == i;
    T c[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
    const T *rp = rb + (size_t)(i > 1 ? i - 1 : 0) * stride;
    T *ri_p = rb + (size_t)i * stride;
    GRID_STRIDE(e, n) {
        const T pv = formed ? rp[e] : (T)0;
        T ri = ri_p[e];
        if (head) {
            if (t[0] != (T)0) {
                ri += t[0] * pv;
                ri_p[e] = ri;
            }
            c[0] += ri * ri;
            c[1] += rb[e] * ri;
        }
#pragma unroll
        for (int k = 0; k < 6; k++) {
            if (k >= cnt || (head && k == 0)) continue;
            T *rj = rb + (size_t)(j0 + k) * stride;
            T v = rj[e];
            if (t[k] != (T)0) {
                v += t[k] * pv;
                rj[e] = v;
            }
            c[head ? k + 1 : k] += v * ri;
        }
    }
    block_partial<T, 6>(c, partials, gridDim.x, blockIdx.x);
}

// gamma = tau^{-1} gamma', omega = gamma(l), gamma'' = T S gamma
// (:144-162), by one thread
template <typename T>
__global__ __launch_bounds__(64) void k_bl_solve(KState<T> *st, const T *__restrict__ tau, T *__restrict__ gamma,
                                                 const T *__restrict__ gp, T *__restrict__ gt, int l)
{
    if (st->skip || threadIdx.x != 0) return;
    const int ld = l + 1;
    gamma[l] = gp[l];
    st->bl_omega = gamma[l];
    for (int j = l - 1; j >= 1; j--) {
        T sigma = (T)0;
        for (int i = j + 1; i <= l; i++) sigma += tau[j * ld + i] * gamma[i];
        gamma[j] = gp[j] - sigma;
    }
    for (int j = 1; j <= l - 1; j++) {
        T sigma = (T)0;
        for (int i = j + 1; i <= l - 1; i++) sigma += tau[j * ld + i] * gamma[i + 1];
        gt[j] = gamma[j + 1] + sigma;
    }
}

// x += gamma_1 r_0; r_0 -= gamma'_l r_l; u_0 -= gamma_l u_l; then for j = 1
// .. l-1: u_0 -= gamma_j u_j, x += gamma''_j r_j, r_0 -= gamma'_j r_j
// (:165-174); sums |r_0|^2 and (r_0, r0~) for the next loop head
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bl_up(int64_t n, T *__restrict__ rb, T *__restrict__ ub,
                                                  int64_t stride, int l, const T *__restrict__ gamma,
                                                  const T *__restrict__ gp, const T *__restrict__ gt,
                                                  const T *__restrict__ rt, T *__restrict__ x,
                                                  const KState<T> *st, T *partials)
{
    if (st->skip) return;
    const T g1 = gamma[1], gpl = -gp[l], gl = -gamma[l];
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(k, n) {
        T r0 = rb[k], u0 = ub[k], xv = x[k];
        if (g1 != (T)0) xv += g1 * r0;
        if (gpl != (T)0) r0 += gpl * rb[(size_t)l * stride + k];
        if (gl != (T)0) u0 += gl * ub[(size_t)l * stride + k];
        for (int j = 1; j <= l - 1; j++) {
            const T a = -gamma[j], bt = gt[j], cp = -gp[j];
            if (a != (T)0) u0 += a * ub[(size_t)j * stride + k];
            if (bt != (T)0) xv += bt * rb[(size_t)j * stride + k];
            if (cp != (T)0) r0 += cp * rb[(size_t)j * stride + k];
        }
        rb[k] = r0;
        ub[k] = u0;
        x[k] = xv;
        c[0] += r0 * r0;
        c[1] += r0 * rt[k];
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// the change of coordinates of the right preconditioning (:177): x =
// M^{-1} x; not after a failed Init(b), which returns before it
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bl_final(int64_t n, const T *__restrict__ dg, T *__restrict__ x,
                                                     const KState<T> *st)
{
    if (st->init_failed) return;
    GRID_STRIDE(i, n) x[i] = x[i] / dg[i];
}

// ------------------------------------------------------------- Gmres ----
// x = 0 (Gmres.cxx:86); r = M^{-1} b (:89); sum r^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_init0(int64_t n, const T *__restrict__ b, const T *__restrict__ d, T *__restrict__ x, T *__restrict__ r,
                                                     KState<T> *st, T *partials)
{
    if (st->skip) return;
    T c[2] = {(T)0, (T)0};
    GRID_STRIDE(i, n) {
        x[i] = (T)0;
        T v = d ? b[i] / d[i] : b[i];
        r[i] = v;
        c[0] += v * v;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// V_0 = r * (1/beta)  (:105-106)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_v0(int64_t n, const T *__restrict__ r, T *__restrict__ v0, const KState<T> *st)
{
    if (st->cyc_skip) return;
    const T sc = st->gm_scale;
    GRID_STRIDE(i, n) v0[i] = r[i] * sc;
}

// H(k, i) from the previous kernel's partials of (V_k, w) (the finish folded
// into this kernel: every work-group forms the same fixed-order total, the
// first one stores it for the rotations and the least squares).
template <typename T>
__device__ __forceinline__ T gm_fold_h(const T *prev, int prev_nb, int hpos, KState<T> *st)
{
    __shared__ double h_sh;
    T tot[2];
    grid_total<T>(prev, prev_nb, tot);
    if (threadIdx.x == 0) {
        const double h = (double)tot[0];
        h_sh = h;
        if (blockIdx.x == 0) st->H[hpos] = h;
    }
    __syncthreads();
    return (T)(-h_sh);
}

// MGS step k >= 1 of Arnoldi step i: H(k-1,i) from the previous partials;
// w += (-H(k-1,i)) V_{k-1}; partials of (V_k, w) for H(k,i)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_mgs(int64_t n, const T *__restrict__ vprev, const T *__restrict__ vk, T *__restrict__ w,
                                                   int hpos, KState<T> *st, const T *prev, int prev_nb, T *partials)
{
    if (st->skip) return;
    const int64_t j0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T wp[kPre], vp[kPre], kp[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t j = j0 + u * stride;
        if (j < n) {
            wp[u] = w[j];
            vp[u] = vprev[j];
            kp[u] = vk[j];
        }
    }
    const T a = gm_fold_h<T>(prev, prev_nb, hpos, st);
    const bool aa = a != (T)0;
    T c[2] = {(T)0, (T)0};
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t j = j0 + u * stride;
        if (j < n) {
            T wv = wp[u];
            if (aa) {
                wv += a * vp[u];
                w[j] = wv;
            }
            c[0] += kp[u] * wv;
        }
    }
    for (int64_t j = j0 + kPre * stride; j < n; j += stride) {
        T wv = w[j];
        if (aa) {
            wv += a * vprev[j];
            w[j] = wv;
        }
        c[0] += vk[j] * wv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// The same step with 16-byte accesses: thread t takes the N = 16 / sizeof(T)
// consecutive elements N t .. N t + N - 1 of a grid stride (n a multiple of
// N, vectors 16-byte aligned; the host checks).  Each element gets the same
// arithmetic; a thread's partial adds its group's N products in element
// order.
template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
    using type = double2;
};
template <>
struct Vec16<float> {
    using type = float4;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_mgs2(int64_t n, const T *__restrict__ vprev, const T *__restrict__ vk,
                                                    T *__restrict__ w, int hpos, KState<T> *st, const T *prev,
                                                    int prev_nb, T *partials)
{
    using V = typename Vec16<T>::type;
    constexpr int N = 16 / (int)sizeof(T);
    if (st->skip) return;
    const int64_t nv = n / N;
    const int64_t j0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    const V *vpv = (const V *)vprev, *vkv = (const V *)vk;
    V *wv = (V *)w;
    constexpr int U = kPre / 2;
    V wp[U], vp[U], kp[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t j = j0 + u * stride;
        if (j < nv) {
            wp[u] = wv[j];
            vp[u] = vpv[j];
            kp[u] = vkv[j];
        }
    }
    const T a = gm_fold_h<T>(prev, prev_nb, hpos, st);
    const bool aa = a != (T)0;
    T c[2] = {(T)0, (T)0};
    auto step = [&](V &x, const V &v, const V &k, int64_t j) __attribute__((always_inline)) {
        T *xe = (T *)&x;
        const T *ve = (const T *)&v, *ke = (const T *)&k;
        if (aa) {
#pragma unroll
            for (int e = 0; e < N; e++) xe[e] += a * ve[e];
            wv[j] = x;
        }
#pragma unroll
        for (int e = 0; e < N; e++) c[0] += ke[e] * xe[e];
    };
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t j = j0 + u * stride;
        if (j < nv) step(wp[u], vp[u], kp[u], j);
    }
    for (int64_t j = j0 + U * stride; j < nv; j += stride) {
        V x = wv[j];
        const V v = vpv[j], k = vkv[j];
        step(x, v, k, j);
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// H(i,i) from the previous partials; last MGS update w += (-H(i,i)) V_i,
// then the partials of |w|^2 for hi_ip1 (:131-135)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_last(int64_t n, const T *__restrict__ vi, T *__restrict__ w, int hpos, KState<T> *st,
                                                    const T *prev, int prev_nb, T *partials)
{
    if (st->skip) return;
    const int64_t j0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    T wp[kPre], vp[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t j = j0 + u * stride;
        if (j < n) {
            wp[u] = w[j];
            vp[u] = vi[j];
        }
    }
    const T a = gm_fold_h<T>(prev, prev_nb, hpos, st);
    const bool aa = a != (T)0;
    T c[2] = {(T)0, (T)0};
#pragma unroll
    for (int u = 0; u < kPre; u++) {
        const int64_t j = j0 + u * stride;
        if (j < n) {
            T wv = wp[u];
            if (aa) {
                wv += a * vp[u];
                w[j] = wv;
            }
            c[0] += wv * wv;
        }
    }
    for (int64_t j = j0 + kPre * stride; j < n; j += stride) {
        T wv = w[j];
        if (aa) {
            wv += a * vi[j];
            w[j] = wv;
        }
        c[0] += wv * wv;
    }
    block_partial<T, 2>(c, partials, gridDim.x, blockIdx.x);
}

// V_{i+1} = w, scaled by 1/hi when hi != 0 (:136-140); runs before the
// cycle-done test of the step has any effect (V_{i+1} is unused after it)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_vnext(int64_t n, const T *__restrict__ w, T *__restrict__ vnext, const KState<T> *st)
{
    if (st->skip) return;
    const T sc = st->gm_scale;
    GRID_STRIDE(j, n) vnext[j] = w[j] * sc;
}

// H.Resize(i,i); Solve(H, s): back substitution (:165-166, the dtrsv loop)
template <typename T>
__global__ void k_gm_solve(KState<T> *st)
{
    if (st->cyc_skip || threadIdx.x != 0) return;
    const int i = st->gm_i, ld = st->gm_m + 1;
    double *H = st->H, *s = st->gs;
    for (int j = i - 1; j >= 0; j--) {
        if (s[j] != 0.0) {
            s[j] = s[j] / H[j * ld + j];
            double temp = s[j];
            for (int q = j - 1; q >= 0; q--) s[q] -= temp * H[j * ld + q];
        }
    }
}

// x += sum_k s(k) V_k, Adds in k order per element (:169-170)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_xupd(int64_t n, const T *__restrict__ V, T *__restrict__ x, const KState<T> *st)
{
    if (st->cyc_skip) return;
    __shared__ T coef[1024];
    const int i = st->gm_i;
    for (int k = threadIdx.x; k < i; k += kBlock) coef[k] = (T)st->gs[k];
    __syncthreads();
    GRID_STRIDE(j, n) {
        T xv = x[j];
        for (int k = 0; k < i; k++) {
            const T a = coef[k];
            if (a != (T)0) xv += a * V[(size_t)k * n + j];
        }
        x[j] = xv;
    }
}

#undef GRID_STRIDE

// -------------------------------------------------------------- host ----

template <typename T>
struct Solve {
    shk_ctx ctx;
    shk_csr A;
    int64_t n;
    int nb;                 // reduction grid
    KState<T> *dst;         // device state
    KState<T> *hst;         // pinned mirror
    double *dhist = nullptr;
    int *dhist_it = nullptr;
    T *partials;
    T *part[3];             // partial buffers: two alternating pairs (Gmres's folded finishes) and one of
                            // six values (QCgs's smoothing step)
    char *ws;
    size_t off = 0;
    // caller-provided work memory (shk_cg_noalloc): vectors, history and
    // partials all come from [ext, ext + ext_cap), nothing is allocated
    char *ext = nullptr;
    size_t ext_cap = 0;

    T *take(int64_t count)
    {
        T *p = (T *)(ws + off);
        off += ((size_t)count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }

    int setup(shk_ctx c, shk_csr a, size_t vec_count, size_t extra_bytes, const shk_iteration *it)
    {
        ctx = c;
        A = a;
        n = a->m;
        nb = red_blocks(n, SHK_SOLVE_ITEMS);
        // a small system gets more work-groups (up to one per CU, down to two
        // elements per thread): its kernels are latency-bound, not streaming
        if (nb < kSolveMinBlocks)
            nb = (int)std::min<int64_t>(kSolveMinBlocks, std::max<int64_t>(nb, (n + 2 * kBlock - 1) / (2 * kBlock)));
        size_t vb = (((size_t)n * sizeof(T) + 255) & ~(size_t)255);
        size_t hist_cap = it->hist && it->hist_cap > 0 ? (size_t)it->hist_cap : 0;
        size_t need = vec_count * vb + extra_bytes + 256 + hist_cap * (sizeof(double) + sizeof(int)) + 512;
        const size_t maxb = (size_t)std::max(nb, a->nblk_launch);
        const size_t red = sizeof(T) * 10 * maxb;  // two buffers of [2][maxb], one of [6][maxb]
        if (ext) {
            need = (need + 255) & ~(size_t)255;
            SHK_CHECK(need + red <= ext_cap, SHK_ERR_ARG,
                      "shk_cg_noalloc: the system needs " + std::to_string(need + red) +
                          " bytes of work memory, more than the " + std::to_string(ext_cap) +
                          " bytes of the library's fixed scratch (at most 32 MiB: about 1M FP64 or 2M FP32 rows)");
            ws = ext;
            partials = (T *)(ext + need);
        } else {
            SHK_TRY(ensure_ws(ctx, need));
            SHK_TRY(ensure_red(ctx, red));
            ws = (char *)ctx->ws;
            partials = (T *)ctx->red.partials;
        }
        off = 0;
        part[0] = partials;
        part[1] = partials + 2 * maxb;
        part[2] = partials + 4 * maxb;
        dst = (KState<T> *)ctx->state;
        hst = (KState<T> *)ctx->state_host;
        if (hist_cap) {
            dhist = (double *)(ws + off);
            off += (hist_cap * sizeof(double) + 255) & ~(size_t)255;
            dhist_it = (int *)(ws + off);
            off += (hist_cap * sizeof(int) + 255) & ~(size_t)255;
        }
        KState<T> s;
        std::memset(&s, 0, sizeof(s));
        s.tol = it->tolerance;
        s.facteur = it->facteur_reste;
        s.max_iter = it->max_iter;
        s.nb_iter = it->nb_iter;
        s.hist = dhist;
        s.hist_it = dhist_it;
        s.hist_cap = (int)hist_cap;
        s.stop_slot = -1;
        s.cyc_skip = 1;
        std::memcpy(hst, &s, sizeof(s));
        SHK_HIP(hipMemcpyAsync(dst, hst, sizeof(s), hipMemcpyHostToDevice, ctx->stream));
        return SHK_OK;
    }

    int read_state()
    {
        SHK_HIP(hipMemcpyAsync(hst, dst, sizeof(KState<T>), hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        return SHK_OK;
    }

    // Copy the outcome into the caller's Iteration mirror; returns ErrorCode().
    int finish(shk_iteration *it, int *code)
    {
        SHK_TRY(read_state());
        const KState<T> &s = *hst;
        if (!s.init_failed) {
            it->facteur_reste = s.facteur;
            it->nb_iter = s.nb_iter;
        }
        it->fail_convergence = s.fail;
        if (s.fail) it->error_code = s.error_code;
        it->hist_len = s.hist_len;
        int ncopy = std::min(s.hist_len, s.hist_cap);
        if (ncopy > 0) {
            SHK_HIP(hipMemcpyAsync(it->hist, dhist, sizeof(double) * ncopy, hipMemcpyDeviceToHost, ctx->stream));
            if (it->hist_it)
                SHK_HIP(hipMemcpyAsync(it->hist_it, dhist_it, sizeof(int) * ncopy, hipMemcpyDeviceToHost,
                                       ctx->stream));
            SHK_HIP(hipStreamSynchronize(ctx->stream));
        }
        // Iteration::ErrorCode (Iterative.cxx:344-354)
        if (it->nb_iter >= it->max_iter) *code = -2;
        else if (s.fail) *code = s.error_code;
        else *code = 0;
        return SHK_OK;
    }

    template <typename Fin>
    int spmv(SpmvArgs<T> &a, Fin fin)
    {
        return launch_spmv<T, Fin>(A, a, fin);
    }

    // The iteration loop: body(stream) enqueues one chunk of iterations on
    // the context's stream and the state is read back after each chunk,
    // until the solve stops or `limit` iterations have been enqueued.
    // (Recording the chunk once into a hipGraph and replaying it was
    // measured: no gain, the per-kernel cost is on the device side.)
    template <class Body>
    int run_chunks(int64_t limit, int chunk, Body body)
    {
        for (int64_t done_its = 0; done_its <= limit; done_its += chunk) {
            SHK_TRY(body(ctx->stream));
            SHK_TRY(read_state());
            if (hst->done) break;
        }
        return SHK_OK;
    }

    // Second half of the split reduction of the kernel just enqueued (which
    // skipped itself, and wrote no partials, when the solve had ended).
    template <typename Fin>
    void fin(Fin f)
    {
        grid_finish<T>(ctx->stream, partials, nb, (const int *)&dst->skip, f);
    }
};

constexpr int kChunk = 16;

template <typename T>
int cg_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code,
            char *ext = nullptr, size_t ext_cap = 0)
{
    Solve<T> S;
    S.ext = ext;
    S.ext_cap = ext_cap;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, jac ? 4 : 3, 0, it));
    const int64_t n = S.n;
    T *p = S.take(n), *q = S.take(n), *r = S.take(n), *z = jac ? S.take(n) : r;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (Cg.cxx:58)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    // r = b - A x or x = 0, r = b; then the first Finished(r)
    FinCgTop<T> top0{S.dst, 0, jac ? 0 : 1, 0};
    if (it->init_guess_null) {
        hipLaunchKernelGGL(k_cg_init0<T>, g, blk, 0, st, n, b, x, r, S.dst, S.partials);
        S.fin(top0);
    } else {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1;
        a.dot0 = kDotSelf; a.skip = &S.dst->skip;
        a.partials = S.partials;
        SHK_TRY(S.spmv(a, top0));
    }
    SHK_HIP(hipGetLastError());
    const T *dg = (const T *)A->diag;
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) + 2;
    // the loop head of iterations 1 .. kChunk - 1 of a chunk is folded into
    // the next kernel (identity: k_cg_p_top; Jacobi: k_cg_z_top, and rho =
    // (r, z) into k_cg_p_rho); the head before a chunk's first iteration runs
    // as its own finish (slot 0), so each chunk ends with the state read back
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            if (jac) {
                if (k == 0)
                    hipLaunchKernelGGL(k_cg_z<T>, g, blk, 0, cs, n, r, dg, z, S.dst, S.part[0]);
                else
                    hipLaunchKernelGGL(k_cg_z_top<T>, g, blk, 0, cs, n, r, dg, z, S.dst, (const T *)S.part[1], S.nb,
                                       S.part[0]);
                hipLaunchKernelGGL(k_cg_p_rho<T>, g, blk, 0, cs, n, z, p, S.dst, (const T *)S.part[0], S.nb);
            } else if (k == 0) {
                hipLaunchKernelGGL(k_cg_p<T>, g, blk, 0, cs, n, z, p, S.dst);
            } else {
                hipLaunchKernelGGL(k_cg_p_top<T>, g, blk, 0, cs, n, z, p, S.dst, (const T *)S.part[1], S.nb,
                                   k & 1);
            }
            SpmvArgs<T> a{};
            a.x = p; a.y_out = q; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = p; a.skip = &S.dst->skip;
            a.partials = S.part[0];
            SHK_TRY(S.spmv(a, FinDefer{}));  // (p, q): finished inside k_cg_xr
            hipLaunchKernelGGL(k_cg_xr<T>, g, blk, 0, cs, n, p, q, x, r, S.dst, (const T *)S.part[0],
                               a.grid_used, S.part[1]);
            if (k == kChunk - 1)
                grid_finish<T>(cs, (const T *)S.part[1], S.nb, (const int *)&S.dst->skip,
                               FinCgTop<T>{S.dst, 1, jac ? 0 : 1, 0});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "cg: solver did not reach a stopping state");
    return S.finish(it, code);
}

template <typename T>
int bicgstab_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, jac ? 8 : 6, 0, it));
    const int64_t n = S.n;
    T *p = S.take(n), *s = S.take(n), *t = S.take(n), *v = S.take(n), *r = S.take(n), *rt = S.take(n);
    T *phat = jac ? S.take(n) : p, *shat = jac ? S.take(n) : s;
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
    }
    hipLaunchKernelGGL(k_bi_init<T>, g, blk, 0, st, n, b, x, r, rt, it->init_guess_null ? 1 : 0,
                       S.dst, S.partials);
    S.fin(FinBiTop<T>{S.dst, 0, 0});
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    // a loop's slot (its place in the chunk) names the Finished(s) that
    // stopped the solve; the chunk is replayed, so slots repeat across
    // chunks, but a stopped solve replays no further chunk
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int slot = 0; slot < kChunk; slot++) {
            // the loop head of slots 1 .. kChunk - 1 is folded into k_bi_p_top;
            // the head before a chunk's first slot runs as its own finish
            if (slot == 0)
                hipLaunchKernelGGL(k_bi_p<T>, g, blk, 0, cs, n, r, v, p, dg, phat, S.dst);
            else
                hipLaunchKernelGGL(k_bi_p_top<T>, g, blk, 0, cs, n, r, v, p, dg, phat, S.dst,
                                   (const T *)S.part[1], S.nb, slot & 1);
            SpmvArgs<T> a{};
            a.x = phat; a.y_out = v; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = rt; a.skip = &S.dst->skip;
            a.partials = S.part[0];
            SHK_TRY(S.spmv(a, FinDefer{}));  // sigma: finished inside k_bi_s
            hipLaunchKernelGGL(k_bi_s<T>, g, blk, 0, cs, n, r, v, s, dg, shat, S.dst, (const T *)S.part[0],
                               a.grid_used, S.part[1]);
            grid_finish<T>(cs, (const T *)S.part[1], S.nb, (const int *)&S.dst->skip, FinBiS<T>{S.dst, slot});
            hipLaunchKernelGGL(k_bi_fix<T>, g, blk, 0, cs, n, phat, x, S.dst, slot);
            SpmvArgs<T> b2{};
            b2.x = shat; b2.y_out = t; b2.alpha = (T)1; b2.mlt = 1;
            b2.dot0 = kDotVec; b2.u0 = s; b2.dot1 = kDotSelf; b2.skip = &S.dst->skip;
            b2.partials = S.part[0];
            SHK_TRY(S.spmv(b2, FinDefer{}));  // omega: finished inside k_bi_xr
            hipLaunchKernelGGL(k_bi_xr<T>, g, blk, 0, cs, n, phat, shat, s, t, rt, x, r, S.dst,
                               (const T *)S.part[0], b2.grid_used, S.part[1]);
            if (slot == kChunk - 1)
                grid_finish<T>(cs, (const T *)S.part[1], S.nb, (const int *)&S.dst->skip,
                               FinBiTop<T>{S.dst, 1, 0});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "bicgstab: solver did not reach a stopping state");
    return S.finish(it, code);
}

template <typename T>
int gmres_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    const int m = it->parameter_restart;
    SHK_CHECK(m >= 1 && m <= 1000, SHK_ERR_ARG, "gmres: restart must be in [1, 1000]");
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    const size_t small = sizeof(double) * ((size_t)(m + 1) * (m + 1) + 3 * (size_t)(m + 1)) + 1024;
    SHK_TRY(S.setup(ctx, A, (size_t)m + 3, small, it));
    const int64_t n = S.n;
    T *V = S.take(n * (int64_t)(m + 1));
    T *w = S.take(n), *r = S.take(n);
    double *H = (double *)S.take((int64_t)((small + sizeof(T) - 1) / sizeof(T)));
    double *gs = H + (size_t)(m + 1) * (m + 1), *gc = gs + (m + 1), *gsn = gc + (m + 1);
    SHK_HIP(hipMemsetAsync(H, 0, small, ctx->stream));
    {
        KState<T> *h = S.hst;
        h->gm_m = m;
        h->H = H;
        h->gs = gs;
        h->gc = gc;
        h->gsn = gsn;
        SHK_HIP(hipMemcpyAsync(S.dst, h, sizeof(KState<T>), hipMemcpyHostToDevice, ctx->stream));
    }
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    auto Vk = [&](int k) { return V + (size_t)k * n; };
    // residual, preconditioning, beta, outer.Init (Gmres.cxx:82-95)
    if (it->init_guess_null) {
        hipLaunchKernelGGL(k_gm_init0<T>, g, blk, 0, st, n, b, dg, x, r, S.dst, S.partials);
        S.fin(FinGmInit<T>{S.dst});
    } else {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.diag = dg;
        a.dot0 = kDotSelf; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinGmInit<T>{S.dst}));
    }
    hipLaunchKernelGGL(k_gm_v0<T>, g, blk, 0, st, n, r, Vk(0), S.dst);
    SHK_HIP(hipGetLastError());
    SHK_TRY(S.read_state());
    const int step_chunk = 4;
    int64_t guard = 0;
    while (!S.hst->done) {
        SHK_CHECK(++guard <= (int64_t)it->max_iter + 2, SHK_ERR_STATE, "gmres: no progress");
        // Arnoldi steps of one cycle, enqueued in chunks
        for (int i0 = 0; i0 < m; i0 += step_chunk) {
            for (int i = i0; i < std::min(m, i0 + step_chunk); i++) {
                SpmvArgs<T> a{};
                a.x = Vk(i); a.y_out = w; a.alpha = (T)1; a.mlt = 1; a.diag = dg;
                a.dot0 = kDotVec; a.u0 = Vk(0); a.skip = &S.dst->skip;
                // each dot's finish is folded into the next kernel; the two
                // partial buffers alternate so a kernel never overwrites the
                // partials it is still reading
                a.partials = S.part[0];
                SHK_TRY(S.spmv(a, FinDefer{}));
                int prev_nb = a.grid_used;
                for (int k = 1; k <= i; k++) {
                    // 16-byte accesses where the vectors allow them
                    const bool pairs = n % (16 / (int64_t)sizeof(T)) == 0 &&
                                       (((uintptr_t)Vk(k - 1) | (uintptr_t)Vk(k) | (uintptr_t)w) & 15) == 0;
                    if (pairs)
                        hipLaunchKernelGGL(k_gm_mgs2<T>, g, blk, 0, st, n, (const T *)Vk(k - 1), (const T *)Vk(k), w,
                                           i * (m + 1) + (k - 1), S.dst, (const T *)S.part[(k - 1) & 1], prev_nb,
                                           S.part[k & 1]);
                    else
                        hipLaunchKernelGGL(k_gm_mgs<T>, g, blk, 0, st, n, (const T *)Vk(k - 1), (const T *)Vk(k), w,
                                           i * (m + 1) + (k - 1), S.dst, (const T *)S.part[(k - 1) & 1], prev_nb,
                                           S.part[k & 1]);
                    prev_nb = S.nb;
                }
                hipLaunchKernelGGL(k_gm_last<T>, g, blk, 0, st, n, (const T *)Vk(i), w, i * (m + 1) + i, S.dst,
                                   (const T *)S.part[i & 1], prev_nb, S.part[(i + 1) & 1]);
                grid_finish<T>(st, (const T *)S.part[(i + 1) & 1], S.nb, (const int *)&S.dst->skip,
                               FinGmStep<T>{S.dst, i});
                hipLaunchKernelGGL(k_gm_vnext<T>, g, blk, 0, st, n, (const T *)w, Vk(i + 1), S.dst);
            }
            SHK_HIP(hipGetLastError());
            SHK_TRY(S.read_state());
            if (S.hst->gm_cycle_done || S.hst->done) break;
        }
        // cycle end: least squares, update, true residual (Gmres.cxx:165-178)
        hipLaunchKernelGGL(k_gm_solve<T>, dim3(1), dim3(64), 0, st, S.dst);
        hipLaunchKernelGGL(k_gm_xupd<T>, g, blk, 0, st, n, (const T *)V, x, S.dst);
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.diag = dg;
        a.dot0 = kDotSelf; a.skip = &S.dst->cyc_skip;
        SHK_TRY(S.spmv(a, FinGmOuter<T>{S.dst}));
        hipLaunchKernelGGL(k_gm_v0<T>, g, blk, 0, st, n, r, Vk(0), S.dst);
        SHK_HIP(hipGetLastError());
        SHK_TRY(S.read_state());
    }
    return S.finish(it, code);
}

// Cgs (Cgs.cxx:46-137): per loop (two of Seldon's iterations) the p / u
// update, vhat = A phat with (rtilde, vhat), the q / u / x update,
// qhat = A uhat, the r update with |r|^2 and (rtilde, r), the loop head.
template <typename T>
int cgs_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, jac ? 9 : 7, 0, it));
    const int64_t n = S.n;
    T *p = S.take(n), *q = S.take(n), *u = S.take(n), *vhat = S.take(n), *qhat = S.take(n), *r = S.take(n),
      *rt = S.take(n);
    T *phat = jac ? S.take(n) : p, *uhat = jac ? S.take(n) : u;
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (Cgs.cxx:58)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
    }
    hipLaunchKernelGGL(k_lz_init<T>, g, blk, 0, st, n, b, x, r, rt, (const T *)nullptr, (T *)nullptr,
                       (T *)nullptr, it->init_guess_null ? 1 : 0, S.dst, S.partials);
    S.fin(FinLanczosTop<T>{S.dst, 0});
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            hipLaunchKernelGGL(k_cgs_pu<T>, g, blk, 0, cs, n, r, q, p, u, dg, phat, S.dst);
            SpmvArgs<T> a{};
            a.x = phat; a.y_out = vhat; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = rt; a.skip = &S.dst->skip;
            a.partials = S.part[0];
            SHK_TRY(S.spmv(a, FinDefer{}));  // delta: finished inside k_cgs_qux
            hipLaunchKernelGGL(k_cgs_qux<T>, g, blk, 0, cs, n, vhat, q, u, dg, uhat, x, S.dst,
                               (const T *)S.part[0], a.grid_used);
            SpmvArgs<T> a2{};
            a2.x = uhat; a2.y_out = qhat; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
            SHK_TRY(S.spmv(a2, FinNone{}));
            hipLaunchKernelGGL(k_cgs_r<T>, g, blk, 0, cs, n, qhat, rt, r, S.dst, S.partials);
            S.fin(FinLanczosTop<T>{S.dst, 1});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "cgs: solver did not reach a stopping state");
    return S.finish(it, code);
}

// BiCg (BiCg.cxx:47-131): per loop the p / p~ update, q = A p with (p~, q),
// q~ = A^T p~ (the library's device-built transpose, Seldon's order), the
// x / r / r~ update with z, z~, |r|^2 and (z, r~), the loop head.
template <typename T>
int bicg_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    shk_csr At = nullptr;
    SHK_TRY(transpose_of(A, &At));
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, jac ? 8 : 6, 0, it));
    const int64_t n = S.n;
    T *p = S.take(n), *pt = S.take(n), *q = S.take(n), *qt = S.take(n), *r = S.take(n), *rt = S.take(n);
    T *z = jac ? S.take(n) : r, *zt = jac ? S.take(n) : rt;
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
    }
    hipLaunchKernelGGL(k_lz_init<T>, g, blk, 0, st, n, b, x, r, rt, dg, z, zt, it->init_guess_null ? 1 : 0,
                       S.dst, S.partials);
    S.fin(FinLanczosTop<T>{S.dst, 0});
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            hipLaunchKernelGGL(k_bicg_p<T>, g, blk, 0, cs, n, z, zt, p, pt, S.dst);
            SpmvArgs<T> a{};
            a.x = p; a.y_out = q; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = pt; a.skip = &S.dst->skip;
            a.partials = S.part[0];
            SHK_TRY(S.spmv(a, FinDefer{}));  // delta: finished inside k_bicg_xr
            SpmvArgs<T> a2{};
            a2.x = pt; a2.y_out = qt; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
            SHK_TRY((launch_spmv<T, FinNone>(At, a2, FinNone{})));
            hipLaunchKernelGGL(k_bicg_xr<T>, g, blk, 0, cs, n, p, q, qt, x, r, rt, dg, z, zt, S.dst,
                               (const T *)S.part[0], a.grid_used, S.partials);
            S.fin(FinLanczosTop<T>{S.dst, 1});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "bicg: solver did not reach a stopping state");
    return S.finish(it, code);
}

// TfQmr (TfQmr.cxx:45-240): per loop (Seldon's two iterations) sigma and
// alpha, y0, h = M^{-1} A y0 with the first half step's vector updates,
// its scalars and Finished, the second half step, its scalars, Finished
// and beta, y1, g = M^{-1} A y1 and v.
template <typename T>
int tfqmr_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 9, 0, it));
    const int64_t n = S.n;
    T *tmp = S.take(n), *v = S.take(n), *h = S.take(n), *w = S.take(n), *y1 = S.take(n), *g = S.take(n),
      *y0 = S.take(n), *rt = S.take(n), *d = S.take(n);
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g_(S.nb), blk(kBlock);
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = tmp; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = tmp;
    }
    hipLaunchKernelGGL(k_tq_init<T>, g_, blk, 0, st, n, src, dg, x, w, y1, rt, d, it->init_guess_null ? 1 : 0,
                       S.dst, S.partials);
    S.fin(FinTqInit<T>{S.dst});
    {
        SpmvArgs<T> a{};
        a.x = y1; a.y_out = tmp; a.alpha = (T)1; a.mlt = 1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
    }
    hipLaunchKernelGGL(k_tq_v<T>, g_, blk, 0, st, n, tmp, dg, g, v, (const T *)h, rt, 1, S.dst, S.partials);
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int slot = 0; slot < kChunk; slot++) {
            S.fin(FinTqSigma<T>{S.dst});
            hipLaunchKernelGGL(k_tq_y0<T>, g_, blk, 0, cs, n, y1, v, y0, S.dst);
            SpmvArgs<T> a{};
            a.x = y0; a.y_out = tmp; a.alpha = (T)1; a.mlt = 1; a.skip = &S.dst->skip;
            SHK_TRY(S.spmv(a, FinNone{}));
            hipLaunchKernelGGL(k_tq_h1<T>, g_, blk, 0, cs, n, tmp, dg, h, g, w, d, y1, S.dst, S.partials);
            S.fin(FinTqHalf<T>{S.dst, 0, slot});
            hipLaunchKernelGGL(k_tq_h2<T>, g_, blk, 0, cs, n, h, g, w, d, y0, x, rt, S.dst, slot, S.partials);
            S.fin(FinTqHalf<T>{S.dst, 1, slot});
            hipLaunchKernelGGL(k_tq_y1<T>, g_, blk, 0, cs, n, w, y0, d, x, y1, S.dst, slot);
            SpmvArgs<T> a2{};
            a2.x = y1; a2.y_out = tmp; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
            SHK_TRY(S.spmv(a2, FinNone{}));
            hipLaunchKernelGGL(k_tq_v<T>, g_, blk, 0, cs, n, tmp, dg, g, v, (const T *)h, rt, 0, S.dst,
                               S.partials);
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "tfqmr: solver did not reach a stopping state");
    return S.finish(it, code);
}

// Qmr (Qmr.cxx:42-190): per loop (Seldon's two iterations) the scaling of
// v, y, w with (w, y), the p / q update, p~ = A p with (q, p~), v and y,
// z~ = A^T q (the device-built transpose) with w, |y| and |w|, the scalar
// step, and the d / s / x / r update with |r| for the loop head.
template <typename T>
int qmr_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    shk_csr At = nullptr;
    SHK_TRY(transpose_of(A, &At));
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 10, 0, it));
    const int64_t n = S.n;
    T *r = S.take(n), *y = S.take(n), *z = S.take(n), *v = S.take(n), *w = S.take(n), *pt = S.take(n),
      *p = S.take(n), *q = S.take(n), *d = S.take(n), *s = S.take(n);
    const T *dg = jac ? (const T *)A->diag : nullptr;
    const T *zsrc = jac ? z : w;  // M^{-T} w: the identity's is w itself
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (Qmr.cxx:60)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = r;
    }
    hipLaunchKernelGGL(k_qmr_init<T>, g, blk, 0, st, n, src, b, dg, x, r, v, y, w, d, s,
                       it->init_guess_null ? 1 : 0, S.dst, S.partials);
    S.fin(FinQmrTop<T>{S.dst, 1});
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            hipLaunchKernelGGL(k_qmr_scale<T>, g, blk, 0, cs, n, v, y, w, dg, z, S.dst, S.partials);
            S.fin(FinQmrDelta<T>{S.dst});
            hipLaunchKernelGGL(k_qmr_pq<T>, g, blk, 0, cs, n, y, zsrc, p, q, S.dst);
            SpmvArgs<T> a{};
            a.x = p; a.y_out = pt; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = q; a.skip = &S.dst->skip;
            a.partials = S.part[1];
            SHK_TRY(S.spmv(a, FinDefer{}));
            grid_finish<T>(cs, (const T *)S.part[1], a.grid_used, (const int *)&S.dst->skip, FinQmrEp<T>{S.dst});
            hipLaunchKernelGGL(k_qmr_vy<T>, g, blk, 0, cs, n, pt, dg, v, y, S.dst);
            SpmvArgs<T> a2{};
            a2.x = q; a2.y_out = z; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
            SHK_TRY((launch_spmv<T, FinNone>(At, a2, FinNone{})));
            hipLaunchKernelGGL(k_qmr_w<T>, g, blk, 0, cs, n, (const T *)z, (const T *)y, w, S.dst, S.partials);
            S.fin(FinQmrTheta<T>{S.dst});
            hipLaunchKernelGGL(k_qmr_dsxr<T>, g, blk, 0, cs, n, p, pt, d, s, x, r, S.dst, S.partials);
            S.fin(FinQmrTop<T>{S.dst, 0});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "qmr: solver did not reach a stopping state");
    return S.finish(it, code);
}

// Cgne (Cgne.cxx:41-124): Cg on A^T A x = A^T b; per loop z = M^{-1}
// M^{-T} r with (r, z), the p update, q = A p, z = A^T q with (p, z)
// (the device-built transpose), the x / r update with |r|.
template <typename T>
int cgne_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    shk_csr At = nullptr;
    SHK_TRY(transpose_of(A, &At));
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 4, 0, it));
    const int64_t n = S.n;
    T *p = S.take(n), *q = S.take(n), *r = S.take(n), *z = S.take(n);
    const T *dg = jac ? (const T *)A->diag : nullptr;
    const T *zsrc = jac ? z : r;  // the identity's M^{-1} M^{-T} r is r itself
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    SpmvArgs<T> a0{};
    a0.x = b; a0.y_out = q; a0.alpha = (T)1; a0.mlt = 1; a0.skip = &S.dst->skip;
    SHK_TRY((launch_spmv<T, FinNone>(At, a0, FinNone{})));
    hipLaunchKernelGGL(k_sumsq<T>, g, blk, 0, st, n, (const T *)q, S.dst, S.partials);
    S.fin(FinCgneInit<T>{S.dst});
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_out = p; a.alpha = (T)1; a.mlt = 1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        SpmvArgs<T> a2{};
        a2.x = p; a2.y_out = r; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
        SHK_TRY((launch_spmv<T, FinNone>(At, a2, FinNone{})));
    }
    hipLaunchKernelGGL(k_cgne_r0<T>, g, blk, 0, st, n, (const T *)q, r, x, it->init_guess_null ? 1 : 0, S.dst);
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            hipLaunchKernelGGL(k_cgne_z<T>, g, blk, 0, cs, n, (const T *)r, dg, z, S.dst, S.partials);
            S.fin(FinCgneRho<T>{S.dst});
            hipLaunchKernelGGL(k_cgne_p<T>, g, blk, 0, cs, n, zsrc, p, S.dst);
            SpmvArgs<T> a{};
            a.x = p; a.y_out = q; a.alpha = (T)1; a.mlt = 1; a.skip = &S.dst->skip;
            SHK_TRY(S.spmv(a, FinNone{}));
            SpmvArgs<T> a2{};
            a2.x = q; a2.y_out = z; a2.alpha = (T)1; a2.mlt = 1;
            a2.dot0 = kDotVec; a2.u0 = p; a2.skip = &S.dst->skip;
            a2.partials = S.part[1];
            SHK_TRY((launch_spmv<T, FinDefer>(At, a2, FinDefer{})));
            grid_finish<T>(cs, (const T *)S.part[1], a2.grid_used, (const int *)&S.dst->skip,
                           FinCgneDelta<T>{S.dst});
            hipLaunchKernelGGL(k_cgne_xr<T>, g, blk, 0, cs, n, (const T *)p, (const T *)z, x, r, S.dst,
                               S.partials);
            S.fin(FinCgneTop<T>{S.dst});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "cgne: solver did not reach a stopping state");
    return S.finish(it, code);
}

// Lsqr (Lsqr.cxx:41-150): per loop u1 = A v, u1 -= alpha u with |u1|, the
// scaling, v1 = A^T u1 (the device-built transpose), v1 -= beta v with
// |v1|, the rotation's scalars, and the x / w update; u / u1 and v / v1
// swap roles from loop to loop (Swap).  The template applies no
// preconditioner, so none is applied here either.
template <typename T>
int lsqr_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, shk_iteration *it, int *code)
{
    shk_csr At = nullptr;
    SHK_TRY(transpose_of(A, &At));
    Solve<T> S;
    SHK_TRY(S.setup(ctx, A, 5, 0, it));
    const int64_t n = S.n;
    T *u = S.take(n), *u1 = S.take(n), *v = S.take(n), *v1 = S.take(n), *w = S.take(n);
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (Lsqr.cxx:55)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = u; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = u;
    }
    hipLaunchKernelGGL(k_lsqr_norms<T>, g, blk, 0, st, n, src, b, S.dst, S.partials);
    S.fin(FinLsqrNorms<T>{S.dst});
    hipLaunchKernelGGL(k_lsqr_u0<T>, g, blk, 0, st, n, b, u, x, S.dst);
    {
        SpmvArgs<T> a{};
        a.x = u; a.y_out = v; a.alpha = (T)1; a.mlt = 1; a.dot0 = kDotSelf; a.skip = &S.dst->skip;
        a.partials = S.part[1];
        SHK_TRY((launch_spmv<T, FinDefer>(At, a, FinDefer{})));
        grid_finish<T>(st, (const T *)S.part[1], a.grid_used, (const int *)&S.dst->skip, FinLsqrStart<T>{S.dst});
    }
    hipLaunchKernelGGL(k_lsqr_v0<T>, g, blk, 0, st, n, v, w, S.dst);
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            // Swap(u1, u); Swap(v1, v) after every loop: even slots use the
            // first assignment, odd ones the other (a chunk has an even count)
            T *uc = (k & 1) ? u1 : u, *un = (k & 1) ? u : u1;
            T *vc = (k & 1) ? v1 : v, *vn = (k & 1) ? v : v1;
            SpmvArgs<T> a{};
            a.x = vc; a.y_out = un; a.alpha = (T)1; a.mlt = 1; a.skip = &S.dst->skip;
            SHK_TRY(S.spmv(a, FinNone{}));
            hipLaunchKernelGGL(k_lsqr_sub<T>, g, blk, 0, cs, n, (const T *)uc, un, 0, S.dst, S.partials);
            S.fin(FinLsqrBeta<T>{S.dst});
            hipLaunchKernelGGL(k_lsqr_scale<T>, g, blk, 0, cs, n, un, S.dst);
            SpmvArgs<T> a2{};
            a2.x = un; a2.y_out = vn; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
            SHK_TRY((launch_spmv<T, FinNone>(At, a2, FinNone{})));
            hipLaunchKernelGGL(k_lsqr_sub<T>, g, blk, 0, cs, n, (const T *)vc, vn, 1, S.dst, S.partials);
            S.fin(FinLsqrAlpha<T>{S.dst, k});
            hipLaunchKernelGGL(k_lsqr_xw<T>, g, blk, 0, cs, n, vn, w, x, S.dst, k);
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "lsqr: solver did not reach a stopping state");
    return S.finish(it, code);
}

// BiCgcr (BiCgcr.cxx:45-117): per iteration rho = (w, v) and mu = (w, a),
// the x / v / s update with |v|, z = A s with (w, z), and the p / a / w
// update with the next (w, v), (w, a); one product per iteration.
template <typename T>
int bicgcr_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, jac ? 6 : 5, 0, it));
    const int64_t n = S.n;
    T *v = S.take(n), *s = S.take(n), *z = S.take(n), *p = S.take(n), *a = S.take(n);
    T *w = jac ? S.take(n) : a;  // the identity's M^{-1} a is a itself
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (BiCgcr.cxx:58)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a0{};
        a0.x = x; a0.y_in = b; a0.y_out = v; a0.alpha = (T)-1; a0.beta = (T)1; a0.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a0, FinNone{}));
        src = v;
    }
    hipLaunchKernelGGL(k_bcr_init<T>, g, blk, 0, st, n, src, dg, x, v, s, p, it->init_guess_null ? 1 : 0, S.dst,
                       S.part[0]);
    {
        SpmvArgs<T> a1{};
        a1.x = p; a1.y_out = a; a1.alpha = (T)1; a1.mlt = 1; a1.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a1, FinNone{}));
    }
    hipLaunchKernelGGL(k_bcr_paw<T>, g, blk, 0, st, n, (const T *)s, (const T *)z, dg, (const T *)v, p, a, w, 1,
                       S.dst, S.part[1]);
    grid_finish<T>(st, (const T *)S.part[0], S.nb, (const int *)&S.dst->skip, FinBcrTop<T>{S.dst, 1});
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            grid_finish<T>(cs, (const T *)S.part[1], S.nb, (const int *)&S.dst->skip, FinBcrMu<T>{S.dst});
            hipLaunchKernelGGL(k_bcr_xvs<T>, g, blk, 0, cs, n, (const T *)p, (const T *)a, (const T *)w, x, v, s,
                               S.dst, S.part[0]);
            SpmvArgs<T> a2{};
            a2.x = s; a2.y_out = z; a2.alpha = (T)1; a2.mlt = 1;
            a2.dot0 = kDotVec; a2.u0 = w; a2.skip = &S.dst->skip;
            a2.partials = S.part[1];
            SHK_TRY(S.spmv(a2, FinDefer{}));
            grid_finish<T>(cs, (const T *)S.part[1], a2.grid_used, (const int *)&S.dst->skip, FinBcrTau<T>{S.dst});
            hipLaunchKernelGGL(k_bcr_paw<T>, g, blk, 0, cs, n, (const T *)s, (const T *)z, dg, (const T *)v, p, a, w,
                               0, S.dst, S.part[1]);
            grid_finish<T>(cs, (const T *)S.part[0], S.nb, (const int *)&S.dst->skip, FinBcrTop<T>{S.dst, 0});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "bicgcr: solver did not reach a stopping state");
    return S.finish(it, code);
}

// QmrSym (QmrSym.cxx:42-166): Qmr without the transposed side; per loop
// the v / y scaling with (v, y), the p update, p~ = A p with (p, p~), v and
// y with |y|, the scalar step, the d / s / x / r update with |r|.
template <typename T>
int qmrsym_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 7, 0, it));
    const int64_t n = S.n;
    T *r = S.take(n), *y = S.take(n), *v = S.take(n), *pt = S.take(n), *p = S.take(n), *d = S.take(n),
      *s = S.take(n);
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (QmrSym.cxx:59)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = r;
    }
    hipLaunchKernelGGL(k_qms_init<T>, g, blk, 0, st, n, src, dg, x, r, v, y, it->init_guess_null ? 1 : 0, S.dst,
                       S.partials);
    S.fin(FinQmsTop<T>{S.dst, 1});
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            hipLaunchKernelGGL(k_qms_scale<T>, g, blk, 0, cs, n, v, y, S.dst, S.partials);
            S.fin(FinQmsDelta<T>{S.dst});
            hipLaunchKernelGGL(k_qms_p<T>, g, blk, 0, cs, n, (const T *)y, p, S.dst);
            SpmvArgs<T> a{};
            a.x = p; a.y_out = pt; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = p; a.skip = &S.dst->skip;
            a.partials = S.part[1];
            SHK_TRY(S.spmv(a, FinDefer{}));
            grid_finish<T>(cs, (const T *)S.part[1], a.grid_used, (const int *)&S.dst->skip, FinQmsEp<T>{S.dst});
            hipLaunchKernelGGL(k_qms_vy<T>, g, blk, 0, cs, n, (const T *)pt, dg, v, y, S.dst, S.partials);
            S.fin(FinQmsTheta<T>{S.dst});
            hipLaunchKernelGGL(k_qms_dsxr<T>, g, blk, 0, cs, n, (const T *)p, (const T *)pt, d, s, x, r, S.dst,
                               S.partials);
            S.fin(FinQmsTop<T>{S.dst, 0});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "qmrsym: solver did not reach a stopping state");
    return S.finish(it, code);
}

// MinRes (MinRes.cxx:47-151): per iteration r = A u with (r, u), z and the
// r / z recurrences with (r, z), the QR step, and the w / x / v / u update;
// the w triple and the v, u pairs rotate through their buffers instead of
// being copied (Copy(w_old, w_oold); Copy(w, w_old); ...).
template <typename T>
int minres_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 9, 0, it));
    const int64_t n = S.n;
    T *r = S.take(n), *z = S.take(n);
    T *wb[3] = {S.take(n), S.take(n), S.take(n)};
    T *vb[2] = {S.take(n), S.take(n)}, *ub[2] = {S.take(n), S.take(n)};
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (MinRes.cxx:63)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = r;
    }
    hipLaunchKernelGGL(k_mr_init<T>, g, blk, 0, st, n, src, b, dg, x, r, z, it->init_guess_null ? 1 : 0, S.dst,
                       S.partials);
    S.fin(FinMrInit<T>{S.dst});
    hipLaunchKernelGGL(k_mr_start<T>, g, blk, 0, st, n, (const T *)r, (const T *)z, vb[0], ub[0], vb[1], ub[1],
                       wb[0], wb[1], wb[2], S.dst);
    SHK_HIP(hipGetLastError());
    int64_t t = 0;  // iterations enqueued: the buffers' rotation
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++, t++) {
            T *u = ub[t % 2], *u_old = ub[(t + 1) % 2], *v = vb[t % 2], *v_old = vb[(t + 1) % 2];
            T *w = wb[t % 3], *w_old = wb[(t + 2) % 3], *w_new = wb[(t + 1) % 3];
            SpmvArgs<T> a{};
            a.x = u; a.y_out = r; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = u; a.skip = &S.dst->skip;
            a.partials = S.part[1];
            SHK_TRY(S.spmv(a, FinDefer{}));
            grid_finish<T>(cs, (const T *)S.part[1], a.grid_used, (const int *)&S.dst->skip, FinMrAlpha<T>{S.dst});
            hipLaunchKernelGGL(k_mr_rz<T>, g, blk, 0, cs, n, dg, (const T *)v, (const T *)u, (const T *)v_old,
                               (const T *)u_old, r, z, S.dst, S.partials);
            S.fin(FinMrStep<T>{S.dst, k});
            hipLaunchKernelGGL(k_mr_wx<T>, g, blk, 0, cs, n, (const T *)u, (const T *)w, (const T *)w_old, w_new, x,
                               (const T *)r, (const T *)z, v_old, u_old, S.dst, k);
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "minres: solver did not reach a stopping state");
    return S.finish(it, code);
}

// Symmlq (Symmlq.cxx:45-161): per iteration (after the first) the v / u /
// w_bar / x update, then r = A u with (u, r), z and the r / z recurrences
// with (r, z) (MinRes's kernel), the QR step; after the loop the last
// x += ceta_bar w_bar.  v and u rotate through two buffers each.
template <typename T>
int symmlq_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 7, 0, it));
    const int64_t n = S.n;
    T *r = S.take(n), *z = S.take(n), *w_bar = S.take(n);
    T *vb[2] = {S.take(n), S.take(n)}, *ub[2] = {S.take(n), S.take(n)};
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (Symmlq.cxx:65)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = r;
    }
    hipLaunchKernelGGL(k_mr_init<T>, g, blk, 0, st, n, src, b, dg, x, r, z, it->init_guess_null ? 1 : 0, S.dst,
                       S.partials);
    S.fin(FinSyInit<T>{S.dst});
    hipLaunchKernelGGL(k_sy_start<T>, g, blk, 0, st, n, (const T *)r, (const T *)z, vb[0], ub[0], vb[1], ub[1],
                       w_bar, S.dst);
    SHK_HIP(hipGetLastError());
    int64_t t = 0;  // iterations enqueued: the v / u buffers' rotation
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) + 2;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++, t++) {
            T *u = ub[t % 2], *u_old = ub[(t + 1) % 2], *v = vb[t % 2], *v_old = vb[(t + 1) % 2];
            if (t > 0)
                hipLaunchKernelGGL(k_sy_upd<T>, g, blk, 0, cs, n, (const T *)r, (const T *)z, v, u, w_bar, x,
                                   S.dst);
            SpmvArgs<T> a{};
            a.x = u; a.y_out = r; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = u; a.skip = &S.dst->skip;
            a.partials = S.part[1];
            SHK_TRY(S.spmv(a, FinDefer{}));
            grid_finish<T>(cs, (const T *)S.part[1], a.grid_used, (const int *)&S.dst->skip, FinMrAlpha<T>{S.dst});
            hipLaunchKernelGGL(k_mr_rz<T>, g, blk, 0, cs, n, dg, (const T *)v, (const T *)u, (const T *)v_old,
                               (const T *)u_old, r, z, S.dst, S.partials);
            S.fin(FinSyStep<T>{S.dst});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "symmlq: solver did not reach a stopping state");
    hipLaunchKernelGGL(k_sy_final<T>, g, blk, 0, st, n, (const T *)w_bar, x, S.dst);
    SHK_HIP(hipGetLastError());
    return S.finish(it, code);
}

// QCgs (QCgs.cxx:47-164): Cgs's loop plus the smoothing step, whose five
// dots come from one six-value reduction; after the loop x = M^{-1}
// x_qcgs, as the template does.
template <typename T>
int qcgs_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, jac ? 11 : 10, 0, it));
    const int64_t n = S.n;
    T *p = S.take(n), *q = S.take(n), *r = S.take(n), *rt = S.take(n), *u = S.take(n), *rq = S.take(n),
      *xq = S.take(n), *v = S.take(n), *w = S.take(n), *phat = S.take(n);
    T *uhat = jac ? S.take(n) : u;
    const T *pin = jac ? phat : p;  // the identity's M^{-1} p is p itself
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // Init(b) (QCgs.cxx:59)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = r; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = r;
    }
    hipLaunchKernelGGL(k_qc_init<T>, g, blk, 0, st, n, src, x, r, rt, rq, xq, q, p, it->init_guess_null ? 1 : 0,
                       S.dst, S.part[0]);
    SHK_HIP(hipGetLastError());
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / 2 + 2;
    bool first = true;
    SHK_TRY(S.run_chunks(limit, kChunk, [&](hipStream_t cs) -> int {
        for (int k = 0; k < kChunk; k++) {
            grid_finish<T>(cs, (const T *)S.part[0], S.nb, (const int *)&S.dst->skip,
                           FinQcTop<T>{S.dst, first ? 1 : 0});
            first = false;
            hipLaunchKernelGGL(k_qc_up<T>, g, blk, 0, cs, n, (const T *)r, (const T *)q, p, u, dg, phat, S.dst);
            SpmvArgs<T> a{};
            a.x = pin; a.y_out = v; a.alpha = (T)1; a.mlt = 1;
            a.dot0 = kDotVec; a.u0 = rt; a.skip = &S.dst->skip;
            a.partials = S.part[1];
            SHK_TRY(S.spmv(a, FinDefer{}));
            grid_finish<T>(cs, (const T *)S.part[1], a.grid_used, (const int *)&S.dst->skip, FinQcSigma<T>{S.dst});
            hipLaunchKernelGGL(k_qc_qu<T>, g, blk, 0, cs, n, (const T *)v, q, u, x, dg, uhat, S.dst);
            SpmvArgs<T> a2{};
            a2.x = uhat; a2.y_out = w; a2.alpha = (T)1; a2.mlt = 1; a2.skip = &S.dst->skip;
            SHK_TRY(S.spmv(a2, FinNone{}));
            hipLaunchKernelGGL(k_qc_r<T>, g, blk, 0, cs, n, (const T *)w, (const T *)rq, (const T *)v, r, u, S.dst,
                               S.part[2]);
            hipLaunchKernelGGL((k_grid_finish6<T, FinQcSmooth<T>>), dim3(1), dim3(kBlock), 0, cs,
                               (const T *)S.part[2], S.nb, (const int *)&S.dst->skip, FinQcSmooth<T>{S.dst});
            hipLaunchKernelGGL(k_qc_smooth<T>, g, blk, 0, cs, n, (const T *)r, (const T *)u, (const T *)v,
                               (const T *)x, (const T *)p, (const T *)rt, rq, xq, S.dst, S.part[0]);
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "qcgs: solver did not reach a stopping state");
    hipLaunchKernelGGL(k_qc_final<T>, g, blk, 0, st, n, (const T *)xq, dg, x, S.dst);
    SHK_HIP(hipGetLastError());
    return S.finish(it, code);
}

// Gcr (Gcr.cxx:45-139), restarted every m = parameter_restart iterations:
// per cycle the outer head and p_0; per inner iteration (two of Seldon's
// iterations, at most ceil(m / 2) per cycle) its head, w_j = M^{-1} A p_j
// with (w_j, w_j) and (w_j, r), the x / r update, q = M^{-1} A r with the
// (w_i, q) in groups of six, and p_{j+1}.
template <typename T>
int gcr_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    const int m = it->parameter_restart;
    SHK_CHECK(m >= 1 && m <= 1000, SHK_ERR_ARG, "gcr: the restart parameter must be in 1 .. 1000");
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    SHK_TRY(S.setup(ctx, A, 3 + 2 * (size_t)(m + 1), 2 * (((size_t)(m + 1) * sizeof(T) + 255) & ~(size_t)255), it));
    const int64_t n = S.n;
    T *r = S.take(n), *q = S.take(n), *u = S.take(n);
    T *pbase = S.take(n);
    for (int i = 1; i <= m; i++) S.take(n);
    T *wbase = S.take(n);
    for (int i = 1; i <= m; i++) S.take(n);
    const int64_t stride = wbase - pbase == 0 ? 0 : (int64_t)((wbase - pbase) / (m + 1));
    SHK_CHECK(stride * (m + 1) == (int64_t)(wbase - pbase), SHK_ERR_STATE, "gcr: work vectors not evenly spaced");
    T *betas = S.take(m + 1), *deltas = S.take(m + 1);
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    const int *gskip = (const int *)&S.dst->gc_skip;
    // no cycle is open until the first outer head opens one
    S.hst->gc_skip = 1;
    SHK_HIP(hipMemcpyAsync(&S.dst->gc_skip, &S.hst->gc_skip, sizeof(int), hipMemcpyHostToDevice, st));
    // Init(b) (Gcr.cxx:55)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = u; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = u;
    }
    hipLaunchKernelGGL(k_gc_init<T>, g, blk, 0, st, n, src, dg, x, r, it->init_guess_null ? 1 : 0, S.dst,
                       S.part[0]);
    SHK_HIP(hipGetLastError());
    const int ni = (m + 1) / 2;  // inner iterations of a full cycle
    const int cycles = std::max(1, kChunk / ni);
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / m + 2;
    SHK_TRY(S.run_chunks(limit, cycles, [&](hipStream_t cs) -> int {
        for (int c = 0; c < cycles; c++) {
            grid_finish<T>(cs, (const T *)S.part[0], S.nb, (const int *)&S.dst->skip, FinGcOuter<T>{S.dst, m});
            hipLaunchKernelGGL(k_gc_p0<T>, g, blk, 0, cs, n, (const T *)r, pbase, S.dst);
            for (int j = 0; j < ni; j++) {
                T *pj = pbase + (size_t)j * stride, *wj = wbase + (size_t)j * stride;
                grid_finish<T>(cs, (const T *)S.part[0], S.nb, gskip, FinGcInner<T>{S.dst});
                SpmvArgs<T> a{};
                a.x = pj; a.y_out = u; a.alpha = (T)1; a.mlt = 1; a.skip = gskip;
                SHK_TRY(S.spmv(a, FinNone{}));
                hipLaunchKernelGGL(k_gc_w<T>, g, blk, 0, cs, n, (const T *)u, dg, (const T *)r, wj, S.dst,
                                   S.part[1]);
                grid_finish<T>(cs, (const T *)S.part[1], S.nb, gskip, FinGcAlpha<T>{S.dst, betas, j});
                hipLaunchKernelGGL(k_gc_xr<T>, g, blk, 0, cs, n, (const T *)pj, (const T *)wj, x, r, S.dst,
                                   S.part[0]);
                SpmvArgs<T> a2{};
                a2.x = r; a2.y_out = u; a2.alpha = (T)1; a2.mlt = 1; a2.skip = gskip;
                SHK_TRY(S.spmv(a2, FinNone{}));
                for (int g0 = 0; g0 <= j; g0 += 6) {
                    const int cnt = std::min(6, j + 1 - g0);
                    hipLaunchKernelGGL(k_gc_q<T>, g, blk, 0, cs, n, (const T *)u, dg, q, g0 == 0 ? 1 : 0,
                                       (const T *)wbase, stride, g0, cnt, S.dst, S.part[2]);
                    hipLaunchKernelGGL((k_grid_finish6<T, FinGcDelta<T>>), dim3(1), dim3(kBlock), 0, cs,
                                       (const T *)S.part[2], S.nb, gskip,
                                       FinGcDelta<T>{(const T *)betas, deltas, g0, cnt});
                }
                hipLaunchKernelGGL(k_gc_p<T>, g, blk, 0, cs, n, (const T *)r, pbase, stride, j,
                                   (const T *)deltas, S.dst);
            }
            // the inner head after the cycle's last iteration: it records its
            // Finished() and ends the inner loop (nb_iter reached the limit)
            grid_finish<T>(cs, (const T *)S.part[0], S.nb, gskip, FinGcInner<T>{S.dst});
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "gcr: solver did not reach a stopping state");
    return S.finish(it, code);
}

// BiCgStabl (BiCgStabl.cxx:43-180), l = parameter_restart: per cycle the
// loop head, the BiCG part (per j: the u update, u_{j+1} = A M^{-1} u_j with
// (u_{j+1}, r0~), the x / r update, r_{j+1} = A M^{-1} r_j with (r_{j+1},
// r0~)), the MR part's modified Gram-Schmidt (per step i one kernel for
// each group of up to six columns j >= i, the subtraction of step i-1 fused
// with the dots of step i), the small solve
// for gamma, and the x / r_0 / u_0 update with the next head's two sums.
// The template's loop goes on after a breakdown (Fail only ends the BiCG
// part); here the solve ends there.
template <typename T>
int bicgstabl_impl(shk_ctx ctx, shk_csr A, T *x, const T *b, int prec, shk_iteration *it, int *code)
{
    const int l = it->parameter_restart;
    SHK_CHECK(l >= 1 && l <= 100, SHK_ERR_ARG, "bicgstabl: the restart parameter (l) must be in 1 .. 100");
    Solve<T> S;
    const bool jac = prec == SHK_PREC_JACOBI;
    const size_t ld = (size_t)l + 1;
    const auto r256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    SHK_TRY(S.setup(ctx, A, 2 * ld + 2, r256(ld * ld * sizeof(T)) + 3 * r256(ld * sizeof(T)), it));
    const int64_t n = S.n;
    T *rbase = S.take(n);
    for (int i = 1; i <= l; i++) S.take(n);
    T *ubase = S.take(n);
    for (int i = 1; i <= l; i++) S.take(n);
    T *rt = S.take(n), *q = S.take(n);
    const int64_t stride = (int64_t)((ubase - rbase) / (int64_t)ld);
    SHK_CHECK(stride * (int64_t)ld == (int64_t)(ubase - rbase) && rt - ubase == stride * (int64_t)ld,
              SHK_ERR_STATE, "bicgstabl: work vectors not evenly spaced");
    T *tau = S.take((int64_t)(ld * ld)), *gamma = S.take((int64_t)ld), *gp = S.take((int64_t)ld),
      *gt = S.take((int64_t)ld);
    const T *dg = jac ? (const T *)A->diag : nullptr;
    hipStream_t st = ctx->stream;
    dim3 g(S.nb), blk(kBlock);
    // u_0, tau, gamma, gamma', gamma'' start at zero (:58-65); rho_0 = 1,
    // alpha = 0, omega = 1 (:85)
    SHK_HIP(hipMemsetAsync(ubase, 0, sizeof(T) * (size_t)n, st));
    SHK_HIP(hipMemsetAsync(tau, 0, sizeof(T) * ld * ld, st));
    SHK_HIP(hipMemsetAsync(gamma, 0, sizeof(T) * ld, st));
    SHK_HIP(hipMemsetAsync(gp, 0, sizeof(T) * ld, st));
    SHK_HIP(hipMemsetAsync(gt, 0, sizeof(T) * ld, st));
    S.hst->bl_rho0 = (T)1;
    S.hst->bl_alpha = (T)0;
    S.hst->bl_omega = (T)1;
    SHK_HIP(hipMemcpyAsync(S.dst, S.hst, sizeof(KState<T>), hipMemcpyHostToDevice, st));
    // Init(b) (:76)
    hipLaunchKernelGGL(k_axpy_dot<T>, g, blk, 0, st, n, (T)0, (const T *)nullptr, (T *)b, b,
                       (const T *)nullptr, S.partials, (const int *)&S.dst->skip);
    S.fin(FinInitNorm<T>{S.dst});
    const T *src = b;
    if (!it->init_guess_null) {
        SpmvArgs<T> a{};
        a.x = x; a.y_in = b; a.y_out = rbase; a.alpha = (T)-1; a.beta = (T)1; a.skip = &S.dst->skip;
        SHK_TRY(S.spmv(a, FinNone{}));
        src = rbase;
    }
    hipLaunchKernelGGL(k_bl_init<T>, g, blk, 0, st, n, src, x, rbase, rt, it->init_guess_null ? 1 : 0, S.dst,
                       S.part[0]);
    SHK_HIP(hipGetLastError());
    const int cycles = std::max(1, kChunk / l);
    const int64_t limit = (int64_t)std::max(it->max_iter, 0) / (2 * l) + 2;
    SHK_TRY(S.run_chunks(limit, cycles, [&](hipStream_t cs) -> int {
        const int *skip = (const int *)&S.dst->skip;
        for (int c = 0; c < cycles; c++) {
            grid_finish<T>(cs, (const T *)S.part[0], S.nb, skip, FinBlTop<T>{S.dst});
            for (int j = 0; j < l; j++) {
                T *uj = ubase + (size_t)j * stride, *rj = rbase + (size_t)j * stride;
                hipLaunchKernelGGL(k_bl_u<T>, g, blk, 0, cs, n, (const T *)rbase, ubase, stride, j, dg, q, S.dst);
                SpmvArgs<T> a{};
                a.x = jac ? q : uj; a.y_out = uj + stride; a.alpha = (T)1; a.mlt = 1;
                a.dot0 = kDotVec; a.u0 = rt; a.skip = skip; a.partials = S.part[1];
                SHK_TRY(S.spmv(a, FinBlSigma<T>{S.dst}));
                hipLaunchKernelGGL(k_bl_xr<T>, g, blk, 0, cs, n, rbase, (const T *)ubase, stride, j, dg, q, x,
                                   S.dst);
                SpmvArgs<T> a2{};
                a2.x = jac ? q : rj; a2.y_out = rj + stride; a2.alpha = (T)1; a2.mlt = 1;
                a2.dot0 = kDotVec; a2.u0 = rt; a2.skip = skip; a2.partials = S.part[1];
                SHK_TRY(S.spmv(a2, FinBlRho<T>{S.dst, j + 1 < l ? 1 : 0}));
            }
            // the MR part, step i over the columns j >= i: the first group
            // holds column i (its last subtraction, gamma(i), gamma'(i))
            // and up to four more, the next groups six each
            for (int i = 1; i <= l; i++)
                for (int j0 = i; j0 <= l;) {
                    const int cnt = std::min(j0 == i ? 5 : 6, l + 1 - j0);
                    hipLaunchKernelGGL(k_bl_mgs<T>, g, blk, 0, cs, n, rbase, stride, i, j0, cnt, (const T *)tau,
                                       (const T *)gamma, (int)ld, S.dst, S.part[2]);
                    hipLaunchKernelGGL((k_grid_finish6<T, FinBlMgs<T>>), dim3(1), dim3(kBlock), 0, cs,
                                       (const T *)S.part[2], S.nb, skip,
                                       FinBlMgs<T>{gamma, gp, tau, i, j0, cnt, (int)ld});
                    j0 += cnt;
                }
            hipLaunchKernelGGL(k_bl_solve<T>, dim3(1), dim3(64), 0, cs, S.dst, (const T *)tau, gamma,
                               (const T *)gp, gt, l);
            hipLaunchKernelGGL(k_bl_up<T>, g, blk, 0, cs, n, rbase, ubase, stride, l, (const T *)gamma,
                               (const T *)gp, (const T *)gt, (const T *)rt, x, S.dst, S.part[0]);
        }
        SHK_HIP(hipGetLastError());
        return SHK_OK;
    }));
    SHK_CHECK(S.hst->done, SHK_ERR_STATE, "bicgstabl: solver did not reach a stopping state");
    if (jac) {
        hipLaunchKernelGGL(k_bl_final<T>, g, blk, 0, st, n, dg, x, S.dst);
        SHK_HIP(hipGetLastError());
    }
    return S.finish(it, code);
}

template <typename T>
int solve_dispatch(int which, shk_ctx ctx, shk_csr A, void *dx, const void *db, int prec,
                   shk_iteration *it, int *code)
{
    if (which == 0) return cg_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 1) return bicgstab_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 3) return cgs_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 4) return bicg_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 5) return tfqmr_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 6) return qmr_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 7) return cgne_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 8) return lsqr_impl<T>(ctx, A, (T *)dx, (const T *)db, it, code);
    if (which == 10) return bicgcr_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 11) return qmrsym_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 12) return minres_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 13) return symmlq_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 14) return qcgs_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 15) return gcr_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    if (which == 16) return bicgstabl_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
    return gmres_impl<T>(ctx, A, (T *)dx, (const T *)db, prec, it, code);
}

static int solve_entry(int which, shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
                       shk_iteration *it, int *code, bool noalloc = false)
{
    SHK_CHECK(ctx && A && it && code, SHK_ERR_ARG, "solver: null argument");
    SHK_CHECK(A->ctx == ctx, SHK_ERR_ARG, "solver: matrix belongs to another context");
    SHK_CHECK(prec == SHK_PREC_IDENTITY || prec == SHK_PREC_JACOBI, SHK_ERR_ARG, "solver: bad preconditioner");
    SHK_CHECK(A->m == A->n, SHK_ERR_DIM, "solver: matrix must be square");
    *code = 0;
    // const int N = A.GetM(); if (N <= 0) return 0;  (Cg.cxx:49-51)
    if (A->m <= 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(db, "b"));
    SHK_CHECK(!(it->hist_cap > 0 && !it->hist), SHK_ERR_ARG, "solver: hist_cap > 0 needs hist");
    it->hist_len = 0;
    if (noalloc) {
        char *ext = (char *)noalloc_scratch(ctx->device);
        SHK_CHECK(ext, SHK_ERR_STATE, "shk_cg_noalloc: the device's scratch address was not set up");
        if (A->dtype == SHK_F64)
            return cg_impl<double>(ctx, A, (double *)dx, (const double *)db, prec, it, code, ext, kNoAllocBytes);
        return cg_impl<float>(ctx, A, (float *)dx, (const float *)db, prec, it, code, ext, kNoAllocBytes);
    }
    if (A->dtype == SHK_F64) return solve_dispatch<double>(which, ctx, A, dx, db, prec, it, code);
    return solve_dispatch<float>(which, ctx, A, dx, db, prec, it, code);
}


// ------------------------------------------------- iterative refinement ----
// FP32 inner Krylov solve + FP64 refinement (no Seldon counterpart).  FP64
// is used only for the true residual r = b - A x, its norm and the update
// x += |r| d; the inner GMRES keeps its Hessenberg / Givens least squares in
// FP64 (shk_solver_state.hpp) while every vector, the SpMV and the dot
// products run in FP32.

// r32 = (float)(r * (1/|r|))  -- the scaled correction equation A d = r/|r|
__global__ __launch_bounds__(kBlock) void k_ir_down(int64_t n, const double *r, double inv, float *r32, float *d32)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        r32[i] = (float)(r[i] * inv);
        d32[i] = 0.0f;
    }
}

// x += |r| * (double)d
__global__ __launch_bounds__(kBlock) void k_ir_up(int64_t n, const float *d32, double scale, double *x)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] += scale * (double)d32[i];
}

// true residual r = b - A x (FP64) and its squared norm to ctx scalars
static int ir_residual(shk_ctx ctx, shk_csr A64, const double *x, const double *b, double *r, double *nrm)
{
    SpmvArgs<double> a{};
    a.x = x; a.y_in = b; a.y_out = r; a.alpha = -1.0; a.beta = 1.0; a.dot0 = kDotSelf;
    SHK_TRY((launch_spmv<double, FinStore<double>>(A64, a, FinStore<double>{(double *)ctx->scalars})));
    double h[2];
    SHK_HIP(hipMemcpyAsync(h, ctx->scalars, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    SHK_HIP(hipStreamSynchronize(ctx->stream));
    *nrm = std::sqrt(h[0]);
    return SHK_OK;
}

static int ir_impl(shk_ctx ctx, shk_csr A64, shk_csr A32, double *x, const double *b, int prec,
                   shk_ir_params *p, int *code)
{
    const int64_t n = A64->m;
    const size_t vb8 = ((size_t)n * 8 + 255) & ~(size_t)255, vb4 = ((size_t)n * 4 + 255) & ~(size_t)255;
    SHK_TRY(ensure_ws2(ctx, vb8 + 2 * vb4));
    SHK_TRY(ensure_red(ctx, sizeof(double) * 2 * (size_t)std::max(red_blocks(n), A64->nblk_launch)));
    char *w = (char *)ctx->ws2;
    double *r = (double *)w;
    float *r32 = (float *)(w + vb8), *d32 = (float *)(w + vb8 + vb4);
    // |b|
    double bn2 = 0;
    {
        const int nb = red_blocks(n);
        hipLaunchKernelGGL(k_axpy_dot<double>, dim3(nb), dim3(kBlock), 0, ctx->stream, n, 0.0,
                           (const double *)nullptr, (double *)b, b, (const double *)nullptr,
                           (double *)ctx->red.partials, (const int *)nullptr);
        grid_finish<double>(ctx->stream, (const double *)ctx->red.partials, nb, nullptr,
                            FinStore<double>{(double *)ctx->scalars});
        SHK_HIP(hipGetLastError());
        SHK_HIP(hipMemcpyAsync(&bn2, ctx->scalars, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
    }
    const double bn = std::sqrt(bn2);
    p->refinements = 0;
    p->total_inner_iter = 0;
    *code = 0;
    if (bn == 0.0) {
        SHK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, ctx->stream));
        p->final_relres = 0.0;
        if (p->relres_hist) p->relres_hist[0] = 0.0;
        return SHK_OK;
    }
    double rn = 0;
    SHK_TRY(ir_residual(ctx, A64, x, b, r, &rn));
    double rel = rn / bn;
    if (p->relres_hist) p->relres_hist[0] = rel;
    const int nb = red_blocks(n);
    while (rel >= p->tolerance && p->refinements < p->max_refinements) {
        hipLaunchKernelGGL(k_ir_down, dim3(nb), dim3(kBlock), 0, ctx->stream, n, (const double *)r, 1.0 / rn, r32,
                           d32);
        SHK_HIP(hipGetLastError());
        shk_iteration it{};
        it.tolerance = p->inner_tolerance;
        it.max_iter = p->inner_max_iter;
        it.init_guess_null = 1;
        it.parameter_restart = p->inner_restart;
        int icode = 0;
        if (p->inner_method == 1) SHK_TRY(cg_impl<float>(ctx, A32, d32, r32, prec, &it, &icode));
        else SHK_TRY(gmres_impl<float>(ctx, A32, d32, r32, prec, &it, &icode));
        p->total_inner_iter += it.nb_iter;
        hipLaunchKernelGGL(k_ir_up, dim3(nb), dim3(kBlock), 0, ctx->stream, n, (const float *)d32, rn, x);
        SHK_HIP(hipGetLastError());
        SHK_TRY(ir_residual(ctx, A64, x, b, r, &rn));
        const double prev = rel;
        rel = rn / bn;
        p->refinements++;
        if (p->relres_hist) p->relres_hist[p->refinements] = rel;
        if (!(rel < prev)) {  // refinement stopped contracting (or NaN)
            *code = 1;
            break;
        }
    }
    p->final_relres = rel;
    if (*code == 0 && rel >= p->tolerance) *code = -2;
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

extern "C" {

int shk_cg(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(0, ctx, A, dx, db, prec, it, code);
}

int shk_cg_noalloc(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it,
                   int *code)
{
    return solve_entry(0, ctx, A, dx, db, prec, it, code, true);
}

int shk_bicgstab(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it,
                 int *code)
{
    return solve_entry(1, ctx, A, dx, db, prec, it, code);
}

int shk_gmres(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it,
              int *code)
{
    return solve_entry(2, ctx, A, dx, db, prec, it, code);
}

int shk_cgs(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(3, ctx, A, dx, db, prec, it, code);
}

int shk_bicg(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(4, ctx, A, dx, db, prec, it, code);
}

int shk_tfqmr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(5, ctx, A, dx, db, prec, it, code);
}

int shk_qmr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(6, ctx, A, dx, db, prec, it, code);
}

int shk_cgne(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(7, ctx, A, dx, db, prec, it, code);
}

int shk_lsqr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(8, ctx, A, dx, db, prec, it, code);
}

int shk_bicgcr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(10, ctx, A, dx, db, prec, it, code);
}

int shk_qmrsym(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(11, ctx, A, dx, db, prec, it, code);
}

int shk_minres(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(12, ctx, A, dx, db, prec, it, code);
}

int shk_symmlq(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(13, ctx, A, dx, db, prec, it, code);
}

int shk_qcgs(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(14, ctx, A, dx, db, prec, it, code);
}

int shk_gcr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it, int *code)
{
    return solve_entry(15, ctx, A, dx, db, prec, it, code);
}

int shk_bicgstabl(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec, shk_iteration *it,
                  int *code)
{
    return solve_entry(16, ctx, A, dx, db, prec, it, code);
}

}  // extern "C"

extern "C" int shk_solve_ir(shk_ctx ctx, shk_csr A64, shk_csr A32, void *dx, const void *db, shk_precond prec,
                            shk_ir_params *p, int *code)
{
    using namespace shk;
    SHK_CHECK(ctx && A64 && A32 && p && code, SHK_ERR_ARG, "shk_solve_ir: null argument");
    SHK_CHECK(A64->ctx == ctx && A32->ctx == ctx, SHK_ERR_ARG, "shk_solve_ir: matrices of another context");
    SHK_CHECK(A64->dtype == SHK_F64 && A32->dtype == SHK_F32, SHK_ERR_ARG,
              "shk_solve_ir: A64 must be FP64 and A32 FP32");
    SHK_CHECK(A64->m == A32->m && A64->n == A32->n && A64->nnz_stored == A32->nnz_stored &&
                  A64->storage == A32->storage,
              SHK_ERR_DIM, "shk_solve_ir: A64 and A32 must share the pattern");
    SHK_CHECK(A64->m == A64->n, SHK_ERR_DIM, "shk_solve_ir: square matrix needed");
    SHK_CHECK(p->max_refinements >= 0 && p->max_refinements < 100000, SHK_ERR_ARG, "shk_solve_ir: bad max_refinements");
    *code = 0;
    if (A64->m == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(db, "b"));
    return ir_impl(ctx, A64, A32, (double *)dx, (const double *)db, prec, p, code);
}

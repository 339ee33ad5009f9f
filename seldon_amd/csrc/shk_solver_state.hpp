// shk_solver_state.hpp -- device-resident Krylov state and the scalar
// "finalize" steps that the last work-group of each fused reduction runs.
//
// The host never reads a scalar inside an iteration: every dot / norm ends in
// a finalize below that applies Seldon's scalar recurrence and its
// Iteration::Finished / Fail logic on the device, then sets `skip` when the
// solve (or a GMRES cycle) is over so the kernels already queued return at
// once.  The host checks the state once per chunk of iterations.
#pragma once

#include "shk_kernels.hpp"

namespace shk {

template <typename T>
struct KState {
    // Iteration<double> (computation/solver/iterative/Iterative.hxx:47-108)
    double tol, facteur, last_res;
    int max_iter, nb_iter, error_code, fail;
    int done, init_failed, hist_len, hist_cap;
    int skip;          // done, or (GMRES) the current cycle is over
    int stop_slot;     // BiCgStab / TfQmr: loop slot (TfQmr: half step) that ended the solve
    double *hist;
    int *hist_it;
    // Cg / BiCgStab scalars, typed like Seldon's `Complexe` (= value type)
    T rho, rho_1, rho_2, alpha, beta, delta, omega, sigma;
    T rho_pp[2];       // Cg rho / BiCgStab rho_1 after the loop head of iteration k, in slot k % 2
    // TfQmr: theta, tau (Titer = double in the template) and eta (value type)
    double tq_theta, tq_tau;
    T tq_eta;
    // Qmr: rho, xi (Titer = double in the template); ep, theta, gamma, eta
    // and the d / s scale theta_1^2 gamma^2 (value type)
    double qm_rho, qm_xi;
    T qm_ep, qm_theta, qm_gamma, qm_eta, qm_tmp;
    // Lsqr: beta, alpha, rnorm (Titer = double); rho_bar, phi_bar and the
    // x / w factors phi / rho, -theta / rho (value type)
    double ls_beta, ls_alpha, ls_rnorm;
    T ls_rho_bar, ls_phi_bar, ls_xs, ls_ws;
    // MinRes: the residual estimate np (Titer = double) and the value-type
    // scalars of its recurrence (indices kMr*)
    double mr_np, sy_sprod;  // Symmlq keeps its np in mr_np too
    T mr[14];
    T qc_mu, qc_nu;  // QCgs's smoothing coefficients
    // Gcr: the cycle's inner limit, its norm |r| and gc_skip (1 when the
    // inner loop of the cycle is over or the solve has ended)
    double gc_normr;
    int gc_inner_max, gc_skip;
    // BiCgStabl: the BiCG part's rho_0, rho_1, alpha, beta, the MR part's
    // omega
    T bl_rho0, bl_rho1, bl_alpha, bl_beta, bl_omega;
    // Gmres(m)
    int gm_m, gm_i, gm_cycle_done, gm_inner_nb, gm_inner_max, gm_active;
    int cyc_skip;      // 1 when no cycle is open (cycle-end kernels return)
    double gm_beta, gm_hi;
    T gm_scale;        // (T)(1 / hi) for V_{i+1}, (T)(1 / beta) for V_0
    double *H;         // (m+1) x (m+1), column-major, H(k,i) = H[i*(m+1)+k]
    double *gs, *gc, *gsn;  // s, rotation cosines, sines
};

__device__ __forceinline__ double sqrt_t(double v) { return sqrt(v); }
__device__ __forceinline__ float sqrt_t(float v) { return sqrtf(v); }

// Iteration::Finished (Iterative.cxx:272-319) with the history record.
template <typename T>
__device__ inline bool it_finished(KState<T> *st, double norm, int nb, int maxit)
{
    double reste = st->facteur * norm;
    st->last_res = reste;
    int k = st->hist_len;
    if (k < st->hist_cap) {
        st->hist[k] = reste;
        st->hist_it[k] = nb;
    }
    st->hist_len = k + 1;
    return (reste < st->tol) || (nb >= maxit);
}

// Iteration::Fail (Iterative.cxx:323-331); the solver then breaks out.
template <typename T>
__device__ inline void it_fail(KState<T> *st, int code)
{
    st->fail = 1;
    st->error_code = code;
    st->done = 1;
    st->skip = 1;
}

template <typename T>
__device__ inline void it_stop(KState<T> *st)
{
    st->done = 1;
    st->skip = 1;
}

// Iteration::Init on |v| (Iterative.cxx:235-249); false when |v| == 0.
template <typename T>
__device__ inline bool it_init(KState<T> *st, T sumsq)
{
    double nr = (double)sqrt_t(sumsq);
    if (nr == 0.0) {
        st->init_failed = 1;
        it_stop(st);
        return false;
    }
    st->facteur = 1.0 / nr;
    st->nb_iter = 0;
    return true;
}

// ---------------------------------------------------------------- Cg ----
// After rho = (r, z) (Cg.cxx:78-94): breakdown test and beta.
template <typename T>
__device__ inline void cg_after_rho(KState<T> *st, T rho)
{
    st->rho = rho;
    if (rho == (T)0) {
        it_fail(st, 1);
        return;
    }
    if (st->nb_iter != 0) st->beta = rho / st->rho_1;
}

template <typename T>
struct FinInitNorm {  // Init(b): 1/|b|
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const { it_init(st, s[0]); }
};

// Loop head of Cg: [rho_1 = rho; ++iter;] Finished(r); identity: rho = |r|^2,
// also kept in rho_pp[slot] for a later kernel that must read it while this
// head runs again (the head folded into the p update).
template <typename T>
struct FinCgTop {
    KState<T> *st;
    int inc, identity, slot;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (inc) {
            st->rho_1 = st->rho;
            st->nb_iter++;
        }
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        if (identity) {
            cg_after_rho(st, s[0]);
            st->rho_pp[slot] = st->rho;
        }
    }
};

template <typename T>
struct FinCgRho {  // Jacobi: rho = (r, z)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const { cg_after_rho(st, s[0]); }
};

template <typename T>
struct FinCgDelta {  // delta = (p, q) (Cg.cxx:98-104)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->delta = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 2);
            return;
        }
        st->alpha = st->rho / s[0];
    }
};

// ---------------------------------------------------------- BiCgStab ----
// Loop head (BiCgStab.cxx:71-96): [rho_2 = rho_1; ++iter;] Finished(r);
// rho_1 = (rtilde, r); breakdowns #1, #2; beta.
template <typename T>
struct FinBiTop {
    KState<T> *st;
    int inc, slot;  // rho_1 also kept in rho_pp[slot] (the head folded into the p update)
    __device__ void operator()(const T (&s)[2]) const
    {
        if (inc) {
            st->rho_2 = st->rho_1;
            st->nb_iter++;
        }
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        st->rho_1 = s[1];
        st->rho_pp[slot] = s[1];
        if (s[1] == (T)0) {
            it_fail(st, 1);
            return;
        }
        if (st->nb_iter != 0) {
            if (st->omega == (T)0) {
                it_fail(st, 2);
                return;
            }
            st->beta = (st->rho_1 / st->rho_2) * (st->alpha / st->omega);
        }
    }
};

template <typename T>
struct FinBiSigma {  // sigma = (rtilde, v); breakdown #3; alpha (:104-110)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->sigma = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 3);
            return;
        }
        st->alpha = st->rho_1 / s[0];
    }
};

template <typename T>
struct FinBiS {  // ++iter; Finished(s) (:115-121)
    KState<T> *st;
    int slot;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->nb_iter++;
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) {
            st->stop_slot = slot;
            it_stop(st);
        }
    }
};

template <typename T>
struct FinBiOmega {  // omega = (t, s) / (t, t) (:129)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const { st->omega = s[0] / s[1]; }
};

// -------------------------------------------------------- Cgs / BiCg ----
// Loop head shared by Cgs (Cgs.cxx:71-79) and BiCg (BiCg.cxx:72-86):
// [rho_2 = rho_1; ++iter;] Finished(r); rho_1 = s[1] ((rtilde, r) for Cgs,
// (z, rtilde) for BiCg); breakdown #1; beta = rho_1 / rho_2 after the
// first iteration.
template <typename T>
struct FinLanczosTop {
    KState<T> *st;
    int inc;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (inc) {
            st->rho_2 = st->rho_1;
            st->nb_iter++;
        }
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        st->rho_1 = s[1];
        if (s[1] == (T)0) {
            it_fail(st, 1);
            return;
        }
        if (st->nb_iter != 0) st->beta = s[1] / st->rho_2;
    }
};

// After the first product of a loop: ++iter; delta; breakdown #2; alpha =
// rho_1 / delta (Cgs.cxx:106-117, BiCg.cxx:104-115).
template <typename T>
struct FinLanczosDelta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->nb_iter++;
        st->delta = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 2);
            return;
        }
        st->alpha = st->rho_1 / s[0];
    }
};

// ------------------------------------------------------------- TfQmr ----
// After r0 = M^{-1} (b - A x): Init(r0); tau = |r0|; theta = eta = 0;
// rho = rho0 = (rtilde, r0) with rtilde = r0 (TfQmr.cxx:72-110).
template <typename T>
struct FinTqInit {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (!it_init(st, s[0])) return;
        st->tq_tau = (double)sqrt_t(s[0]);
        st->tq_theta = 0.0;
        st->tq_eta = (T)0;
        st->rho = s[1];
        st->rho_1 = s[1];
    }
};

// sigma = (rtilde, v); breakdown 1; alpha = rho / sigma; breakdown 2 when
// alpha is 0 (TfQmr.cxx:116-140; the template tests alpha after the w
// update, which changes nothing then)
template <typename T>
struct FinTqSigma {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->sigma = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 1);
            return;
        }
        const T alpha = st->rho / s[0];
        st->alpha = alpha;
        if (alpha == (T)0) it_fail(st, 2);
    }
};

// One half step's scalars from |w|^2 (TfQmr.cxx:145-172 and :180-205):
// breakdown 3 / 4 when tau is 0; theta = |w| / tau; c = 1 / sqrt(1 +
// theta^2); tau *= c theta; eta = c^2 alpha (formed in double, kept in
// the value type); kappa = tau sqrt(2 (iter + 1) [+ 1]); Finished(kappa).
// The x += eta d of this half runs whether or not the solve ends here, so
// an end records its slot (stop_slot) for the kernel that applies it.  The
// second half then takes rho0 = rho, rho = (rtilde, w) = s[1],
// breakdown 5, beta = rho / rho0 and the loop's ++iter (:209-240).
template <typename T>
struct FinTqHalf {
    KState<T> *st;
    int half, slot;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double tau = st->tq_tau;
        if (tau == 0.0) {
            it_fail(st, half ? 4 : 3);
            return;
        }
        const double theta = (double)sqrt_t(s[0]) / tau;
        const double c = 1.0 / sqrt(1.0 + theta * theta);
        const double tau2 = tau * c * theta;
        st->tq_theta = theta;
        st->tq_tau = tau2;
        st->tq_eta = (T)(c * c * (double)st->alpha);
        const double kappa = tau2 * sqrt(half ? 2. * (st->nb_iter + 1) + 1. : 2. * (st->nb_iter + 1));
        if (it_finished(st, kappa, st->nb_iter, st->max_iter)) {
            st->stop_slot = 2 * slot + half;
            it_stop(st);
            return;
        }
        if (!half) {
            st->nb_iter++;
            return;
        }
        const T rho0 = st->rho;
        st->rho_1 = rho0;
        st->rho = s[1];
        if (rho0 == (T)0) {
            st->stop_slot = 2 * slot + half;
            it_fail(st, 5);
            return;
        }
        st->beta = s[1] / rho0;
        st->nb_iter++;
    }
};

// --------------------------------------------------------------- Qmr ----
// Loop head (Qmr.cxx:79-90): [++iter;] Finished(r); breakdowns #1 (rho =
// 0) and #2 (xi = 0).  The first head also takes rho = |y| and xi = |w|
// from s (y = M^{-1} r, w = r) and the scalars' starting values
// (:49-77): theta = 0, gamma = 1, eta = -1, ep = 0; there r = w, so |r| is
// sqrt(s[1]).
template <typename T>
struct FinQmrTop {
    KState<T> *st;
    int first;
    __device__ void operator()(const T (&s)[2]) const
    {
        double rnorm;
        if (first) {
            st->qm_rho = (double)sqrt_t(s[0]);
            st->qm_xi = (double)sqrt_t(s[1]);
            st->qm_theta = (T)0;
            st->qm_gamma = (T)1;
            st->qm_eta = (T)-1;
            st->qm_ep = (T)0;
            st->nb_iter = 0;
            rnorm = (double)sqrt_t(s[1]);
        } else {
            st->nb_iter++;
            rnorm = (double)sqrt_t(s[0]);
        }
        if (it_finished(st, rnorm, st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        if (st->qm_rho == 0.0) {
            it_fail(st, 1);
            return;
        }
        if (st->qm_xi == 0.0) it_fail(st, 2);
    }
};

// delta = (w, y) after the scaling; breakdown #3 (Qmr.cxx:99-104)
template <typename T>
struct FinQmrDelta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->delta = s[0];
        if (s[0] == (T)0) it_fail(st, 3);
    }
};

// ep = (q, p~); breakdown #4; beta = ep / delta; breakdown #5 (:124-137)
template <typename T>
struct FinQmrEp {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->qm_ep = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 4);
            return;
        }
        const T beta = s[0] / st->delta;
        st->beta = beta;
        if (beta == (T)0) it_fail(st, 5);
    }
};

// rho = |y|, xi = |w| (s[0], s[1]); ++iter; theta, gamma, breakdown #6,
// eta and the d / s scale (Qmr.cxx:143-172), each expression in the
// template's types
template <typename T>
struct FinQmrTheta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double rho_1 = st->qm_rho;
        const double rho = (double)sqrt_t(s[0]);
        st->qm_rho = rho;
        st->qm_xi = (double)sqrt_t(s[1]);
        const T gamma_1 = st->qm_gamma, theta_1 = st->qm_theta, beta = st->beta;
        st->nb_iter++;
        const T theta = (T)(rho / (double)(gamma_1 * beta));
        const T gamma = (T)((double)(T)1 / sqrt(1.0 + (double)(theta * theta)));
        st->qm_theta = theta;
        st->qm_gamma = gamma;
        if (gamma == (T)0) {
            it_fail(st, 6);
            return;
        }
        st->qm_eta = (T)((double)(-st->qm_eta) * rho_1 * (double)gamma * (double)gamma /
                         (double)(beta * gamma_1 * gamma_1));
        st->qm_tmp = theta_1 * theta_1 * gamma * gamma;
    }
};

// -------------------------------------------------------------- Cgne ----
// Init(q) with q = A^T b, and the first Finished(|q|) (Cgne.cxx:55-81)
template <typename T>
struct FinCgneInit {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (!it_init(st, s[0])) return;
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) it_stop(st);
    }
};

// rho = (r, z); breakdown #1; beta (:88-101)
template <typename T>
struct FinCgneRho {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->rho = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 1);
            return;
        }
        if (st->nb_iter != 0) st->beta = s[0] / st->rho_1;
    }
};

// delta = (p, z) with z = A^T A p; breakdown #2; alpha (:104-113)
template <typename T>
struct FinCgneDelta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->delta = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 2);
            return;
        }
        st->alpha = st->rho / s[0];
    }
};

// rho_1 = rho; dp = |r|; two ++iter; Finished(dp) (:115-123, :82)
template <typename T>
struct FinCgneTop {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->rho_1 = st->rho;
        st->nb_iter += 2;
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) it_stop(st);
    }
};

// -------------------------------------------------------------- Lsqr ----
// rnorm = |b - A x| (s[0]) and beta = |b| (s[1]) (Lsqr.cxx:57-67)
template <typename T>
struct FinLsqrNorms {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->ls_rnorm = (double)sqrt_t(s[0]);
        st->ls_beta = (double)sqrt_t(s[1]);
    }
};

// alpha = |A^T u|; phi_bar = beta; rho_bar = alpha; the first
// Finished(rnorm) (:70-80)
template <typename T>
struct FinLsqrStart {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double alpha = (double)sqrt_t(s[0]);
        st->ls_alpha = alpha;
        st->ls_phi_bar = (T)st->ls_beta;
        st->ls_rho_bar = (T)alpha;
        st->nb_iter = 0;
        if (it_finished(st, st->ls_rnorm, st->nb_iter, st->max_iter)) it_stop(st);
    }
};

// beta = |u1|; breakdown #1 (:86-92)
template <typename T>
struct FinLsqrBeta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double beta = (double)sqrt_t(s[0]);
        st->ls_beta = beta;
        if (beta == 0.0) it_fail(st, 1);
    }
};

// alpha = |v1|; breakdown #2; the rotation (rho, breakdown #3, c, s, theta,
// rho_bar, phi, phi_bar) and the x / w factors, each in the template's
// types; rnorm = |phi_bar|; two ++iter; Finished(rnorm), whose end records
// its slot: the x update of this loop still runs (:97-145)
template <typename T>
struct FinLsqrAlpha {
    KState<T> *st;
    int slot;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double alpha = (double)sqrt_t(s[0]), beta = st->ls_beta;
        st->ls_alpha = alpha;
        if (alpha == 0.0) {
            it_fail(st, 2);
            return;
        }
        const T rho_bar = st->ls_rho_bar;
        const T rho = (T)sqrt((double)(rho_bar * rho_bar) + beta * beta);
        if (rho == (T)0) {
            it_fail(st, 3);
            return;
        }
        const T c = rho_bar / rho;
        const T sn = (T)(beta / (double)rho);
        const T theta = (T)((double)sn * alpha);
        st->ls_rho_bar = (T)((double)(-c) * alpha);
        const T phi = c * st->ls_phi_bar;
        const T phi_bar = sn * st->ls_phi_bar;
        st->ls_phi_bar = phi_bar;
        st->ls_xs = phi / rho;
        st->ls_ws = -theta / rho;
        st->ls_rnorm = (double)(phi_bar < (T)0 ? -phi_bar : phi_bar);
        st->nb_iter += 2;
        if (it_finished(st, st->ls_rnorm, st->nb_iter, st->max_iter)) {
            st->stop_slot = slot;
            it_stop(st);
        }
    }
};

// ------------------------------------------------------------ BiCgcr ----
// Loop head: [++iter;] Finished(v) (BiCgcr.cxx:73-80); the first head
// counts the start's product (nb_iter = 1)
template <typename T>
struct FinBcrTop {
    KState<T> *st;
    int first;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (first) st->nb_iter = 1;
        else st->nb_iter++;
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) it_stop(st);
    }
};

// rho = (w, v), mu = (w, a); breakdown #1; alpha = rho / mu (:82-90)
template <typename T>
struct FinBcrMu {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->rho = s[0];
        st->delta = s[1];  // mu
        if (s[1] == (T)0) {
            it_fail(st, 1);
            return;
        }
        st->alpha = s[0] / s[1];
    }
};

// tau = (w, z); breakdown #2; beta = tau / mu (:96-104)
template <typename T>
struct FinBcrTau {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->sigma = s[0];  // tau
        if (s[0] == (T)0) {
            it_fail(st, 2);
            return;
        }
        st->beta = s[0] / st->delta;
    }
};

// ------------------------------------------------------------ QmrSym ----
// Loop head (QmrSym.cxx:75-85): [++iter;] Finished(r); breakdown #1
// (rho = 0).  The first head also takes rho = |y| (s[0]) and the scalars'
// starting values; |r| is sqrt(s[1]) there and sqrt(s[0]) later.
template <typename T>
struct FinQmsTop {
    KState<T> *st;
    int first;
    __device__ void operator()(const T (&s)[2]) const
    {
        double rnorm;
        if (first) {
            st->qm_rho = (double)sqrt_t(s[0]);
            st->qm_theta = (T)0;
            st->qm_gamma = (T)1;
            st->qm_eta = (T)-1;
            st->qm_ep = (T)0;
            st->nb_iter = 0;
            rnorm = (double)sqrt_t(s[1]);
        } else {
            st->nb_iter++;
            rnorm = (double)sqrt_t(s[0]);
        }
        if (it_finished(st, rnorm, st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        if (st->qm_rho == 0.0) it_fail(st, 1);
    }
};

// delta = (v, y); its breakdown is error code 3 (:92-96)
template <typename T>
struct FinQmsDelta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->delta = s[0];
        if (s[0] == (T)0) it_fail(st, 3);
    }
};

// ep = (p, p~), code 4; beta = ep / delta, code 5 (:107-119)
template <typename T>
struct FinQmsEp {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->qm_ep = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 4);
            return;
        }
        const T beta = s[0] / st->delta;
        st->beta = beta;
        if (beta == (T)0) it_fail(st, 5);
    }
};

// rho = |y|; theta, gamma (code 6 when 0), eta and the d / s scale, in the
// template's types (:124-140); the first loop's d / s start is chosen by
// the kernel from nb_iter (still 0 there)
template <typename T>
struct FinQmsTheta {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double rho_1 = st->qm_rho;
        const double rho = (double)sqrt_t(s[0]);
        st->qm_rho = rho;
        const T gamma_1 = st->qm_gamma, theta_1 = st->qm_theta, beta = st->beta;
        const T theta = (T)(rho / (double)(gamma_1 * beta));
        const T gamma = (T)((double)(T)1 / sqrt(1.0 + (double)(theta * theta)));
        st->qm_theta = theta;
        st->qm_gamma = gamma;
        if (gamma == (T)0) {
            it_fail(st, 6);
            return;
        }
        st->qm_eta = (T)((double)(-st->qm_eta) * rho_1 * (double)gamma * (double)gamma /
                         (double)(beta * gamma_1 * gamma_1));
        st->qm_tmp = theta_1 * theta_1 * gamma * gamma;
    }
};

// ------------------------------------------------------------ MinRes ----
enum { kMrBeta, kMrEta, kMrIbeta, kMrAlpha, kMrC, kMrS, kMrCold, kMrSold, kMrRho2, kMrRho3, kMrIrho1, kMrCeta };

// dp = sqrt((r, z)); beta = eta = dp; ibeta = 1 / beta; the rotation's
// start; np = |b|; the first Finished(np) (MinRes.cxx:74-89)
template <typename T>
struct FinMrInit {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        T *m = st->mr;
        const T dp = sqrt_t(s[0]);
        m[kMrBeta] = dp;
        m[kMrEta] = dp;
        m[kMrIbeta] = (T)(1.0 / (double)dp);
        m[kMrC] = (T)1;
        m[kMrS] = (T)0;
        m[kMrCold] = (T)1;
        m[kMrSold] = (T)0;
        st->mr_np = (double)sqrt_t(s[1]);
        st->nb_iter = 0;
        if (it_finished(st, st->mr_np, st->nb_iter, st->max_iter)) it_stop(st);
    }
};

template <typename T>
struct FinMrAlpha {  // alpha = (A u, u) (:94)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const { st->mr[kMrAlpha] = s[0]; }
};

// beta = sqrt((r, z)); the QR step (rho0..rho3, breakdown #1, c, s); the
// factors of the w / x update; breakdown #2 (beta = 0, after that update);
// ibeta; np *= |s|; ++iter; Finished(np).  An end after the w / x update
// records its slot: that update still runs (:106-148)
template <typename T>
struct FinMrStep {
    KState<T> *st;
    int slot;
    __device__ void operator()(const T (&sm)[2]) const
    {
        T *m = st->mr;
        const T beta_old = m[kMrBeta], alpha = m[kMrAlpha];
        const T beta = sqrt_t(sm[0]);
        m[kMrBeta] = beta;
        const T coold = m[kMrCold], cold = m[kMrC], soold = m[kMrSold], sold = m[kMrS];
        m[kMrCold] = cold;
        m[kMrSold] = sold;
        const T rho0 = cold * alpha - coold * sold * beta_old;
        const T rho1 = sqrt_t(rho0 * rho0 + beta * beta);
        const T rho2 = sold * alpha + coold * cold * beta_old;
        const T rho3 = soold * beta_old;
        if (rho1 == (T)0) {
            it_fail(st, 1);
            return;
        }
        const T c = rho0 / rho1, s = beta / rho1;
        m[kMrC] = c;
        m[kMrS] = s;
        m[kMrRho2] = rho2;
        m[kMrRho3] = rho3;
        m[kMrIrho1] = (T)(1. / (double)rho1);
        const T eta = m[kMrEta];
        m[kMrCeta] = c * eta;
        m[kMrEta] = -s * eta;
        if (beta == (T)0) {
            st->stop_slot = slot;
            it_fail(st, 2);
            return;
        }
        m[kMrIbeta] = (T)(1.0 / (double)beta);
        st->mr_np *= (double)(s < (T)0 ? -s : s);
        st->nb_iter++;
        if (it_finished(st, st->mr_np, st->nb_iter, st->max_iter)) {
            st->stop_slot = slot;
            it_stop(st);
        }
    }
};

// ------------------------------------------------------------ Symmlq ----
// Symmlq's value-type scalars share MinRes's array: beta, alpha, c, s,
// cold, sold as there, and these
enum { kSyBeta1 = kMrRho2, kSyCeta = kMrRho3, kSyCetaOld = kMrIrho1 };

// dp = sqrt((r, z)); beta = beta1 = dp; s_prod = |beta1|; the rotation's
// start; ceta = ceta_old = 0; np = |b|; the first Finished(np)
// (Symmlq.cxx:74-92)
template <typename T>
struct FinSyInit {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        T *m = st->mr;
        const T dp = sqrt_t(s[0]);
        m[kMrBeta] = dp;
        m[kSyBeta1] = dp;
        st->sy_sprod = (double)(dp < (T)0 ? -dp : dp);
        m[kMrIbeta] = (T)(1.0 / (double)dp);
        m[kMrC] = (T)1;
        m[kMrS] = (T)0;
        m[kMrCold] = (T)1;
        m[kMrSold] = (T)0;
        m[kSyCeta] = (T)0;
        m[kSyCetaOld] = (T)0;
        st->mr_np = (double)sqrt_t(s[1]);
        st->nb_iter = 0;
        if (it_finished(st, st->mr_np, st->nb_iter, st->max_iter)) it_stop(st);
    }
};

// beta = sqrt((r, z)); the QR step and Givens rotation; ceta (the first
// iteration from beta1); s_prod *= |s|; np = s_prod / |c| (or s_prod 1e16
// when c = 0); ++iter; Finished(np) (:122-150).  kMrIbeta keeps 1 / beta
// for the next iteration's v, u (:99)
template <typename T>
struct FinSyStep {
    KState<T> *st;
    __device__ void operator()(const T (&sm)[2]) const
    {
        T *m = st->mr;
        const T beta_old = m[kMrBeta], alpha = m[kMrAlpha];
        const T beta = sqrt_t(sm[0]);
        m[kMrBeta] = beta;
        const T coold = m[kMrCold], cold = m[kMrC], soold = m[kMrSold], sold = m[kMrS];
        m[kMrCold] = cold;
        m[kMrSold] = sold;
        const T rho0 = cold * alpha - coold * sold * beta_old;
        const T rho1 = sqrt_t(rho0 * rho0 + beta * beta);
        const T rho2 = sold * alpha + coold * cold * beta_old;
        const T rho3 = soold * beta_old;
        const T c = rho0 / rho1, s = beta / rho1;
        m[kMrC] = c;
        m[kMrS] = s;
        const T ceta_old = m[kSyCeta], ceta_oold = m[kSyCetaOld];
        const T ceta = st->nb_iter == 0 ? m[kSyBeta1] / rho1 : -(rho2 * ceta_old + rho3 * ceta_oold) / rho1;
        m[kSyCetaOld] = ceta_old;
        m[kSyCeta] = ceta;
        m[kMrIbeta] = (T)(1.0 / (double)beta);
        st->sy_sprod *= (double)(s < (T)0 ? -s : s);
        st->mr_np = c == (T)0 ? st->sy_sprod * 1e16 : st->sy_sprod / (double)(c < (T)0 ? -c : c);
        st->nb_iter++;
        if (it_finished(st, st->mr_np, st->nb_iter, st->max_iter)) it_stop(st);
    }
};

// -------------------------------------------------------------- QCgs ----
// Loop head (QCgs.cxx:80-90): [++iter;] Finished(r_qcgs); rho_2 =
// (rtilde, r); breakdown #1 (rho_1 = 0); beta = rho_2 / rho_1.  The first
// head starts rho_1 at 1 (:70).
template <typename T>
struct FinQcTop {
    KState<T> *st;
    int first;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (first) st->rho_1 = (T)1;
        else st->nb_iter++;
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        st->rho_2 = s[1];
        if (st->rho_1 == (T)0) {
            it_fail(st, 1);
            return;
        }
        st->beta = s[1] / st->rho_1;
    }
};

// ++iter; sigma = (rtilde, v); breakdown #2; alpha = rho_2 / sigma (:103-112)
template <typename T>
struct FinQcSigma {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->nb_iter++;
        st->sigma = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 2);
            return;
        }
        st->alpha = st->rho_2 / s[0];
    }
};

// The smoothing step's 2 x 2 system from (u, u), (v, v), (u, v), (u, r),
// (v, r): breakdown #3 when its determinant is 0; mu, nu; rho_1 = rho_2
// (:130-152, bt_b symmetric: bt_b(0,1) = bt_b(1,0))
template <typename T>
struct FinQcSmooth {
    KState<T> *st;
    __device__ void operator()(const T (&s)[6]) const
    {
        const T b00 = s[0], b11 = s[1], b10 = s[2];
        const T delta = b00 * b11 - b10 * b10;
        if (delta == (T)0) {
            it_fail(st, 3);
            return;
        }
        const T m00 = b11 / delta, m11 = b00 / delta, m10 = -b10 / delta;
        const T rn0 = -s[3], rn1 = -s[4];
        st->qc_mu = m00 * rn0 + m10 * rn1;
        st->qc_nu = m10 * rn0 + m11 * rn1;
        st->rho_1 = st->rho_2;
    }
};

// --------------------------------------------------------------- Gcr ----
// Outer loop head (Gcr.cxx:82-89): Finished(r) on the outer Iteration; a new
// cycle's inner Iteration gets the limit nb_iter + m; normr = |r|
template <typename T>
struct FinGcOuter {
    KState<T> *st;
    int m;
    __device__ void operator()(const T (&s)[2]) const
    {
        const double rn = (double)sqrt_t(s[0]);
        st->gc_normr = rn;
        if (it_finished(st, rn, st->nb_iter, st->max_iter)) {
            it_stop(st);
            st->gc_skip = 1;
            return;
        }
        st->gc_inner_max = st->nb_iter + m;
        st->gc_skip = 0;
    }
};

// Inner loop head (:94): Finished(r) against the cycle's limit; true ends
// the cycle's inner loop (its kernels then return)
template <typename T>
struct FinGcInner {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->gc_inner_max)) st->gc_skip = 1;
    }
};

// beta_j = (w_j, w_j); breakdown #1; alpha = (w_j, r) / beta_j; the two
// ++iter of the inner iteration (:103-133; nothing reads the count between
// them).  The template's outer loop would go on after the breakdown; here
// the solve ends.
template <typename T>
struct FinGcAlpha {
    KState<T> *st;
    T *betas;
    int j;
    __device__ void operator()(const T (&s)[2]) const
    {
        betas[j] = s[0];
        if (s[0] == (T)0) {
            it_fail(st, 1);
            st->gc_skip = 1;
            return;
        }
        st->alpha = s[1] / s[0];
        st->nb_iter += 2;
    }
};

// delta_i = -(w_i, q) / beta_i for the group i = g0 .. g0 + cnt - 1 (:124-128)
template <typename T>
struct FinGcDelta {
    const T *betas;
    T *deltas;
    int g0, cnt;
    __device__ void operator()(const T (&s)[6]) const
    {
        for (int k = 0; k < cnt; k++) deltas[g0 + k] = -s[k] / betas[g0 + k];
    }
};

// --------------------------------------------------------- BiCgStabl ----
// Loop head (BiCgStabl.cxx:87-101): Finished(r_0); rho_0 *= -omega; then the
// BiCG part's j = 0 head: rho_1 = (r_0, r0~), breakdown #1, beta, rho_0.
// s = {|r_0|^2, (r_0, r0~)}.
template <typename T>
__device__ inline void bl_rho_head(KState<T> *st, T rho1)
{
    st->bl_rho1 = rho1;
    if (st->bl_rho0 == (T)0) {
        it_fail(st, 1);
        return;
    }
    st->bl_beta = st->bl_alpha * (rho1 / st->bl_rho0);
    st->bl_rho0 = rho1;
}

template <typename T>
struct FinBlTop {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (it_finished(st, (double)sqrt_t(s[0]), st->nb_iter, st->max_iter)) {
            it_stop(st);
            return;
        }
        st->bl_rho0 *= -st->bl_omega;
        bl_rho_head(st, s[1]);
    }
};

// ++iter; sigma = (u_{j+1}, r0~); breakdown #2; alpha = rho_1 / sigma
// (:110-118)
template <typename T>
struct FinBlSigma {
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->nb_iter++;
        if (s[0] == (T)0) {
            it_fail(st, 2);
            return;
        }
        st->bl_alpha = st->bl_rho1 / s[0];
    }
};

// ++iter after r_{j+1} = A M^{-1} r_j (:125); for j + 1 < l the next BiCG
// head on s[0] = (r_{j+1}, r0~)
template <typename T>
struct FinBlRho {
    KState<T> *st;
    int head;
    __device__ void operator()(const T (&s)[2]) const
    {
        st->nb_iter++;
        if (head) bl_rho_head(st, s[0]);
    }
};

// Step i of the MR part's Gram-Schmidt for the columns j0 .. j0+cnt-1
// (:131-141): where the group starts at column i, gamma(i) = |r_i|^2 and
// gamma'(i) = (r_0, r_i) / gamma(i) (where gamma(i) != 0) from s[0], s[1];
// then tau(i, j) = (r_j, r_i) / gamma(i) for the group's j > i, formed only
// where gamma(i) != 0 (tau keeps its old value otherwise).
template <typename T>
struct FinBlMgs {
    T *gamma, *gp, *tau;
    int i, j0, cnt, ld;
    __device__ void operator()(const T (&s)[6]) const
    {
        const bool head = j0 == i;
        if (head) {
            gamma[i] = s[0];
            if (s[0] != (T)0) gp[i] = s[1] / s[0];
        }
        const T gi = gamma[i];
        if (gi == (T)0) return;
        for (int k = head ? 1 : 0; k < cnt; k++) tau[(size_t)i * ld + j0 + k] = s[head ? k + 1 : k] / gi;
    }
};

// ------------------------------------------------------------- Gmres ----
// GenRot "old BLAS version" (Functions_Vector.cxx:612-649) and ApplyRot
// (:681-687), in double.
__device__ inline void dev_genrot(double &a_in, double &b_in, double &c_, double &s_)
{
    double roe = fabs(a_in) > fabs(b_in) ? a_in : b_in;
    double scal = fabs(a_in) + fabs(b_in);
    double r, z;
    if (scal != 0.0) {
        double a_scl = a_in / scal;
        double b_scl = b_in / scal;
        r = scal * sqrt(a_scl * a_scl + b_scl * b_scl);
        if (roe < 0.0) r *= -1.0;
        c_ = a_in / r;
        s_ = b_in / r;
        z = 1.0;
        if (fabs(a_in) > fabs(b_in)) z = s_;
        else if (fabs(b_in) >= fabs(a_in) && c_ != 0.0) z = 1.0 / c_;
    } else {
        c_ = 1.0;
        s_ = 0.0;
        r = 0.0;
        z = 0.0;
    }
    a_in = r;
    b_in = z;
}

__device__ inline void dev_applyrot(double &x, double &y, double c_, double s_)
{
    double temp = c_ * x + s_ * y;
    y = c_ * y - s_ * x;
    x = temp;
}

// Start of an outer cycle (Gmres.cxx:102-116) after outer.Finished(beta).
template <typename T>
__device__ inline void gm_outer_check(KState<T> *st, double beta)
{
    st->gm_beta = beta;
    if (it_finished(st, beta, st->nb_iter, st->max_iter)) {
        st->gm_active = 0;
        st->cyc_skip = 1;
        it_stop(st);
        return;
    }
    const int m = st->gm_m;
    for (int k = 0; k < m + 1; k++) st->gs[k] = 0.0;
    st->gs[0] = beta;
    st->gm_i = 0;
    st->gm_cycle_done = 0;
    st->gm_active = 1;
    st->cyc_skip = 0;
    st->gm_inner_nb = st->nb_iter;
    st->gm_inner_max = st->nb_iter + m;
    st->gm_scale = (T)(1.0 / beta);
    st->skip = 0;
}

template <typename T>
struct FinGmInit {  // beta = |r|; outer.Init(r); outer.Finished(beta)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const
    {
        if (!it_init(st, s[0])) {
            st->gm_active = 0;
            st->cyc_skip = 1;
            return;
        }
        gm_outer_check(st, (double)sqrt_t(s[0]));
    }
};

template <typename T>
struct FinGmOuter {  // true residual after the update (Gmres.cxx:173-178)
    KState<T> *st;
    __device__ void operator()(const T (&s)[2]) const { gm_outer_check(st, (double)sqrt_t(s[0])); }
};

template <typename T>
struct FinGmH {  // H(k, i) = (V_k, w)
    KState<T> *st;
    int k, i;
    __device__ void operator()(const T (&s)[2]) const { st->H[i * (st->gm_m + 1) + k] = (double)s[0]; }
};

// End of Arnoldi step i (Gmres.cxx:135-162).
template <typename T>
struct FinGmStep {
    KState<T> *st;
    int i;
    __device__ void operator()(const T (&s)[2]) const
    {
        const int ld = st->gm_m + 1;
        double *Hc = st->H + (size_t)i * ld;
        double hi = (double)sqrt_t(s[0]);
        st->gm_hi = hi;
        st->gm_scale = hi != 0.0 ? (T)(1.0 / hi) : (T)1;
        for (int k = 0; k < i; k++) dev_applyrot(Hc[k], Hc[k + 1], st->gc[k], st->gsn[k]);
        if (hi != 0.0) {
            dev_genrot(Hc[i], hi, st->gc[i], st->gsn[i]);
            dev_applyrot(st->gs[i], st->gs[i + 1], st->gc[i], st->gsn[i]);
        }
        st->gm_inner_nb++;
        st->nb_iter++;
        st->gm_i = i + 1;
        if (it_finished(st, fabs(st->gs[i + 1]), st->gm_inner_nb, st->gm_inner_max)) {
            st->gm_cycle_done = 1;
            st->skip = 1;
        }
    }
};

}  // namespace shk

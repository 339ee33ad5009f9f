/*
 * seldon_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of Seldon's sparse Krylov path (SpMV on the
 * RowSparse / RowSymSparse / ArrayRowSparse storages and its transposed
 * forms, the BLAS-1 steps, the Iteration stopping logic, Cg, BiCgStab,
 * Gmres, Cgs, QCgs, BiCg, TfQmr, Qmr, Cgne, Lsqr, Gcr, BiCgStabl, BiCgcr,
 * QmrSym, MinRes and Symmlq, SOR / SSOR,
 * ConvertMatrix_from_Coordinates): the oracle the HIP library's results are
 * compared with.  Cgs, QCgs, BiCg, TfQmr, Qmr, Cgne, Lsqr, Gcr, BiCgStabl, BiCgcr, QmrSym, MinRes
 * and Symmlq are pinned bitwise to
 * Seldon's own runs of them (the overlay program's CPU legs), and by closed
 * one-loop cases, BiCg following Cg on an SPD matrix, TfQmr's
 * quasi-residual bound, Qmr's carried residual, Cgne as Cg on the explicit
 * A^T A and Lsqr's residual estimate; their FP32 versions by kappa(A)
 * eps_32.  The FP32 family (orc_s_*; Gmres with FP32 vectors and
 * an FP64 Hessenberg) has no Seldon build to compare with; it is pinned by
 * what the mathematics fixes instead (tests/test_oracle.py): A = I converges
 * in one step (Gmres through the happy breakdown), a diagonal matrix with k
 * distinct values in k, the Arnoldi relation M^-1 A V = V Hbar and V's
 * orthonormality hold to FP32 rounding (with planted slips -- a dropped
 * term, a transposed Hessenberg, the preconditioner left out -- shown to
 * fail that check), the FP32 solutions lie within kappa(A) eps_32 of the
 * FP64 oracle's, and the arithmetic is float, not rounded double.  Only the
 * tests, the smoke check and the benchmark's CPU baseline leg may load it;
 * the product never does.
 *
 * Every function follows the reference file:line named in its comment, in
 * the same arithmetic order, and is built with -ffp-contract=off so that
 * each a*b+c rounds twice, as the reference's C++ does under ISO rules.
 *
 * Two families are generated from one body: orc_d_* (double vectors) and
 * orc_s_* (float vectors).  Iteration bookkeeping (tolerance, 1/|r0|,
 * residual history) is always double, matching Iteration<double>.
 */
#ifndef SELDON_ORACLE_H
#define SELDON_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ROWSPARSE = 0, ORC_ROWSYMSPARSE = 1, ORC_ARRAYROWSPARSE = 2 };
enum { ORC_PREC_IDENTITY = 0, ORC_PREC_JACOBI = 1 };

/* A sparse matrix in one of the three storages, as flat CSR arrays:
 * ROWSPARSE: full CSR; ROWSYMSPARSE: upper triangle incl. diagonal;
 * ARRAYROWSPARSE: the per-row sparse vectors laid end to end (the
 * ConvertToCSR flattening, matrix_sparse/Matrix_Conversions.cxx:3676-3705). */
typedef struct {
    int storage;
    int m, n;
    int64_t nnz;
    const int *ptr;
    const int *ind;
    const void *val; /* double* or float* by family */
} orc_mat;

/* Mirrors Iteration<double> (computation/solver/iterative/Iterative.hxx:47-108). */
typedef struct {
    double tolerance;
    double facteur_reste;
    int max_iter;
    int nb_iter;
    int error_code;
    int fail_convergence;
    int print_level; /* kept for completeness; the oracle never prints */
    int init_guess_null;
    int parameter_restart;
    /* residual history: every value Finished() compared with the
     * tolerance, in call order, with the iteration number at the call */
    double *hist;
    int *hist_it;
    int hist_len;
    int hist_cap;
    /* optional (Gmres only): the first restart cycle's Arnoldi basis and
     * Hessenberg matrix before any rotation, for the invariant
     * A M^{-1}... V_k = V_{k+1} Hbar_k.  arnoldi_v holds (m+1) vectors of n
     * values of the family's type, arnoldi_h (m+1) x m doubles, column
     * major; arnoldi_steps = the k the cycle reached. */
    double *arnoldi_h;
    void *arnoldi_v;
    int arnoldi_steps;
} orc_iter;

void orc_iter_init(orc_iter *it, int max_iter, double tol);
void orc_coo_dims(int nnz, const int *row, const int *col, int index, int m_min, int n_min, int *m, int *n);

/* ---- family-generated entry points (SFX = d | s, T = double | float) ---- */
#define ORC_DECLARE(SFX, T)                                                    \
    void orc_##SFX##_mltadd(T alpha, const orc_mat *A, const T *x, T beta,     \
                            T *y);                                             \
    void orc_##SFX##_mlt(const orc_mat *A, const T *x, T *y);                  \
    void orc_##SFX##_mltadd_trans(T alpha, const orc_mat *A, const T *x,       \
                                  T beta, T *y);                               \
    void orc_##SFX##_mlt_trans(const orc_mat *A, const T *x, T *y);            \
    void orc_##SFX##_sor(const orc_mat *A, T *x, const T *b, double omega,     \
                         int wide, int iter, int type_ssor);                   \
    void orc_##SFX##_coo_to_csr(int nnz, const int *row, const int *col,       \
                                const T *val, int index, int m, int *ptr,      \
                                int *ind, T *out);                             \
    T orc_##SFX##_dot(int n, const T *x, const T *y);                          \
    T orc_##SFX##_norm2(int n, const T *x);                                    \
    void orc_##SFX##_add(int n, T alpha, const T *x, T *y);                    \
    void orc_##SFX##_scal(int n, T alpha, T *x);                               \
    void orc_##SFX##_copy(int n, const T *x, T *y);                            \
    void orc_##SFX##_jacobi_diag(const orc_mat *A, T *diag);                   \
    int orc_##SFX##_cg(const orc_mat *A, T *x, const T *b, int precond,        \
                       const T *diag, orc_iter *it);                           \
    int orc_##SFX##_bicgstab(const orc_mat *A, T *x, const T *b, int precond,  \
                             const T *diag, orc_iter *it);                     \
    int orc_##SFX##_cgs(const orc_mat *A, T *x, const T *b, int precond,       \
                        const T *diag, orc_iter *it);                          \
    int orc_##SFX##_bicg(const orc_mat *A, T *x, const T *b, int precond,      \
                         const T *diag, orc_iter *it);                         \
    int orc_##SFX##_tfqmr(const orc_mat *A, T *x, const T *b, int precond,     \
                          const T *diag, orc_iter *it);                        \
    int orc_##SFX##_qmr(const orc_mat *A, T *x, const T *b, int precond,       \
                        const T *diag, orc_iter *it);                          \
    int orc_##SFX##_cgne(const orc_mat *A, T *x, const T *b, int precond,      \
                         const T *diag, orc_iter *it);                         \
    int orc_##SFX##_lsqr(const orc_mat *A, T *x, const T *b, int precond,      \
                         const T *diag, orc_iter *it);                         \
    int orc_##SFX##_bicgcr(const orc_mat *A, T *x, const T *b, int precond,    \
                           const T *diag, orc_iter *it);                       \
    int orc_##SFX##_qmrsym(const orc_mat *A, T *x, const T *b, int precond,    \
                           const T *diag, orc_iter *it);                       \
    int orc_##SFX##_minres(const orc_mat *A, T *x, const T *b, int precond,    \
                           const T *diag, orc_iter *it);                       \
    int orc_##SFX##_symmlq(const orc_mat *A, T *x, const T *b, int precond,    \
                           const T *diag, orc_iter *it);                       \
    int orc_##SFX##_qcgs(const orc_mat *A, T *x, const T *b, int precond,      \
                         const T *diag, orc_iter *it);                         \
    int orc_##SFX##_gcr(const orc_mat *A, T *x, const T *b, int precond,       \
                        const T *diag, orc_iter *it);                          \
    int orc_##SFX##_bicgstabl(const orc_mat *A, T *x, const T *b, int precond, \
                              const T *diag, orc_iter *it);                    \
    int orc_##SFX##_gmres(const orc_mat *A, T *x, const T *b, int precond,     \
                          const T *diag, orc_iter *it);                        \
    void orc_##SFX##_sddmm(int m, const int *ptr, const int *ind,              \
                           const T *lam, const T *x, T *g);

ORC_DECLARE(d, double)
ORC_DECLARE(s, float)

/* Back substitution on the leading n x n upper triangle of a column-major
 * array with leading dimension ld (the role of cblas_dtrsv in
 * computation/interfaces/Blas_2.cxx:3226-3238). */
void orc_trsv_upper(int n, const double *H, int ld, double *s);
/* Old-BLAS Givens generation (computation/basic_functions/
 * Functions_Vector.cxx:612-649) and application (:681-687). */
void orc_genrot(double *a, double *b, double *c, double *s);
void orc_applyrot(double *x, double *y, double c, double s);

#ifdef __cplusplus
}
#endif
#endif

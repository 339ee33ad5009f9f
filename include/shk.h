/*
 * shk.h -- C ABI of the MI355X (gfx950) sparse Krylov library "shk"
 * (Seldon HIP Krylov).
 *
 * This is the drop-in boundary for Seldon's sparse Krylov hot path.  Seldon
 * has no runtime plugin layer: its solvers are templates that reach the
 * kernels through unqualified calls (Mlt, MltAdd, DotProd, Norm2, Add, Copy,
 * Mlt(alpha, v), M.Solve) resolved at compile time
 * (computation/solver/iterative/Cg.cxx:45-116, BiCgStab.cxx:44-145,
 * Gmres.cxx:45-182; doc/guide/iterative.dox:62,223).  Each entry point below
 * replaces one of those calls (or a whole solver template) and is cited to it.
 * The Seldon-compatible C++ overloads in include/seldon_amd/ call these.
 *
 * Conventions
 *   - Every function returns an int status: SHK_OK (0) or a negative
 *     SHK_ERR_*; shk_last_error() gives the message (thread-local).  No C++
 *     exception crosses this boundary.
 *   - Vectors are raw device pointers of the matrix's value type; every call
 *     is ordered on the context's stream.  Scalars that Seldon returns by
 *     value (DotProd, Norm2) come back on the host (the call synchronises);
 *     the *_device variants leave them in device memory instead.
 *   - A context is single-threaded (Seldon is not thread-safe either,
 *     vector/Vector.hxx:50); separate contexts may be used from separate
 *     threads.  Device memory is owned by the handles.
 *   - Arithmetic order: SpMV row sums are taken sequentially in stored order,
 *     exactly as Functions_MatVect.cxx:424-430 does, and FP64/FP32 products
 *     and sums are not contracted into FMAs, so SpMV results are bitwise
 *     equal to Seldon's generic templates.  Reductions (dot, nrm2) are
 *     deterministic fixed-order trees.
 */
#ifndef SHK_H
#define SHK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHK_OK 0
#define SHK_ERR_ARG (-1)     /* bad argument (null handle, negative size, ...) */
#define SHK_ERR_HIP (-2)     /* HIP runtime error; message has the HIP text */
#define SHK_ERR_NOMEM (-3)   /* device allocation failed (Seldon NoMemory) */
#define SHK_ERR_DIM (-4)     /* dimension mismatch (Seldon WrongDim) */
#define SHK_ERR_NODEV (-5)   /* no usable gfx950 device */
#define SHK_ERR_STATE (-6)   /* call not valid in the current state */

typedef enum { SHK_F64 = 0, SHK_F32 = 1 } shk_dtype;

/* The three Seldon storages of the path (share/Storage.hxx:82,115,138). */
typedef enum {
    SHK_ROWSPARSE = 0,       /* Matrix<T,General,RowSparse>: full CSR */
    SHK_ROWSYMSPARSE = 1,    /* Matrix<T,Symmetric,RowSymSparse>: upper CSR */
    SHK_ARRAYROWSPARSE = 2   /* Matrix<T,General,ArrayRowSparse>: rows laid
                                end to end as by ConvertToCSR
                                (matrix_sparse/Matrix_Conversions.cxx:3676-3705) */
} shk_storage;

typedef enum { SHK_PREC_IDENTITY = 0, SHK_PREC_JACOBI = 1 } shk_precond;

/* Device layout of a matrix (an internal choice; results are identical).
 *   ROWS   -- the caller's CSR order, cut into row blocks;
 *   PANELS -- panels of 2048 rows whose entries are regrouped by column
 *             windows so the x entries they read are staged on chip; for
 *             rows with sorted columns near the diagonal (banded, stencil
 *             and FEM matrices in natural order; RowSymSparse upper parts
 *             and their transposes are walked one after the other);
 *   AUTO   -- PANELS where it applies, else ROWS. */
/* Device layouts of a matrix: ROWS (row blocks, the caller's order), PANELS
 * (column-window chunks of 2048-row panels), SLICED (RowSparse only: the row
 * layout plus a copy of the entries in 8192-row panels whose lanes own rows,
 * for the SpMV).  AUTO picks one from the matrix's shape. */
typedef enum { SHK_LAYOUT_AUTO = 0, SHK_LAYOUT_ROWS = 1, SHK_LAYOUT_PANELS = 2, SHK_LAYOUT_SLICED = 3 } shk_layout;

typedef struct shk_ctx_s *shk_ctx;
typedef struct shk_csr_s *shk_csr;

const char *shk_last_error(void);
const char *shk_version(void);

/* ---- context (device + stream + workspace) ----------------------------- */
int shk_ctx_create(int device, shk_ctx *out);
int shk_ctx_destroy(shk_ctx ctx);
/* Run on an external hipStream_t (e.g. torch.cuda.current_stream()), taken
 * verbatim: NULL is the default (null) stream.  shk_ctx_create starts on a
 * stream of the context's own (default priority). */
int shk_ctx_set_stream(shk_ctx ctx, void *hip_stream);
int shk_ctx_synchronize(shk_ctx ctx);
int shk_ctx_device_info(shk_ctx ctx, int *num_cus, int *num_xcds, char *arch, int arch_len);

/* ---- device memory (Allocator role: share/Allocator.cxx:34-65) --------- */
int shk_malloc(shk_ctx ctx, size_t bytes, void **dptr);
int shk_free(shk_ctx ctx, void *dptr);
int shk_memcpy_h2d(shk_ctx ctx, void *dst, const void *src, size_t bytes);
int shk_memcpy_d2h(shk_ctx ctx, void *dst, const void *src, size_t bytes);
int shk_memcpy_d2d(shk_ctx ctx, void *dst, const void *src, size_t bytes);

/* ---- sparse matrices --------------------------------------------------- */
/* Upload a matrix.  ptr/ind/val describe the storage as Seldon holds it
 * (0-based; RowSymSparse = upper triangle with diagonal).  src_on_device
 * says where the three arrays live.  The caller keeps ownership of them.
 * Replaces Matrix_Sparse::SetData / the Matrix(m,n,values,ptr,ind)
 * constructor (matrix_sparse/Matrix_Sparse.cxx:105-181, 300-412) for a
 * device-resident copy. */
int shk_csr_create(shk_ctx ctx, shk_storage storage, shk_dtype dtype, int m, int n,
                   int64_t nnz, const int *ptr, const int *ind, const void *val,
                   int src_on_device, shk_csr *out);
/* Same, with an explicit device layout (SHK_LAYOUT_PANELS and
 * SHK_LAYOUT_SLICED fail with SHK_ERR_ARG where they do not apply). */
int shk_csr_create_layout(shk_ctx ctx, shk_storage storage, shk_dtype dtype, int m, int n,
                          int64_t nnz, const int *ptr, const int *ind, const void *val,
                          int src_on_device, int layout, shk_csr *out);
int shk_csr_layout(shk_csr A, int *layout);
/* Seldon's binary CSR file (int m, n, nz; ptr[m+1]; ind[nz]; values[nz]):
 * Matrix_Sparse::Write / Read (matrix_sparse/Matrix_Sparse.cxx:1532-1556,
 * 1649-1682), the same layout for RowSymSparse's upper part
 * (Matrix_SymSparse.cxx).  shk_csr_write writes the stored arrays back in
 * the caller's order (panel layout: rows by increasing column, which that
 * layout requires); shk_csr_read creates a matrix of the given storage,
 * value type (the file's) and layout. */
int shk_csr_write(shk_ctx ctx, shk_csr A, const char *path);
int shk_csr_read(shk_ctx ctx, const char *path, shk_storage storage, shk_dtype dtype, int layout, shk_csr *out);
/* Harwell-Boeing files (matrix_sparse/IOMatrixMarket.cxx).  shk_csr_write_hb
 * is WriteHarwellBoeing(A, path) (:397-510): the file Seldon writes for the
 * same storage and arrays, byte for byte (compressed columns in ConvertToCSC's
 * order, values at 15 digits; the symmetric storage is written whole).
 * shk_csr_read_hb is ReadHarwellBoeing(path, A) (:80-360) into a ColSparse
 * matrix followed by ConvertMatrix_to_Coordinates / _from_Coordinates into a
 * RowSparse matrix of the same size: it creates that RowSparse matrix, its
 * rows assembled on the device.  Real and pattern matrices, as Seldon reads;
 * Seldon's reader cannot parse a D-format value line such as the "(3D23.15)"
 * its own writer puts, and neither can this one. */
/* Seldon's binary vector file (vector/Vector.cxx:846-874 Write, :1060-1091
 * Read): the length as an int when with_size, then the n elements of the
 * device vector dx.  shk_vec_read puts the file's length in *n (dx NULL: the
 * length only) and refuses more than n_cap elements; without the length it
 * reads n_cap elements, as Seldon's Read(stream, false) reads the vector's
 * current length. */
int shk_vec_write(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, const char *path, int with_size);
int shk_vec_read(shk_ctx ctx, shk_dtype dt, const char *path, int with_size, int64_t n_cap, void *dx, int64_t *n);
int shk_csr_write_hb(shk_ctx ctx, shk_csr A, const char *path);
int shk_csr_read_hb(shk_ctx ctx, const char *path, shk_dtype dtype, int layout, shk_csr *out);
int shk_csr_destroy(shk_csr A);
/* New values on the same pattern, in the caller's stored order (RowSymSparse:
 * upper part).  Used when A depends on parameters (the differentiable solve). */
int shk_csr_set_values(shk_ctx ctx, shk_csr A, const void *val, int src_on_device);
/* m, n, stored nnz (as the caller gave it), nnz held on device after the
 * symmetric expansion, value type, storage. */
int shk_csr_info(shk_csr A, int *m, int *n, int64_t *nnz_stored, int64_t *nnz_device,
                 int *dtype, int *storage);
/* Algorithmic bytes of one MltAdd on this matrix (what a Seldon-style CSR
 * traversal must move: stored ind+val, ptr, x, y write, + y read if beta!=0). */
double shk_csr_spmv_bytes(shk_csr A, int beta_nonzero);

/* ---- SpMV -------------------------------------------------------------- */
/* MltAdd(alpha, A, X, beta, Y): Y <- beta*Y + alpha*A*X
 *   RowSparse      Functions_MatVect.cxx:398-431 (beta multiplies, NaN kept)
 *   RowSymSparse   Functions_MatVect.cxx:526-564
 *   ArrayRowSparse matrix_sparse/Functions_MatrixArray.cxx:534-575
 *                  (beta == 0 zero-fills; alpha != 1 applied per term)
 * Trans/SeldonNoTrans for the symmetric storage forward here
 * (Functions_MatVect.cxx:905-927). */
int shk_spmv(shk_ctx ctx, shk_csr A, double alpha, const void *dx, double beta, void *dy);
/* Out-of-place MltAdd: Y_out <- beta*Y_in + alpha*A*X (Y_in may equal Y_out). */
int shk_spmv_out(shk_ctx ctx, shk_csr A, double alpha, const void *dx, double beta,
                 const void *dy_in, void *dy_out);
/* Mlt(M, X, Y): Y.Fill(0); MltAdd(1, M, X, 0, Y) (Functions_MatVect.cxx:72-81). */
int shk_mlt(shk_ctx ctx, shk_csr A, const void *dx, void *dy);
/* MltAdd(alpha, SeldonTrans, A, X, beta, Y): Y <- beta*Y + alpha*A^T*X
 *   RowSparse      Functions_MatVect.cxx:729-758 (Mlt(beta, Y), then
 *                  Y(ind[j]) += alpha * data[j] * X(i), rows i ascending)
 *   ArrayRowSparse matrix_sparse/Functions_MatrixArray.cxx:591-632
 *                  (beta == 0 zero-fills)
 *   RowSymSparse   forwards to shk_spmv (Functions_MatVect.cxx:918-927)
 * X has A's row count, Y its column count.  The first call builds A^T on
 * the device (kept with A; rebuilt after shk_csr_set_values).
 * SeldonConjTrans is the same for real values. */
int shk_spmv_trans(shk_ctx ctx, shk_csr A, double alpha, const void *dx, double beta, void *dy);
/* Mlt(SeldonTrans, M, X, Y): Y.Fill(0); MltAdd(1, SeldonTrans, M, X, 0, Y)
 * (Functions_MatVect.cxx:123-130). */
int shk_mlt_trans(shk_ctx ctx, shk_csr A, const void *dx, void *dy);
/* Fused: Y = A*X and returns (X, Y) -- the CG step q = A p, (p, q)
 * (computation/solver/iterative/Cg.cxx:97-98). */
int shk_mlt_dot(shk_ctx ctx, shk_csr A, const void *dx, void *dy, double *result);

/* ---- relaxation --------------------------------------------------------- */
/* SOR(A, X, B, omega, iter, type_ssor): type_ssor 2 forward sweeps, 3
 * backward, 0 both (all forward sweeps, then all backward ones), bitwise
 * Seldon's (matrix_sparse/Relaxation_MatVect.cxx:41-95 RowSparse, :106-167
 * ArrayRowSparse, :179-300 RowSymSparse).  Rows run level by level (no row
 * of a level reads another's X); the levels are built on the first call and
 * kept with A.  omega_type is the C++ type of Seldon's omega argument:
 * SHK_F64 with float values makes each relaxation step double, as C++
 * promotion does; SHK_F32 keeps it in float.  Seldon's SorPreconditioner
 * (computation/solver/preconditioner/Precond_Ssor.cxx:24-90) calls SOR with
 * type 0 (symmetric) or 2 / 3 and omega in its value type. */
int shk_sor(shk_ctx ctx, shk_csr A, void *dx, const void *db, double omega, shk_dtype omega_type, int iter,
            int type_ssor);

/* ---- assembly ----------------------------------------------------------- */
/* ConvertMatrix_from_Coordinates(IndRow, IndCol, Val, A, index) for RowSparse
 * (matrix_sparse/Matrix_Conversions.cxx:1288-1335), on the device: the
 * entries ordered by (row, column), duplicates kept in their input order
 * (Seldon's MergeSort is stable, vector/Functions_Arrays.cxx:412-480);
 * m = max(row - index) + 1 but at least m_min (A.GetM() before the call),
 * n likewise.  Inputs and outputs are device arrays: dptr holds m_cap + 1
 * ints, dind and dval_out nnz entries.  SHK_ERR_DIM when m > m_cap (*m then
 * says how many rows are needed); dptr = NULL only computes *m and *n.  The
 * result feeds shk_csr_create with src_on_device = 1. */
int shk_coo_to_csr(shk_ctx ctx, shk_dtype dt, int64_t nnz, const int *drow, const int *dcol, const void *dval,
                   int index, int m_min, int n_min, int m_cap, int *m, int *n, int *dptr, int *dind,
                   void *dval_out);

/* ---- BLAS-1 (computation/basic_functions/Functions_Vector.cxx) --------- */
/* DotProd / DotProdConj (:330-346, :391-397): accumulated in the vector type. */
int shk_dot(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, const void *dy, double *result);
int shk_dot_device(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, const void *dy,
                   void *dresult);
/* Norm2 (:540-549): sqrt of the sum of squares, in the vector type. */
int shk_nrm2(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, double *result);
/* Add(alpha, X, Y) (:85-105): Y += alpha X; alpha == 0 is a no-op. */
int shk_axpy(shk_ctx ctx, shk_dtype dt, int64_t n, double alpha, const void *dx, void *dy);
/* Mlt(alpha, X) (:69-74 -> vector/Vector.cxx:607-615): X *= alpha. */
int shk_scal(shk_ctx ctx, shk_dtype dt, int64_t n, double alpha, void *dx);
/* Copy(X, Y) (:240-246). */
int shk_copy(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, void *dy);
/* Vector::Fill(x) / Zero() (vector/Vector.cxx:686-713). */
int shk_fill(shk_ctx ctx, shk_dtype dt, int64_t n, double value, void *dx);
/* Fused Add + DotProd: Y += alpha X, then returns (Y, Z). */
int shk_axpy_dot(shk_ctx ctx, shk_dtype dt, int64_t n, double alpha, const void *dx,
                 void *dy, const void *dz, double *result);
/* Cast between the two value types (FP64 <-> FP32 copies for refinement). */
int shk_convert(shk_ctx ctx, int64_t n, shk_dtype src_dt, const void *src, shk_dtype dst_dt,
                void *dst);

/* ---- preconditioners (Preconditioner_Base, computation/solver/iterative/
 *      Iterative.hxx:25-39 / Iterative.cxx:56-77) -------------------------- */
/* Jacobi: z_i = r_i / a_ii, with a_ii from A (new: Seldon has no Jacobi; the
 * division form follows SOR's / a_jj, matrix_sparse/Relaxation_MatVect.cxx:79).
 * Identity (Preconditioner_Base::Solve) is shk_copy. */
int shk_jacobi(shk_ctx ctx, shk_csr A, const void *dr, void *dz);

/* ---- Iteration + solvers ----------------------------------------------- */
/* Mirrors Iteration<double> (computation/solver/iterative/Iterative.hxx:47-108).
 * In: tolerance, max_iter, nb_iter (kept when Init fails, as Seldon does),
 * init_guess_null, parameter_restart, hist/hist_it/hist_cap (optional host
 * arrays for the residual history).  Out: facteur_reste, nb_iter,
 * error_code/fail_convergence, hist_len (number of Finished() calls, which
 * may exceed hist_cap; only hist_cap are stored). */
typedef struct {
    double tolerance;
    double facteur_reste;
    int max_iter;
    int nb_iter;
    int error_code;
    int fail_convergence;
    int print_level;
    int init_guess_null;
    int parameter_restart;
    int hist_cap;
    int hist_len;
    double *hist;
    int *hist_it;
} shk_iteration;

/* Return value of the three solvers = Iteration::ErrorCode()
 * (Iterative.cxx:344-354): 0, -2 at max_iter, or the breakdown code.
 * A negative SHK_ERR_* (<= -1, != -2 ...) cannot be confused: solver status
 * is returned through *code, the function result is the call status. */
/* Cg (computation/solver/iterative/Cg.cxx:45-116).  For real values it is
 * also Seldon's CoCg (CoCg.cxx:44-115): that loop is Cg's operation for
 * operation (DotProd and DotProdConj agree on reals), so the C++ overlay
 * routes CoCg here. */
int shk_cg(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
           shk_iteration *iter, int *code);
/* shk_cg without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory during the solve (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels, arguments and results as shk_cg, but the work vectors, the
 * device residual history and the reduction partials live in a fixed 32 MiB
 * array of the library (one per device, held until the process ends) instead
 * of the context's grow-only buffers: about 1M FP64 or 2M FP32 rows with the
 * Jacobi preconditioner; a larger system returns SHK_ERR_ARG.  Calls that
 * overlap on one device race on that array. */
int shk_cg_noalloc(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
                   shk_iteration *iter, int *code);
/* BiCgStab (computation/solver/iterative/BiCgStab.cxx:44-145). */
int shk_bicgstab(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
                 shk_iteration *iter, int *code);
/* Gmres(m) (computation/solver/iterative/Gmres.cxx:45-182), m =
 * iter->parameter_restart; Hessenberg, Givens rotations and the triangular
 * solve stay in FP64 on the device for both value types. */
int shk_gmres(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
              shk_iteration *iter, int *code);
/* Cgs (computation/solver/iterative/Cgs.cxx:46-137): the loop fused on the
 * device like BiCgStab's (scalars and Finished() in device state, the host
 * reads the state once per 16 loops); Seldon's vector updates element for
 * element.  Error codes 1 / 2 are its breakdowns #1 / #2. */
int shk_cgs(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
            shk_iteration *iter, int *code);
/* BiCg (computation/solver/iterative/BiCg.cxx:47-131), fused likewise; its
 * A^T p~ products run on the transpose the library builds on first use (the
 * one behind shk_spmv_trans, Seldon's summation order).  With the Jacobi
 * preconditioner, z~ = r~ / a_ii as z = r / a_ii. */
int shk_bicg(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
             shk_iteration *iter, int *code);
/* TfQmr (computation/solver/iterative/TfQmr.cxx:45-240), fused likewise
 * (left preconditioned, as the template: r0 = M^{-1} (b - A x), the
 * stopping test on kappa = tau sqrt(2 iter + 1 or 2)); its theta, tau and c
 * are double for both value types, as Titer is in the template.  Error
 * codes 1-5 are its breakdowns (sigma = 0, alpha = 0, tau = 0 in either
 * half step, rho0 = 0). */
int shk_tfqmr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
              shk_iteration *iter, int *code);
/* Qmr (computation/solver/iterative/Qmr.cxx:42-190), fused likewise; its
 * A^T q products run on the library's device-built transpose; rho and xi
 * are double for both value types (Titer in the template), the other
 * scalars the value type.  Error codes 1-6 are its breakdowns #1-#6. */
int shk_qmr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
            shk_iteration *iter, int *code);
/* Cgne (computation/solver/iterative/Cgne.cxx:41-124), fused likewise: Cg
 * on A^T A x = A^T b with M M^T as the preconditioner, the A^T products on
 * the device-built transpose; two iterations counted per loop. */
int shk_cgne(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
             shk_iteration *iter, int *code);
/* Lsqr (computation/solver/iterative/Lsqr.cxx:41-150), fused likewise.  As
 * the template, it applies no preconditioner (prec is accepted and
 * ignored) and restarts from x = 0 whatever the initial guess (which only
 * sets the first residual norm).  Error codes 1-3 are its breakdowns. */
int shk_lsqr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
             shk_iteration *iter, int *code);
/* BiCgcr (computation/solver/iterative/BiCgcr.cxx:45-117), fused likewise:
 * one product per iteration, (w, v), (w, a) and (w, z) from fused
 * reductions; the start's product counts as the first iteration, as in
 * the template.  Error codes 1 / 2 are its breakdowns. */
int shk_bicgcr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
               shk_iteration *iter, int *code);
/* QmrSym (computation/solver/iterative/QmrSym.cxx:42-166), fused likewise:
 * Qmr for symmetric A and M, no transposed products; rho double, the other
 * scalars the value type.  Error codes 1, 3-6 are its breakdowns. */
int shk_qmrsym(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
               shk_iteration *iter, int *code);
/* MinRes (computation/solver/iterative/MinRes.cxx:47-151), fused likewise:
 * the stopping test on the residual estimate np (|b| times the product of
 * the rotations' |s|, double as Titer in the template); the other scalars
 * the value type.  Error codes 1 / 2 are its breakdowns. */
int shk_minres(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
               shk_iteration *iter, int *code);
/* Symmlq (computation/solver/iterative/Symmlq.cxx:45-161), fused likewise:
 * the stopping test on np = s_prod / |c| (double, as Titer), the last
 * x += ceta_bar w_bar after the loop, as the template; it reports no
 * breakdown (error code 0 or -2). */
int shk_symmlq(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
               shk_iteration *iter, int *code);
/* QCgs (computation/solver/iterative/QCgs.cxx:47-164), fused likewise: Cgs
 * with the quasi-minimal-residual smoothing step (its five dot products in
 * one reduction), the stopping test on the smoothed residual, and x =
 * M^{-1} x_qcgs at the end, as the template.  Error codes 1-3 are its
 * breakdowns. */
int shk_qcgs(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
             shk_iteration *iter, int *code);
/* Gcr (computation/solver/iterative/Gcr.cxx:45-139), fused likewise,
 * restarted every m = iter->parameter_restart iterations; both its outer and
 * inner stopping tests enter the history, as their Finished() calls do in
 * Seldon.  Error code 1 is its breakdown (w_j = 0), which here ends the
 * solve (Seldon's outer loop would go on). */
int shk_gcr(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
            shk_iteration *iter, int *code);
/* BiCgStabl (computation/solver/iterative/BiCgStabl.cxx:43-180), fused
 * likewise, with l = iter->parameter_restart (1 .. 100) BiCG steps per cycle
 * and right preconditioning (x is mapped by M^{-1} at the end, as there).
 * Error codes 1 and 2 are its breakdowns (rho_0 = 0, sigma = 0), which here
 * end the solve (Seldon's loop would go on after the cycle's MR part). */
int shk_bicgstabl(shk_ctx ctx, shk_csr A, void *dx, const void *db, shk_precond prec,
                  shk_iteration *iter, int *code);

/* ---- FP32 inner solve + FP64 iterative refinement (no counterpart in
 *      Seldon; its Gmres is the inner solver) ------------------------------ */
typedef struct {
    double tolerance;       /* on ||b - A x||_2 / ||b||_2, true FP64 residual */
    int max_refinements;
    int inner_restart;      /* GMRES(m) restart of the FP32 inner solve */
    int inner_max_iter;     /* per refinement step */
    double inner_tolerance; /* relative, on the preconditioned FP32 residual */
    int inner_method;       /* 0 = GMRES, 1 = CG */
    /* out */
    int refinements;
    int total_inner_iter;
    double final_relres;
    double *relres_hist;    /* optional, max_refinements + 1 entries */
} shk_ir_params;
/* A64 and A32 hold the same pattern (FP64 and FP32 values).  x, b: FP64. */
int shk_solve_ir(shk_ctx ctx, shk_csr A64, shk_csr A32, void *dx, const void *db,
                 shk_precond prec, shk_ir_params *p, int *code);

/* ---- adjoint gradient on the pattern (no counterpart in Seldon) -------- */
/* g_j = -(lam_i * x_col(j)) for every stored entry j of row i: dL/dA_ij for
 * x = A^{-1} b and adjoint lam = A^{-T} dL/dx.  For RowSymSparse the value of
 * an upper entry (i,j), i != j, also stands for (j,i):
 * g = -(lam_i x_j + lam_j x_i). */
int shk_sddmm(shk_ctx ctx, shk_csr A, const void *dlam, const void *dx, void *dg);

/* ---- batched CG over many independent systems (no counterpart) --------- */
/* `count` systems stacked block-diagonally in one CSR: system s owns rows
 * [row_off[s], row_off[s+1]).  One persistent kernel runs every CG (identity
 * preconditioner, Seldon's Cg recurrences per system) to tol relative to
 * ||b_s||; per-system iterations and codes come back in its[], codes[].
 * group = threads per system work-group (64..1024, power of two; 0 = 512);
 * rows_per_pass = rows a thread keeps in flight per pass (1, 2, 4 or 8;
 * 0 = 4). */
int shk_cg_batched(shk_ctx ctx, shk_csr A, int count, const int *row_off_host,
                   void *dx, const void *db, double tol, int max_iter, int group,
                   int rows_per_pass, int use_graph, int *its, int *codes);
/* Mixed variant: FP32 inner CG per system with FP64 residual refinement
 * in the same persistent kernel; A64/A32 share the pattern. */
int shk_cg_batched_ir(shk_ctx ctx, shk_csr A64, shk_csr A32, int count,
                      const int *row_off_host, void *dx, const void *db, double tol,
                      int max_iter, double inner_tol, int group, int rows_per_pass,
                      int use_graph, int *its, int *codes);

/* ---- timing of the last launch of a kernel family (HIP events on the
 *      stream the kernel ran on) ---------------------------------------- */
int shk_time_spmv(shk_ctx ctx, shk_csr A, int mode, const void *dx, void *dy, int reps,
                  double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif

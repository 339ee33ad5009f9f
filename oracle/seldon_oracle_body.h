/*
 * seldon_oracle_body.h -- TEST INFRASTRUCTURE ONLY (see seldon_oracle.h).
 *
 * Type-generic body, included by seldon_oracle.c once per family with
 *   T   = double | float      (vector / matrix value type)
 *   SFX = d      | s
 * Scalars of the Krylov recurrences are T, as Seldon's `Complexe`
 * (= Vector1::value_type) is; the Iteration bookkeeping is double.
 * The GMRES Hessenberg matrix, its rotations and right-hand side are double
 * in both families (for T = double this is exactly Seldon's arithmetic; for
 * T = float it is the FP32-inner definition used by the mixed-precision
 * solver, whose vector arithmetic stays in float).
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(CAT(orc_, SFX), CAT(_, name))

/* ---- BLAS-1, generic templates --------------------------------------- */

/* DotProd: computation/basic_functions/Functions_Vector.cxx:330-346
 * (DotProdConj forwards to it for real types, :391-397). */
T FN(dot)(int n, const T *x, const T *y)
{
    T value = 0;
    for (int i = 0; i < n; i++) value += x[i] * y[i];
    return value;
}

/* Norm2: Functions_Vector.cxx:540-549 (unscaled sqrt of the sum). */
T FN(norm2)(int n, const T *x)
{
    T value = 0;
    for (int i = 0; i < n; i++) value += x[i] * x[i];
    return (T)sqrt(value);
}

/* Add(alpha, X, Y): Functions_Vector.cxx:85-105; alpha == 0 is a no-op. */
void FN(add)(int n, T alpha, const T *x, T *y)
{
    if (alpha != (T)0) {
        T alpha_ = alpha;
        for (int i = 0; i < n; i++) y[i] += alpha_ * x[i];
    }
}

/* Mlt(alpha, X) -> X *= alpha: Functions_Vector.cxx:69-74,
 * vector/Vector.cxx:607-615. */
void FN(scal)(int n, T alpha, T *x)
{
    for (int i = 0; i < n; i++) x[i] *= alpha;
}

/* Copy(X, Y): Functions_Vector.cxx:240-246 -> Vector.cxx:579-586. */
void FN(copy)(int n, const T *x, T *y)
{
    memcpy(y, x, sizeof(T) * (size_t)n);
}

static void FN(fill)(int n, T v, T *x)
{
    for (int i = 0; i < n; i++) x[i] = v; /* Vector.cxx:707-713 */
}

static T *FN(dup)(int n, const T *x)
{
    T *y = (T *)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(y, x, sizeof(T) * (size_t)n);
    return y;
}

/* ---- SpMV ---------------------------------------------------------- */

void FN(mltadd)(T alpha, const orc_mat *A, const T *x, T beta, T *y)
{
    const int m = A->m;
    const int *ptr = A->ptr, *ind = A->ind;
    const T *data = (const T *)A->val;
    if (A->storage == ORC_ROWSPARSE) {
        /* computation/basic_functions/Functions_MatVect.cxx:398-431 */
        FN(scal)(m, beta, y);
        for (int i = 0; i < m; i++) {
            T temp = 0;
            for (int j = ptr[i]; j < ptr[i + 1]; j++) temp += data[j] * x[ind[j]];
            y[i] += alpha * temp;
        }
    } else if (A->storage == ORC_ROWSYMSPARSE) {
        /* Functions_MatVect.cxx:526-564: gather over the stored upper part,
         * then the transposed scatter of the strict upper part. */
        FN(scal)(m, beta, y);
        for (int i = 0; i < m; i++) {
            T temp = 0;
            for (int j = ptr[i]; j < ptr[i + 1]; j++) temp += data[j] * x[ind[j]];
            y[i] += alpha * temp;
        }
        for (int i = 0; i < m - 1; i++)
            for (int j = ptr[i]; j < ptr[i + 1]; j++)
                if (ind[j] != i) y[ind[j]] += alpha * data[j] * x[i];
    } else {
        /* matrix_sparse/Functions_MatrixArray.cxx:534-575 */
        if (beta == (T)0) FN(fill)(m, (T)0, y);
        else FN(scal)(m, beta, y);
        if (alpha == (T)1) {
            for (int i = 0; i < m; i++)
                for (int j = ptr[i]; j < ptr[i + 1]; j++) y[i] += data[j] * x[ind[j]];
        } else {
            for (int i = 0; i < m; i++)
                for (int j = ptr[i]; j < ptr[i + 1]; j++)
                    y[i] += alpha * data[j] * x[ind[j]];
        }
    }
}

/* Mlt(M, X, Y): Functions_MatVect.cxx:72-81 = Fill(0) then MltAdd(1,..,0,..). */
void FN(mlt)(const orc_mat *A, const T *x, T *y)
{
    FN(fill)(A->m, (T)0, y);
    FN(mltadd)((T)1, A, x, (T)0, y);
}

/* MltAdd(alpha, SeldonTrans, M, X, beta, Y): Y <- beta Y + alpha M^T X.
 * RowSparse: computation/basic_functions/Functions_MatVect.cxx:729-758
 * (Mlt(beta, Y), then the scatter over the rows in order).  ArrayRowSparse:
 * matrix_sparse/Functions_MatrixArray.cxx:591-632.  RowSymSparse forwards
 * to the plain product (Functions_MatVect.cxx:918-927).  x has m entries,
 * y has n. */
void FN(mltadd_trans)(T alpha, const orc_mat *A, const T *x, T beta, T *y)
{
    const int m = A->m, n = A->n;
    const int *ptr = A->ptr, *ind = A->ind;
    const T *data = (const T *)A->val;
    if (A->storage == ORC_ROWSYMSPARSE) {
        FN(mltadd)(alpha, A, x, beta, y);
    } else if (A->storage == ORC_ROWSPARSE) {
        FN(scal)(n, beta, y);
        for (int i = 0; i < m; i++)
            for (int j = ptr[i]; j < ptr[i + 1]; j++) y[ind[j]] += alpha * data[j] * x[i];
    } else {
        if (beta == (T)0) FN(fill)(n, (T)0, y);
        else FN(scal)(n, beta, y);
        if (alpha == (T)1) {
            for (int i = 0; i < m; i++)
                for (int j = ptr[i]; j < ptr[i + 1]; j++) y[ind[j]] += data[j] * x[i];
        } else {
            for (int i = 0; i < m; i++)
                for (int j = ptr[i]; j < ptr[i + 1]; j++) y[ind[j]] += alpha * data[j] * x[i];
        }
    }
}

/* Mlt(SeldonTrans, M, X, Y): Functions_MatVect.cxx:123-130 fills Y with 0
 * and calls MltAdd(1, Trans, M, X, 0, Y) with Trans a SeldonTranspose
 * reference, which selects the generic MltAdd of :1279-1316 (the
 * storage-specific transposed products take the class_SeldonTrans tag):
 * Y.Fill(0), then for every column i, temp = sum over j = 0 .. m-1 of
 * M(j, i) * X(j) -- absent entries read as 0 (Matrix_Sparse / SymSparse
 * operator(), RowSymSparse swapping to the stored upper entry) -- and
 * Y(i) += 1 * temp. */
void FN(mlt_trans)(const orc_mat *A, const T *x, T *y)
{
    const int ma = A->m, na = A->storage == ORC_ROWSYMSPARSE ? A->m : A->n;
    const int *ptr = A->ptr, *ind = A->ind;
    const T *data = (const T *)A->val;
    const int sym = A->storage == ORC_ROWSYMSPARSE;
    /* the matrix's entries M(j, i) as (j, i, v), the symmetric ones mirrored */
    long nz = 0;
    for (int r = 0; r < ma; r++)
        for (int k = ptr[r]; k < ptr[r + 1]; k++) nz += (sym && ind[k] != r) ? 2 : 1;
    int *er = (int *)malloc(sizeof(int) * (size_t)(nz + 1)), *ec = (int *)malloc(sizeof(int) * (size_t)(nz + 1));
    T *ev = (T *)malloc(sizeof(T) * (size_t)(nz + 1));
    long q = 0;
    for (int r = 0; r < ma; r++)
        for (int k = ptr[r]; k < ptr[r + 1]; k++) {
            er[q] = r; ec[q] = ind[k]; ev[q] = data[k]; q++;
            if (sym && ind[k] != r) { er[q] = ind[k]; ec[q] = r; ev[q] = data[k]; q++; }
        }
    /* column lists in increasing j: stable counting sorts by j, then by i */
    long *cnt = (long *)calloc((size_t)(ma > na ? ma : na) + 2, sizeof(long));
    long *t1 = (long *)malloc(sizeof(long) * (size_t)(nz + 1)), *t2 = (long *)malloc(sizeof(long) * (size_t)(nz + 1));
    for (long k = 0; k < nz; k++) cnt[er[k] + 1]++;
    for (int r = 0; r < ma; r++) cnt[r + 1] += cnt[r];
    for (long k = 0; k < nz; k++) t1[cnt[er[k]]++] = k;
    memset(cnt, 0, sizeof(long) * ((size_t)(ma > na ? ma : na) + 2));
    for (long k = 0; k < nz; k++) cnt[ec[k] + 1]++;
    for (int c = 0; c < na; c++) cnt[c + 1] += cnt[c];
    long *cptr = (long *)malloc(sizeof(long) * ((size_t)na + 1));
    for (int c = 0; c <= na; c++) cptr[c] = cnt[c];
    for (long k = 0; k < nz; k++) { long e = t1[k]; t2[cnt[ec[e]]++] = e; }
    FN(fill)(na, (T)0, y);
    for (int i = 0; i < na; i++) {
        T temp = 0;
        long k = cptr[i];
        for (int j = 0; j < ma; j++) {
            T mji = 0;
            if (k < cptr[i + 1] && er[t2[k]] == j) mji = ev[t2[k++]];
            temp += mji * x[j];
        }
        y[i] += (T)1 * temp;
    }
    free(er); free(ec); free(ev); free(cnt); free(t1); free(t2); free(cptr);
}

/* Seldon's A(j, j) on RowSparse (matrix_sparse/Matrix_Sparse.cxx:1054-1084):
 * the scan moves on while the column is below j and stops at the row's last
 * entry; the entry it stops on is returned if its column is j, else 0.  On a
 * row with unsorted columns that can miss a diagonal entry stored after a
 * larger column. */
static T FN(rowsparse_diag)(const int *ptr, const int *ind, const T *data, int j)
{
    const int first = ptr[j], last = ptr[j + 1] - 1;
    if (last < first) return 0;
    int k = first;
    while (k < last && ind[k] < j) k++;
    return ind[k] == j ? data[k] : (T)0;
}

/* (1 - omega) x_j + omega temp / a_jj, in the type of (value, omega):
 * omega's own when it is double and the values float (C++ promotion in
 * Relaxation_MatVect.cxx:80 with T3 = double), the value type otherwise. */
static T FN(sor_relax)(T xj, T temp, T ajj, double omega, int wide)
{
    if (wide) return (T)((1.0 - omega) * (double)xj + omega * (double)temp / (double)ajj);
    const T om = (T)omega;
    return ((T)1 - om) * xj + om * temp / ajj;
}

/* RowSymSparse (matrix_sparse/Relaxation_MatVect.cxx:179-300), upper triangle
 * stored.  Forward: x_j = (1-w)/w a_jj x_j + (b_j - sum_{p>j} a_jp x_p) for
 * every j with the old x, then the lower solve by columns (x_j *= w / a_jj,
 * then x_p -= a_jp x_j for the row's other entries).  Backward: y_p +=
 * a_jp x_j (old x) and x_j = (1-w)/w a_jj x_j + b_j - y_j row by row, then the
 * upper solve in decreasing j: x_j = (x_j - sum_p a_jp x_p) w / a_jj.  a_jj is
 * 0 plus the row's diagonal entries. */
static void FN(sor_sym)(const orc_mat *A, T *x, const T *b, double omega, int wide, int iter, int type_ssor)
{
    const int ma = A->m;
    const int *ptr = A->ptr, *ind = A->ind;
    const T *data = (const T *)A->val;
    const T omt = (T)omega;
    if (type_ssor % 2 == 0)
        for (int it = 0; it < iter; it++) {
            for (int j = 0; j < ma; j++) {
                T temp = 0, ajj = 0;
                for (int k = ptr[j]; k < ptr[j + 1]; k++) {
                    if (ind[k] == j) ajj += data[k];
                    else temp += data[k] * x[ind[k]];
                }
                temp = b[j] - temp;
                x[j] = wide ? (T)((1.0 - omega) / omega * (double)ajj * (double)x[j] + (double)temp)
                            : ((T)1 - omt) / omt * ajj * x[j] + temp;
            }
            for (int j = 0; j < ma; j++) {
                T ajj = 0;
                for (int k = ptr[j]; k < ptr[j + 1]; k++)
                    if (ind[k] == j) ajj += data[k];
                x[j] = wide ? (T)((double)x[j] * (omega / (double)ajj)) : x[j] * (omt / ajj);
                for (int k = ptr[j]; k < ptr[j + 1]; k++)
                    if (ind[k] != j) x[ind[k]] -= data[k] * x[j];
            }
        }
    if (type_ssor % 3 == 0) {
        T *y = (T *)malloc(sizeof(T) * (size_t)(ma > 0 ? ma : 1));
        for (int it = 0; it < iter; it++) {
            for (int j = 0; j < ma; j++) y[j] = 0;
            for (int j = 0; j < ma; j++) {
                T ajj = 0;
                for (int k = ptr[j]; k < ptr[j + 1]; k++) {
                    if (ind[k] == j) ajj += data[k];
                    else y[ind[k]] += data[k] * x[j];
                }
                x[j] = wide ? (T)((1.0 - omega) / omega * (double)ajj * (double)x[j] + (double)b[j] - (double)y[j])
                            : ((T)1 - omt) / omt * ajj * x[j] + b[j] - y[j];
            }
            for (int j = ma - 1; j >= 0; j--) {
                T temp = 0, ajj = 0;
                for (int k = ptr[j]; k < ptr[j + 1]; k++) {
                    if (ind[k] == j) ajj += data[k];
                    else temp += data[k] * x[ind[k]];
                }
                x[j] = wide ? (T)((double)(x[j] - temp) * omega / (double)ajj) : (x[j] - temp) * omt / ajj;
            }
        }
        free(y);
    }
}

/* SOR(A, X, B, omega, iter, type_ssor): matrix_sparse/Relaxation_MatVect.cxx:
 * 41-95 (RowSparse: a_jj = A(j, j)), :106-167 (ArrayRowSparse: a_jj = 0 plus
 * the row's diagonal entries), :179-300 (RowSymSparse, above).  type_ssor % 2
 * == 0: iter forward sweeps; then type_ssor % 3 == 0: iter backward sweeps.
 * wide: omega is a double applied to float values. */
void FN(sor)(const orc_mat *A, T *x, const T *b, double omega, int wide, int iter, int type_ssor)
{
    if (A->storage == ORC_ROWSYMSPARSE) {
        FN(sor_sym)(A, x, b, omega, wide, iter, type_ssor);
        return;
    }
    const int ma = A->m;
    const int *ptr = A->ptr, *ind = A->ind;
    const T *data = (const T *)A->val;
    for (int d = 0; d < 2; d++) {
        if (d == 0 && type_ssor % 2 != 0) continue;
        if (d == 1 && type_ssor % 3 != 0) continue;
        for (int it = 0; it < iter; it++)
            for (int q = 0; q < ma; q++) {
                const int j = d == 0 ? q : ma - 1 - q;
                T temp = 0, ajj = 0;
                if (A->storage == ORC_ARRAYROWSPARSE) {
                    for (int k = ptr[j]; k < ptr[j + 1]; k++) {
                        temp += data[k] * x[ind[k]];
                        if (ind[k] == j) ajj += data[k];
                    }
                } else {
                    ajj = FN(rowsparse_diag)(ptr, ind, data, j);
                    for (int k = ptr[j]; k < ptr[j + 1]; k++) temp += data[k] * x[ind[k]];
                }
                temp = b[j] - temp + ajj * x[j];
                x[j] = FN(sor_relax)(x[j], temp, ajj, omega, wide);
            }
    }
}

/* MergeSort(m, n, tab1, tab2, tab3) of vector/Functions_Arrays.cxx:412-480:
 * bottom-up, runs of inc, the right run's element taken only when strictly
 * smaller (so the sort is stable); tab1 the keys. */
static void FN(mergesort3)(int lo, int hi, int *tab1, int *tab2, T *tab3)
{
    if (lo >= hi) return;
    const int len = hi - lo + 1;
    int *t1 = (int *)malloc(sizeof(int) * (size_t)len), *t2 = (int *)malloc(sizeof(int) * (size_t)len);
    T *t3 = (T *)malloc(sizeof(T) * (size_t)len);
    int inc = 1, ind = 0, current, i, j, sup;
    while (inc < len) {
        for (i = lo; i <= hi - inc; i += 2 * inc) {
            ind = i;
            current = i + inc;
            sup = i + 2 * inc;
            if (sup >= hi + 1) sup = hi + 1;
            j = i;
            while (j < i + inc) {
                if (current < sup) {
                    while (current < sup && tab1[j] > tab1[current]) {
                        t1[ind - lo] = tab1[current]; t2[ind - lo] = tab2[current]; t3[ind - lo] = tab3[current];
                        current++; ind++;
                    }
                    t1[ind - lo] = tab1[j]; t2[ind - lo] = tab2[j]; t3[ind - lo] = tab3[j];
                    ind++; j++;
                    if (j == i + inc)
                        for (j = current; j < sup; j++) {
                            t1[ind - lo] = tab1[j]; t2[ind - lo] = tab2[j]; t3[ind - lo] = tab3[j];
                            ind++;
                        }
                } else {
                    for (current = j; current < i + inc; current++) {
                        t1[ind - lo] = tab1[current]; t2[ind - lo] = tab2[current]; t3[ind - lo] = tab3[current];
                        ind++;
                    }
                    j = current + 1;
                }
            }
        }
        for (i = lo; i < ind; i++) { tab1[i] = t1[i - lo]; tab2[i] = t2[i - lo]; tab3[i] = t3[i - lo]; }
        inc *= 2;
    }
    free(t1); free(t2); free(t3);
}

/* ConvertMatrix_from_Coordinates(IndRow, IndCol, Val, A, index) for RowSparse:
 * matrix_sparse/Matrix_Conversions.cxx:1288-1335.  Shifts the indices by
 * `index`, m = max(row) + 1 but at least m_min (A.GetM()), n likewise (the
 * dimensions come from orc_coo_dims), Sort(IndRow, IndCol, Val) by rows,
 * counts Ptr, then Sort(Ptr(i), Ptr(i + 1) - 1, IndCol, Val) per row.
 * ptr has m + 1 entries; ind / out have nnz. */
void FN(coo_to_csr)(int nnz, const int *row, const int *col, const T *val, int index, int m, int *ptr, int *ind,
                    T *out)
{
    int *r = (int *)malloc(sizeof(int) * (size_t)(nnz + 1));
    for (int i = 0; i < nnz; i++) {
        r[i] = row[i] - index;
        ind[i] = col[i] - index;
        out[i] = val[i];
    }
    FN(mergesort3)(0, nnz - 1, r, ind, out);
    for (int i = 0; i <= m; i++) ptr[i] = 0;
    for (int i = 0; i < nnz; i++) ptr[r[i] + 1]++;
    for (int i = 0; i < m; i++) ptr[i + 1] += ptr[i];
    /* Sort(Ptr(i), Ptr(i + 1) - 1, IndCol, Val): the two-array MergeSort
     * (:337-410) is the same merge on (IndCol, Val); the row array is
     * carried along here unchanged (constant within the row) */
    for (int i = 0; i < m; i++) FN(mergesort3)(ptr[i], ptr[i + 1] - 1, ind, r, out);
    free(r);
}

/* Diagonal for the Jacobi preconditioner (an entry with ind == row; 0 when
 * the row stores none). */
void FN(jacobi_diag)(const orc_mat *A, T *diag)
{
    const T *data = (const T *)A->val;
    for (int i = 0; i < A->m; i++) {
        diag[i] = 0;
        for (int j = A->ptr[i]; j < A->ptr[i + 1]; j++)
            if (A->ind[j] == i) diag[i] = data[j];
    }
}

/* Preconditioner application M.Solve(A, r, z):
 * identity = Copy(r, z) (computation/solver/iterative/Iterative.cxx:60-65);
 * Jacobi  = z_i = r_i / a_ii (the division form of SOR's /a_jj,
 * matrix_sparse/Relaxation_MatVect.cxx:79). */
static void FN(precond)(int precond, int n, const T *diag, const T *r, T *z)
{
    if (precond == ORC_PREC_JACOBI) {
        for (int i = 0; i < n; i++) z[i] = r[i] / diag[i];
    } else {
        FN(copy)(n, r, z);
    }
}

/* Iteration::Init (Iterative.cxx:235-249): 1/|r| or -1 for a null rhs. */
static int FN(iter_init)(orc_iter *it, int n, const T *r)
{
    double norme_rhs = (double)FN(norm2)(n, r);
    if (norme_rhs == 0.0) return -1;
    it->facteur_reste = 1.0 / norme_rhs;
    it->nb_iter = 0;
    return 0;
}

/* ---- Cg: computation/solver/iterative/Cg.cxx:45-116 ---------------- */
int FN(cg)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
           orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho = 1, rho_1 = 1, alpha, beta, delta;
    T *p = FN(dup)(N, b), *q = FN(dup)(N, b), *r = FN(dup)(N, b), *z = FN(dup)(N, b);
    (void)rho;
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        FN(precond)(precond, N, diag, r, z);
        rho = FN(dot)(N, r, z);
        if (rho == (T)0) { orc_fail(it, 1); break; }
        if (it->nb_iter == 0) {
            FN(copy)(N, z, p);
        } else {
            beta = rho / rho_1;
            FN(scal)(N, beta, p);
            FN(add)(N, (T)1, z, p);
        }
        FN(mlt)(A, p, q);
        delta = FN(dot)(N, p, q);
        if (delta == (T)0) { orc_fail(it, 2); break; }
        alpha = rho / delta;
        FN(add)(N, alpha, p, x);
        FN(add)(N, -alpha, q, r);
        rho_1 = rho;
        it->nb_iter++;
    }
done:
    free(p); free(q); free(r); free(z);
    return orc_error_code(it);
}

/* ---- BiCgStab: computation/solver/iterative/BiCgStab.cxx:44-145 ---- */
int FN(bicgstab)(const orc_mat *A, T *x, const T *b, int precond,
                 const T *diag, orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho_1, rho_2 = 0, alpha = 0, beta, omega = 0, sigma;
    T *p = FN(dup)(N, b), *phat = FN(dup)(N, b), *s = FN(dup)(N, b),
      *shat = FN(dup)(N, b), *t = FN(dup)(N, b), *v = FN(dup)(N, b),
      *r = FN(dup)(N, b), *rtilde = FN(dup)(N, b);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(copy)(N, r, rtilde);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        rho_1 = FN(dot)(N, rtilde, r);
        if (rho_1 == (T)0) { orc_fail(it, 1); break; }
        if (it->nb_iter == 0) {
            FN(copy)(N, r, p);
        } else {
            if (omega == (T)0) { orc_fail(it, 2); break; }
            beta = (rho_1 / rho_2) * (alpha / omega);
            FN(add)(N, -omega, v, p);
            FN(scal)(N, beta, p);
            FN(add)(N, (T)1, r, p);
        }
        FN(precond)(precond, N, diag, p, phat);
        FN(mlt)(A, phat, v);
        sigma = FN(dot)(N, rtilde, v);
        if (sigma == (T)0) { orc_fail(it, 3); break; }
        alpha = rho_1 / sigma;
        FN(copy)(N, r, s);
        FN(add)(N, -alpha, v, s);
        it->nb_iter++;
        if (orc_finished(it, (double)FN(norm2)(N, s), it->nb_iter, it->max_iter)) {
            FN(add)(N, alpha, phat, x);
            break;
        }
        FN(precond)(precond, N, diag, s, shat);
        FN(mlt)(A, shat, t);
        omega = FN(dot)(N, t, s) / FN(dot)(N, t, t);
        FN(add)(N, alpha, phat, x);
        FN(add)(N, omega, shat, x);
        FN(copy)(N, s, r);
        FN(add)(N, -omega, t, r);
        rho_2 = rho_1;
        it->nb_iter++;
    }
done:
    free(p); free(phat); free(s); free(shat); free(t); free(v); free(r); free(rtilde);
    return orc_error_code(it);
}

/* ---- Cgs: computation/solver/iterative/Cgs.cxx:46-137 -------------- */
int FN(cgs)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
            orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho_1, rho_2 = 0, alpha, beta, delta;
    T *p = FN(dup)(N, b), *phat = FN(dup)(N, b), *q = FN(dup)(N, b),
      *qhat = FN(dup)(N, b), *vhat = FN(dup)(N, b), *u = FN(dup)(N, b),
      *uhat = FN(dup)(N, b), *r = FN(dup)(N, b), *rtilde = FN(dup)(N, b);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(copy)(N, r, rtilde);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        rho_1 = FN(dot)(N, rtilde, r);
        if (rho_1 == (T)0) { orc_fail(it, 1); break; }
        if (it->nb_iter == 0) {
            FN(copy)(N, r, u);
            FN(copy)(N, u, p);
        } else {
            /* u = r + beta q; p = beta (beta p + q) + u */
            beta = rho_1 / rho_2;
            FN(copy)(N, r, u);
            FN(add)(N, beta, q, u);
            FN(scal)(N, beta, p);
            FN(add)(N, (T)1, q, p);
            FN(scal)(N, beta, p);
            FN(add)(N, (T)1, u, p);
        }
        FN(precond)(precond, N, diag, p, phat);
        FN(mlt)(A, phat, vhat);
        it->nb_iter++;
        delta = FN(dot)(N, rtilde, vhat);
        if (delta == (T)0) { orc_fail(it, 2); break; }
        alpha = rho_1 / delta;
        FN(copy)(N, u, q);
        FN(add)(N, -alpha, vhat, q);
        FN(add)(N, (T)1, q, u);
        FN(precond)(precond, N, diag, u, uhat);
        FN(add)(N, alpha, uhat, x);
        FN(mlt)(A, uhat, qhat);
        FN(add)(N, -alpha, qhat, r);
        rho_2 = rho_1;
        it->nb_iter++;
    }
done:
    free(p); free(phat); free(q); free(qhat); free(vhat); free(u); free(uhat);
    free(r); free(rtilde);
    return orc_error_code(it);
}

/* ---- BiCg: computation/solver/iterative/BiCg.cxx:47-131 ------------ *
 * M.TransSolve of the identity is Copy, of the Jacobi preconditioner the
 * same division by a_ii (A^T has A's diagonal). */
int FN(bicg)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
             orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho_1, rho_2 = 0, alpha, beta, delta;
    T *r = FN(dup)(N, b), *z = FN(dup)(N, b), *p = FN(dup)(N, b), *q = FN(dup)(N, b),
      *r_tilde = FN(dup)(N, b), *z_tilde = FN(dup)(N, b), *p_tilde = FN(dup)(N, b),
      *q_tilde = FN(dup)(N, b);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(copy)(N, r, r_tilde);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        FN(precond)(precond, N, diag, r, z);
        FN(precond)(precond, N, diag, r_tilde, z_tilde);
        rho_1 = FN(dot)(N, z, r_tilde);
        if (rho_1 == (T)0) { orc_fail(it, 1); break; }
        if (it->nb_iter == 0) {
            FN(copy)(N, z, p);
            FN(copy)(N, z_tilde, p_tilde);
        } else {
            beta = rho_1 / rho_2;
            FN(scal)(N, beta, p);
            FN(add)(N, (T)1, z, p);
            FN(scal)(N, beta, p_tilde);
            FN(add)(N, (T)1, z_tilde, p_tilde);
        }
        FN(mlt)(A, p, q);
        it->nb_iter++;
        FN(mlt_trans)(A, p_tilde, q_tilde);
        delta = FN(dot)(N, p_tilde, q);
        if (delta == (T)0) { orc_fail(it, 2); break; }
        alpha = rho_1 / delta;
        FN(add)(N, alpha, p, x);
        FN(add)(N, -alpha, q, r);
        FN(add)(N, -alpha, q_tilde, r_tilde);
        rho_2 = rho_1;
        it->nb_iter++;
    }
done:
    free(r); free(z); free(p); free(q); free(r_tilde); free(z_tilde); free(p_tilde);
    free(q_tilde);
    return orc_error_code(it);
}

/* ---- TfQmr: computation/solver/iterative/TfQmr.cxx:45-240 ----------- *
 * Scalars typed as in the template: sigma, alpha, beta, eta, rho, rho0 are
 * the value type T; c, kappa, tau, theta are Titer = double.  Mlt(s, d)
 * with a double s scales each T element in double and rounds back
 * (`d(i) *= s`); eta = c * c * alpha is formed in double and stored as T. */
static void FN(scal_d)(int n, double s, T *x)
{
    for (int i = 0; i < n; i++) x[i] *= s;
}

int FN(tfqmr)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
              orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T *tmp = FN(dup)(N, b), *r0 = FN(dup)(N, b), *v = FN(dup)(N, b), *h = FN(dup)(N, b),
      *w = FN(dup)(N, b), *y1 = FN(dup)(N, b), *g = FN(dup)(N, b), *y0 = FN(dup)(N, b),
      *rtilde = FN(dup)(N, b), *d = FN(dup)(N, b);
    T sigma, alpha, beta, eta, rho, rho0;
    double c, kappa, tau, theta;
    FN(fill)(N, (T)0, tmp);
    /* 1. r0 = M^{-1} (b - A x) */
    FN(copy)(N, b, tmp);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, tmp);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(precond)(precond, N, diag, tmp, r0);
    if (FN(iter_init)(it, N, r0) != 0) goto done;
    /* 2.-8. */
    FN(copy)(N, r0, w);
    FN(copy)(N, r0, y1);
    FN(copy)(N, y1, v);
    FN(mlt)(A, v, tmp);
    FN(precond)(precond, N, diag, tmp, v);
    it->nb_iter++;
    FN(copy)(N, v, g);
    FN(fill)(N, (T)0, d);
    tau = (double)FN(norm2)(N, r0);
    theta = 0.0;
    eta = (T)0;
    FN(copy)(N, r0, rtilde);
    rho = FN(dot)(N, rtilde, r0);
    rho0 = rho;
    it->nb_iter = 0;
    for (;;) {
        /* 9.-11. */
        sigma = FN(dot)(N, rtilde, v);
        if (sigma == (T)0) { orc_fail(it, 1); break; }
        alpha = rho / sigma;
        FN(copy)(N, y1, y0);
        FN(add)(N, -alpha, v, y0);
        /* 12. h = M^{-1} A y0 */
        FN(copy)(N, y0, h);
        FN(mlt)(A, h, tmp);
        FN(precond)(precond, N, diag, tmp, h);
        /* first half: 13., 18., 14.-17., 19.-21. */
        FN(add)(N, -alpha, g, w);
        if (alpha == (T)0) { orc_fail(it, 2); break; }
        FN(scal_d)(N, theta * theta * (double)eta / (double)alpha, d);
        FN(add)(N, (T)1, y1, d);
        if (tau == 0.0) { orc_fail(it, 3); break; }
        theta = (double)FN(norm2)(N, w) / tau;
        c = 1.0 / sqrt(1.0 + theta * theta);
        tau = tau * c * theta;
        eta = (T)(c * c * (double)alpha);
        FN(add)(N, eta, d, x);
        kappa = tau * sqrt(2. * (it->nb_iter + 1));
        if (orc_finished(it, kappa, it->nb_iter, it->max_iter)) break;
        it->nb_iter++;
        FN(copy)(N, h, g);
        /* second half */
        FN(add)(N, -alpha, g, w);
        FN(scal_d)(N, theta * theta * (double)eta / (double)alpha, d);
        FN(add)(N, (T)1, y0, d);
        if (tau == 0.0) { orc_fail(it, 4); break; }
        theta = (double)FN(norm2)(N, w) / tau;
        c = 1.0 / sqrt(1.0 + theta * theta);
        tau = tau * c * theta;
        eta = (T)(c * c * (double)alpha);
        FN(add)(N, eta, d, x);
        kappa = tau * sqrt(2. * (it->nb_iter + 1) + 1.);
        if (orc_finished(it, kappa, it->nb_iter, it->max_iter)) break;
        /* 22.-26. */
        rho0 = rho;
        rho = FN(dot)(N, rtilde, w);
        if (rho0 == (T)0) { orc_fail(it, 5); break; }
        beta = rho / rho0;
        FN(copy)(N, w, y1);
        FN(add)(N, beta, y0, y1);
        FN(copy)(N, y1, g);
        FN(mlt)(A, g, tmp);
        FN(precond)(precond, N, diag, tmp, g);
        FN(scal)(N, beta * beta, v);
        FN(add)(N, beta, h, v);
        FN(add)(N, (T)1, g, v);
        it->nb_iter++;
    }
done:
    free(tmp); free(r0); free(v); free(h); free(w); free(y1); free(g); free(y0);
    free(rtilde); free(d);
    return orc_error_code(it);
}

/* ---- Qmr: computation/solver/iterative/Qmr.cxx:42-190 --------------- *
 * Scalars typed as in the template: rho, rho_1, xi are Titer = double;
 * delta, ep, beta, theta, gamma, eta (and their _1 copies) are the value
 * type T, each expression evaluated in C++'s types (a double operand
 * makes the operation double) and rounded where it is stored.  d and s
 * start as copies of b (Vector1 d(b), s(b)); the branch that would copy p
 * into d is never taken (First() is tested after ++iter). */
int FN(qmr)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
            orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T delta, ep = 0, beta, theta_1, gamma_1, theta = 0, gamma = 1, eta = -1;
    double rho, rho_1, xi;
    T *r = FN(dup)(N, b), *y = FN(dup)(N, b), *z_tld = FN(dup)(N, b), *v = FN(dup)(N, b),
      *w = FN(dup)(N, b), *p_tld = FN(dup)(N, b), *p = FN(dup)(N, b), *q = FN(dup)(N, b),
      *d = FN(dup)(N, b), *s = FN(dup)(N, b);
    FN(fill)(N, (T)0, r);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(copy)(N, r, v);
    FN(precond)(precond, N, diag, v, y);
    rho = (double)FN(norm2)(N, y);
    FN(copy)(N, r, w);
    xi = (double)FN(norm2)(N, w);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        if (rho == 0.0) { orc_fail(it, 1); break; }
        if (xi == 0.0) { orc_fail(it, 2); break; }
        FN(scal)(N, (T)(1. / rho), v);
        FN(scal)(N, (T)(1. / rho), y);
        FN(scal)(N, (T)(1. / xi), w);
        delta = FN(dot)(N, w, y);
        if (delta == (T)0) { orc_fail(it, 3); break; }
        FN(precond)(precond, N, diag, w, z_tld);  /* TransSolve */
        if (it->nb_iter == 0) {
            FN(copy)(N, y, p);
            FN(copy)(N, z_tld, q);
        } else {
            FN(scal)(N, (T)(-(xi * (double)delta / (double)ep)), p);
            FN(add)(N, (T)1, y, p);
            FN(scal)(N, (T)(-(rho * (double)delta / (double)ep)), q);
            FN(add)(N, (T)1, z_tld, q);
        }
        FN(mlt)(A, p, p_tld);
        ep = FN(dot)(N, q, p_tld);
        if (ep == (T)0) { orc_fail(it, 4); break; }
        beta = ep / delta;
        if (beta == (T)0) { orc_fail(it, 5); break; }
        FN(scal)(N, -beta, v);
        FN(add)(N, (T)1, p_tld, v);
        FN(precond)(precond, N, diag, v, y);
        rho_1 = rho;
        rho = (double)FN(norm2)(N, y);
        FN(mlt_trans)(A, q, z_tld);
        FN(scal)(N, -beta, w);
        FN(add)(N, (T)1, z_tld, w);
        xi = (double)FN(norm2)(N, w);
        gamma_1 = gamma;
        theta_1 = theta;
        it->nb_iter++;
        theta = (T)(rho / (double)(gamma_1 * beta));
        gamma = (T)((double)(T)1 / sqrt(1.0 + (double)(theta * theta)));
        if (gamma == (T)0) { orc_fail(it, 6); break; }
        eta = (T)((double)(-eta) * rho_1 * (double)gamma * (double)gamma /
                  (double)(beta * gamma_1 * gamma_1));
        {
            const T tmp = theta_1 * theta_1 * gamma * gamma;
            FN(scal)(N, tmp, d);
            FN(add)(N, eta, p, d);
            FN(scal)(N, tmp, s);
            FN(add)(N, eta, p_tld, s);
        }
        FN(add)(N, (T)1, d, x);
        FN(add)(N, -(T)1, s, r);
        it->nb_iter++;
    }
done:
    free(r); free(y); free(z_tld); free(v); free(w); free(p_tld); free(p); free(q);
    free(d); free(s);
    return orc_error_code(it);
}

/* ---- Cgne: computation/solver/iterative/Cgne.cxx:41-124 ------------ *
 * Cg on A^T A x = A^T b with M M^T as the preconditioner: z = M^{-1}
 * M^{-T} r; the stopping test on |r| (dp, Titer = double), first on
 * |A^T b|; two iterations counted per loop. */
int FN(cgne)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
             orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho = 1, rho_1 = 0, alpha, beta, delta;
    double dp;
    T *p = FN(dup)(N, b), *q = FN(dup)(N, b), *r = FN(dup)(N, b), *z = FN(dup)(N, b);
    FN(mlt_trans)(A, b, q);
    if (FN(iter_init)(it, N, q) != 0) goto done;
    if (!it->init_guess_null) {
        FN(mlt)(A, x, p);
        FN(mlt_trans)(A, p, r);
        FN(scal)(N, (T)-1, r);
        FN(add)(N, (T)1, q, r);
    } else {
        FN(copy)(N, q, r);
        memset(x, 0, sizeof(T) * (size_t)N);
    }
    dp = (double)FN(norm2)(N, q);
    it->nb_iter = 0;
    while (!orc_finished(it, dp, it->nb_iter, it->max_iter)) {
        FN(precond)(precond, N, diag, r, q);  /* TransSolve */
        FN(precond)(precond, N, diag, q, z);
        rho = FN(dot)(N, r, z);
        if (rho == (T)0) { orc_fail(it, 1); break; }
        if (it->nb_iter == 0) {
            FN(copy)(N, z, p);
        } else {
            beta = rho / rho_1;
            FN(scal)(N, beta, p);
            FN(add)(N, (T)1, z, p);
        }
        FN(mlt)(A, p, q);
        FN(mlt_trans)(A, q, z);
        delta = FN(dot)(N, p, z);
        if (delta == (T)0) { orc_fail(it, 2); break; }
        alpha = rho / delta;
        FN(add)(N, alpha, p, x);
        FN(add)(N, -alpha, z, r);
        rho_1 = rho;
        dp = (double)FN(norm2)(N, r);
        it->nb_iter += 2;
    }
done:
    free(p); free(q); free(r); free(z);
    return orc_error_code(it);
}

/* ---- Lsqr: computation/solver/iterative/Lsqr.cxx:41-150 ------------ *
 * Golub-Kahan bidiagonalisation.  The template never applies M, starts
 * from x = 0 whatever the guess (only the first residual norm sees x),
 * and types its scalars as: rho, rho_bar, phi, phi_bar, theta, c, s, tmp
 * the value type T; beta, alpha, rnorm Titer = double. */
int FN(lsqr)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
             orc_iter *it)
{
    (void)precond; (void)diag;
    const int N = A->m;
    if (N <= 0) return 0;
    T rho, rho_bar, phi, phi_bar, theta, c, s, tmp;
    double beta, alpha, rnorm;
    T *v = FN(dup)(N, b), *v1 = FN(dup)(N, b), *u = FN(dup)(N, b), *u1 = FN(dup)(N, b),
      *w = FN(dup)(N, b);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, u);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, u);
    else memset(x, 0, sizeof(T) * (size_t)N);
    rnorm = (double)FN(norm2)(N, u);
    FN(copy)(N, b, u);
    beta = (double)FN(norm2)(N, u);
    tmp = (T)(1.0 / beta);
    FN(scal)(N, tmp, u);
    FN(mlt_trans)(A, u, v);
    alpha = (double)FN(norm2)(N, v);
    tmp = (T)(1.0 / alpha);
    FN(scal)(N, tmp, v);
    FN(copy)(N, v, w);
    memset(x, 0, sizeof(T) * (size_t)N);
    phi_bar = (T)beta;
    rho_bar = (T)alpha;
    it->nb_iter = 0;
    while (!orc_finished(it, rnorm, it->nb_iter, it->max_iter)) {
        FN(mlt)(A, v, u1);
        FN(add)(N, (T)(-alpha), u, u1);
        beta = (double)FN(norm2)(N, u1);
        if (beta == 0.0) { orc_fail(it, 1); break; }
        tmp = (T)(1.0 / beta);
        FN(scal)(N, tmp, u1);
        FN(mlt_trans)(A, u1, v1);
        FN(add)(N, (T)(-beta), v, v1);
        alpha = (double)FN(norm2)(N, v1);
        if (alpha == 0.0) { orc_fail(it, 2); break; }
        tmp = (T)(1.0 / alpha);
        FN(scal)(N, tmp, v1);
        rho = (T)sqrt((double)(rho_bar * rho_bar) + beta * beta);
        if (rho == (T)0) { orc_fail(it, 3); break; }
        c = rho_bar / rho;
        s = (T)(beta / (double)rho);
        theta = (T)((double)s * alpha);
        rho_bar = (T)((double)(-c) * alpha);
        phi = c * phi_bar;
        phi_bar = s * phi_bar;
        tmp = phi / rho;
        FN(add)(N, tmp, w, x);
        tmp = -theta / rho;
        FN(scal)(N, tmp, w);
        FN(add)(N, (T)1, v1, w);
        rnorm = (double)(phi_bar < 0 ? -phi_bar : phi_bar);
        { T *t_ = u1; u1 = u; u = t_; }
        { T *t_ = v1; v1 = v; v = t_; }
        it->nb_iter += 2;
    }
done:
    free(v); free(v1); free(u); free(u1); free(w);
    return orc_error_code(it);
}

/* ---- BiCgcr: computation/solver/iterative/BiCgcr.cxx:45-117 --------- */
int FN(bicgcr)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
               orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho, mu, alpha, beta, tau;
    T *v = FN(dup)(N, b), *w = FN(dup)(N, b), *s = FN(dup)(N, b), *z = FN(dup)(N, b),
      *p = FN(dup)(N, b), *a = FN(dup)(N, b);
    FN(fill)(N, (T)0, v); FN(fill)(N, (T)0, w); FN(fill)(N, (T)0, s);
    FN(fill)(N, (T)0, z); FN(fill)(N, (T)0, p); FN(fill)(N, (T)0, a);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, v);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, v);
    else memset(x, 0, sizeof(T) * (size_t)N);
    it->nb_iter = 0;
    FN(precond)(precond, N, diag, v, s);
    FN(copy)(N, s, p);
    FN(mlt)(A, p, a);
    FN(precond)(precond, N, diag, a, w);
    it->nb_iter++;
    while (!orc_finished(it, (double)FN(norm2)(N, v), it->nb_iter, it->max_iter)) {
        rho = FN(dot)(N, w, v);
        mu = FN(dot)(N, w, a);
        if (mu == (T)0) { orc_fail(it, 1); break; }
        alpha = rho / mu;
        FN(add)(N, alpha, p, x);
        FN(add)(N, -alpha, a, v);
        FN(add)(N, -alpha, w, s);
        FN(mlt)(A, s, z);
        tau = FN(dot)(N, w, z);
        if (tau == (T)0) { orc_fail(it, 2); break; }
        beta = tau / mu;
        FN(scal)(N, -beta, p);
        FN(add)(N, (T)1, s, p);
        FN(scal)(N, -beta, a);
        FN(add)(N, (T)1, z, a);
        FN(precond)(precond, N, diag, a, w);
        it->nb_iter++;
    }
done:
    free(v); free(w); free(s); free(z); free(p); free(a);
    return orc_error_code(it);
}

/* ---- QmrSym: computation/solver/iterative/QmrSym.cxx:42-166 ---------- *
 * Qmr for a symmetric A and M (no transposed products); scalar types as
 * in Qmr (rho, rho_1 double; the others the value type).  Here First() is
 * tested before ++iter, so the first loop starts d and s from p, p~. */
int FN(qmrsym)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
               orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T delta, ep = 0, beta, theta_1, gamma_1, theta = 0, gamma = 1, eta = -1;
    double rho, rho_1;
    T *r = FN(dup)(N, b), *y = FN(dup)(N, b), *v = FN(dup)(N, b), *p_tld = FN(dup)(N, b),
      *p = FN(dup)(N, b), *d = FN(dup)(N, b), *s = FN(dup)(N, b);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(copy)(N, r, v);
    FN(precond)(precond, N, diag, v, y);
    rho = (double)FN(norm2)(N, y);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        if (rho == 0.0) { orc_fail(it, 1); break; }
        FN(scal)(N, (T)(1. / rho), v);
        FN(scal)(N, (T)(1. / rho), y);
        delta = FN(dot)(N, v, y);
        if (delta == (T)0) { orc_fail(it, 3); break; }
        if (it->nb_iter == 0) {
            FN(copy)(N, y, p);
        } else {
            FN(scal)(N, (T)(-(rho * (double)delta / (double)ep)), p);
            FN(add)(N, (T)1, y, p);
        }
        FN(mlt)(A, p, p_tld);
        ep = FN(dot)(N, p, p_tld);
        if (ep == (T)0) { orc_fail(it, 4); break; }
        beta = ep / delta;
        if (beta == (T)0) { orc_fail(it, 5); break; }
        FN(scal)(N, -beta, v);
        FN(add)(N, (T)1, p_tld, v);
        FN(precond)(precond, N, diag, v, y);
        rho_1 = rho;
        rho = (double)FN(norm2)(N, y);
        gamma_1 = gamma;
        theta_1 = theta;
        theta = (T)(rho / (double)(gamma_1 * beta));
        gamma = (T)((double)(T)1 / sqrt(1.0 + (double)(theta * theta)));
        if (gamma == (T)0) { orc_fail(it, 6); break; }
        eta = (T)((double)(-eta) * rho_1 * (double)gamma * (double)gamma /
                  (double)(beta * gamma_1 * gamma_1));
        if (it->nb_iter == 0) {
            FN(copy)(N, p, d);
            FN(scal)(N, eta, d);
            FN(copy)(N, p_tld, s);
            FN(scal)(N, eta, s);
        } else {
            const T tmp = theta_1 * theta_1 * gamma * gamma;
            FN(scal)(N, tmp, d);
            FN(add)(N, eta, p, d);
            FN(scal)(N, tmp, s);
            FN(add)(N, eta, p_tld, s);
        }
        FN(add)(N, (T)1, d, x);
        FN(add)(N, -(T)1, s, r);
        it->nb_iter++;
    }
done:
    free(r); free(y); free(v); free(p_tld); free(p); free(d); free(s);
    return orc_error_code(it);
}

/* ---- MinRes: computation/solver/iterative/MinRes.cxx:47-151 ---------- *
 * All scalars the value type T except the residual estimate np (Titer =
 * double); the stopping test is on np, which starts at |b| and is
 * multiplied by |s| each iteration. */
int FN(minres)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
               orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T *u_old = FN(dup)(N, b), *u = FN(dup)(N, b), *r = FN(dup)(N, b), *v_old = FN(dup)(N, b),
      *v = FN(dup)(N, b), *w_old = FN(dup)(N, b), *w = FN(dup)(N, b), *z = FN(dup)(N, b),
      *w_oold = FN(dup)(N, b);
    T dp, beta, ibeta, beta_old, alpha, eta, ceta;
    T cold, coold, c, soold, sold, s, rho0, rho1, rho2, rho3;
    double np;
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(fill)(N, (T)0, u_old); FN(fill)(N, (T)0, v_old); FN(fill)(N, (T)0, w_old);
    FN(fill)(N, (T)0, w); FN(fill)(N, (T)0, w_oold);
    FN(precond)(precond, N, diag, r, z);
    dp = FN(dot)(N, r, z);
    dp = (T)sqrt(dp);
    beta = dp;
    eta = beta;
    FN(copy)(N, r, v);
    FN(copy)(N, z, u);
    ibeta = (T)(1.0 / (double)beta);
    FN(scal)(N, ibeta, v);
    FN(scal)(N, ibeta, u);
    c = 1; s = 0; cold = 1; sold = 0;
    np = (double)FN(norm2)(N, b);
    it->nb_iter = 0;
    while (!orc_finished(it, np, it->nb_iter, it->max_iter)) {
        FN(mlt)(A, u, r);
        alpha = FN(dot)(N, r, u);
        FN(precond)(precond, N, diag, r, z);
        FN(add)(N, -alpha, v, r);
        FN(add)(N, -alpha, u, z);
        FN(add)(N, -beta, v_old, r);
        FN(add)(N, -beta, u_old, z);
        beta_old = beta;
        dp = FN(dot)(N, r, z);
        beta = (T)sqrt(dp);
        coold = cold; cold = c; soold = sold; sold = s;
        rho0 = cold * alpha - coold * sold * beta_old;
        rho1 = (T)sqrt(rho0 * rho0 + beta * beta);
        rho2 = sold * alpha + coold * cold * beta_old;
        rho3 = soold * beta_old;
        if (rho1 == (T)0) { orc_fail(it, 1); break; }
        c = rho0 / rho1;
        s = beta / rho1;
        FN(copy)(N, w_old, w_oold);
        FN(copy)(N, w, w_old);
        FN(copy)(N, u, w);
        FN(add)(N, -rho2, w_old, w);
        FN(add)(N, -rho3, w_oold, w);
        FN(scal)(N, (T)(1. / (double)rho1), w);
        ceta = c * eta;
        FN(add)(N, ceta, w, x);
        eta = -s * eta;
        FN(copy)(N, v, v_old);
        FN(copy)(N, u, u_old);
        FN(copy)(N, r, v);
        FN(copy)(N, z, u);
        if (beta == (T)0) { orc_fail(it, 2); break; }
        ibeta = (T)(1.0 / (double)beta);
        FN(scal)(N, ibeta, v);
        FN(scal)(N, ibeta, u);
        np *= (double)(s < 0 ? -s : s);
        it->nb_iter++;
    }
done:
    free(u_old); free(u); free(r); free(v_old); free(v); free(w_old); free(w); free(z); free(w_oold);
    return orc_error_code(it);
}

/* ---- Symmlq: computation/solver/iterative/Symmlq.cxx:45-161 ---------- *
 * All scalars the value type T except np and s_prod (Titer = double); the
 * last step x += ceta_bar w_bar follows the loop, whatever ended it. */
int FN(symmlq)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
               orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T alpha, beta, ibeta, beta_old, beta1, ceta = 0, ceta_oold, ceta_old, ceta_bar;
    T c, cold, s, sold, coold, soold, rho0, rho1, rho2, rho3, dp;
    double np, s_prod;
    T *r = FN(dup)(N, b), *z = FN(dup)(N, b), *u = FN(dup)(N, b), *v = FN(dup)(N, b),
      *w = FN(dup)(N, b), *u_old = FN(dup)(N, b), *v_old = FN(dup)(N, b), *w_bar = FN(dup)(N, b);
    FN(fill)(N, (T)0, u_old); FN(fill)(N, (T)0, v_old); FN(fill)(N, (T)0, w); FN(fill)(N, (T)0, w_bar);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    ceta_oold = 0; ceta_old = 0;
    c = 1; cold = 1; s = 0; sold = 0;
    FN(precond)(precond, N, diag, r, z);
    dp = FN(dot)(N, r, z);
    dp = (T)sqrt(dp);
    beta = dp;
    beta1 = beta;
    s_prod = (double)(beta1 < 0 ? -beta1 : beta1);
    FN(copy)(N, r, v);
    FN(copy)(N, z, u);
    ibeta = (T)(1.0 / (double)beta);
    FN(scal)(N, ibeta, v);
    FN(scal)(N, ibeta, u);
    FN(copy)(N, u, w_bar);
    np = (double)FN(norm2)(N, b);
    it->nb_iter = 0;
    while (!orc_finished(it, np, it->nb_iter, it->max_iter)) {
        if (it->nb_iter != 0) {
            FN(copy)(N, v, v_old);
            FN(copy)(N, u, u_old);
            ibeta = (T)(1.0 / (double)beta);
            FN(copy)(N, r, v);
            FN(scal)(N, ibeta, v);
            FN(copy)(N, z, u);
            FN(scal)(N, ibeta, u);
            FN(copy)(N, w_bar, w);
            FN(scal)(N, c, w);
            FN(add)(N, s, u, w);
            FN(scal)(N, -s, w_bar);
            FN(add)(N, c, u, w_bar);
            FN(add)(N, ceta, w, x);
            ceta_oold = ceta_old;
            ceta_old = ceta;
        }
        FN(mlt)(A, u, r);
        alpha = FN(dot)(N, u, r);
        FN(precond)(precond, N, diag, r, z);
        FN(add)(N, -alpha, v, r);
        FN(add)(N, -alpha, u, z);
        FN(add)(N, -beta, v_old, r);
        FN(add)(N, -beta, u_old, z);
        beta_old = beta;
        dp = FN(dot)(N, r, z);
        beta = (T)sqrt(dp);
        coold = cold; cold = c; soold = sold; sold = s;
        rho0 = cold * alpha - coold * sold * beta_old;
        rho1 = (T)sqrt(rho0 * rho0 + beta * beta);
        rho2 = sold * alpha + coold * cold * beta_old;
        rho3 = soold * beta_old;
        c = rho0 / rho1;
        s = beta / rho1;
        if (it->nb_iter == 0) ceta = beta1 / rho1;
        else ceta = -(rho2 * ceta_old + rho3 * ceta_oold) / rho1;
        s_prod *= (double)(s < 0 ? -s : s);
        if (c == (T)0) np = s_prod * 1e16;
        else np = s_prod / (double)(c < 0 ? -c : c);
        it->nb_iter++;
    }
    if (c == (T)0) ceta_bar = (T)((double)ceta * 1e15);
    else ceta_bar = ceta / c;
    FN(add)(N, ceta_bar, w_bar, x);
done:
    free(r); free(z); free(u); free(v); free(w); free(u_old); free(v_old); free(w_bar);
    return orc_error_code(it);
}

/* ---- QCgs: computation/solver/iterative/QCgs.cxx:47-164 ------------- *
 * Cgs with a quasi-minimal-residual smoothing step: the 2 x 2 system of
 * (u, u), (v, v), (u, v) with u = r_qcgs - r; the stopping test is on
 * r_qcgs; after the loop x = M^{-1} x_qcgs, as the template has it. */
int FN(qcgs)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
             orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    T rho_1, rho_2, mu, nu, alpha, beta, sigma, delta;
    T b00, b11, b10, m00, m11, m10, rn0, rn1;
    T *p = FN(dup)(N, b), *q = FN(dup)(N, b), *r = FN(dup)(N, b), *rtilde = FN(dup)(N, b),
      *u = FN(dup)(N, b), *phat = FN(dup)(N, b), *r_qcgs = FN(dup)(N, b), *x_qcgs = FN(dup)(N, b),
      *v = FN(dup)(N, b);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, r);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(copy)(N, r, rtilde);
    rho_1 = (T)1;
    FN(fill)(N, (T)0, q);
    FN(fill)(N, (T)0, p);
    FN(copy)(N, r, r_qcgs);
    FN(copy)(N, x, x_qcgs);
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r_qcgs), it->nb_iter, it->max_iter)) {
        rho_2 = FN(dot)(N, rtilde, r);
        if (rho_1 == (T)0) { orc_fail(it, 1); break; }
        beta = rho_2 / rho_1;
        FN(copy)(N, r, u);
        FN(add)(N, beta, q, u);
        FN(scal)(N, beta, p);
        FN(add)(N, (T)1, q, p);
        FN(scal)(N, beta, p);
        FN(add)(N, (T)1, u, p);
        FN(precond)(precond, N, diag, p, phat);
        FN(mlt)(A, phat, v);
        it->nb_iter++;
        sigma = FN(dot)(N, rtilde, v);
        if (sigma == (T)0) { orc_fail(it, 2); break; }
        alpha = rho_2 / sigma;
        FN(copy)(N, u, q);
        FN(add)(N, -alpha, v, q);
        FN(add)(N, (T)1, q, u);
        FN(add)(N, alpha, u, x);
        FN(precond)(precond, N, diag, u, phat);
        FN(mlt)(A, phat, u);
        FN(add)(N, -alpha, u, r);
        FN(copy)(N, r_qcgs, u);
        FN(add)(N, -(T)1, r, u);
        b00 = FN(dot)(N, u, u);
        b11 = FN(dot)(N, v, v);
        b10 = FN(dot)(N, u, v);
        delta = b00 * b11 - b10 * b10;
        if (delta == (T)0) { orc_fail(it, 3); break; }
        m00 = b11 / delta;
        m11 = b00 / delta;
        m10 = -b10 / delta;
        rn0 = -FN(dot)(N, u, r);
        rn1 = -FN(dot)(N, v, r);
        mu = m00 * rn0 + m10 * rn1;
        nu = m10 * rn0 + m11 * rn1;
        FN(copy)(N, r, r_qcgs);
        FN(add)(N, mu, u, r_qcgs);
        FN(add)(N, nu, v, r_qcgs);
        FN(scal)(N, mu, x_qcgs);
        FN(add)(N, (T)1 - mu, x, x_qcgs);
        FN(add)(N, -nu, p, x_qcgs);
        rho_1 = rho_2;
        it->nb_iter++;
    }
    FN(precond)(precond, N, diag, x_qcgs, x);
done:
    free(p); free(q); free(r); free(rtilde); free(u); free(phat); free(r_qcgs); free(x_qcgs); free(v);
    return orc_error_code(it);
}

/* ---- Gcr: computation/solver/iterative/Gcr.cxx:45-139 --------------- *
 * Restarted every m = parameter_restart iterations; the inner Iteration is
 * the outer one with nb_iter's limit set to (its start + m), and both
 * loops' Finished() calls enter the history.  Seldon's outer loop keeps
 * going after breakdown #1 (Fail does not end it); here the solve ends
 * there. */
int FN(gcr)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
            orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    const int m = it->parameter_restart;
    T **p = (T **)malloc(sizeof(T *) * (size_t)(m + 1)), **w = (T **)malloc(sizeof(T *) * (size_t)(m + 1));
    T *beta = (T *)calloc((size_t)(m + 1), sizeof(T));
    for (int i = 0; i <= m; i++) {
        p[i] = FN(dup)(N, b); FN(fill)(N, (T)0, p[i]);
        w[i] = FN(dup)(N, b); FN(fill)(N, (T)0, w[i]);
    }
    T *r = FN(dup)(N, b), *q = FN(dup)(N, b), *u = FN(dup)(N, b);
    T alpha, delta;
    double normr;
    int failed = 0;
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, b, u);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, u);
    else memset(x, 0, sizeof(T) * (size_t)N);
    FN(precond)(precond, N, diag, u, r);
    normr = (double)FN(norm2)(N, r);
    it->nb_iter = 0;
    while (!failed && !orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, it->max_iter)) {
        const int inner_max = it->nb_iter + m;
        FN(copy)(N, r, p[0]);
        for (int i = 0; i < N; i++) p[0][i] *= 1.0 / normr;  /* Mlt(Titer(1) / normr, p[0]) */
        int j = 0;
        while (!orc_finished(it, (double)FN(norm2)(N, r), it->nb_iter, inner_max)) {
            FN(mlt)(A, p[j], u);
            FN(precond)(precond, N, diag, u, w[j]);
            beta[j] = FN(dot)(N, w[j], w[j]);
            if (beta[j] == (T)0) { orc_fail(it, 1); failed = 1; break; }
            alpha = FN(dot)(N, w[j], r) / beta[j];
            FN(add)(N, alpha, p[j], x);
            FN(add)(N, -alpha, w[j], r);
            it->nb_iter++;
            FN(mlt)(A, r, u);
            FN(precond)(precond, N, diag, u, q);
            FN(copy)(N, r, p[j + 1]);
            for (int i = 0; i <= j; i++) {
                delta = -FN(dot)(N, w[i], q) / beta[i];
                FN(add)(N, delta, p[i], p[j + 1]);
            }
            it->nb_iter++;
            j++;
        }
        normr = (double)FN(norm2)(N, r);
    }
done:
    for (int i = 0; i <= m; i++) { free(p[i]); free(w[i]); }
    free(p); free(w); free(beta); free(r); free(q); free(u);
    return orc_error_code(it);
}

/* ---- BiCgStabl: computation/solver/iterative/BiCgStabl.cxx:43-180 -- *
 * l = parameter_restart BiCG steps per cycle, then the MR part's modified
 * Gram-Schmidt on r_1 .. r_l, the solve for gamma and the update; right
 * preconditioning (x is mapped by M^{-1} at the end).  tau and gamma' keep
 * their values from earlier cycles where a zero gamma(i) skips them, as
 * there.  Seldon's loop goes on after a breakdown (Fail only ends the BiCG
 * part); here the solve ends there, before the MR part. */
int FN(bicgstabl)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
                  orc_iter *it)
{
    const int N = A->m;
    if (N <= 0) return 0;
    const int l = it->parameter_restart, ld = l + 1;
    T **r = (T **)malloc(sizeof(T *) * (size_t)ld), **u = (T **)malloc(sizeof(T *) * (size_t)ld);
    for (int i = 0; i <= l; i++) {
        r[i] = FN(dup)(N, b); FN(fill)(N, (T)0, r[i]);
        u[i] = FN(dup)(N, b); FN(fill)(N, (T)0, u[i]);
    }
    T *q = FN(dup)(N, b), *r0 = FN(dup)(N, b);
    T *gamma = (T *)calloc((size_t)ld, sizeof(T)), *gp = (T *)calloc((size_t)ld, sizeof(T));
    T *gt = (T *)calloc((size_t)ld, sizeof(T)), *tau = (T *)calloc((size_t)ld * ld, sizeof(T));
    T rho_0, rho_1 = 0, alpha, beta, omega, sigma;
    int failed = 0;
    FN(copy)(N, b, r[0]);
    if (!it->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, r[0]);
    else memset(x, 0, sizeof(T) * (size_t)N);
    if (FN(iter_init)(it, N, b) != 0) goto done;
    FN(copy)(N, r[0], r0);
    rho_0 = 1; alpha = 0; omega = 1;
    it->nb_iter = 0;
    while (!orc_finished(it, (double)FN(norm2)(N, r[0]), it->nb_iter, it->max_iter)) {
        rho_0 *= -omega;
        for (int j = 0; j < l; j++) {
            rho_1 = FN(dot)(N, r[j], r0);
            if (rho_0 == (T)0) { orc_fail(it, 1); failed = 1; break; }
            beta = alpha * (rho_1 / rho_0);
            rho_0 = rho_1;
            for (int i = 0; i <= j; i++) {
                FN(scal)(N, -beta, u[i]);
                FN(add)(N, (T)1, r[i], u[i]);
            }
            FN(precond)(precond, N, diag, u[j], q);
            FN(mlt)(A, q, u[j + 1]);
            it->nb_iter++;
            sigma = FN(dot)(N, u[j + 1], r0);
            if (sigma == (T)0) { orc_fail(it, 2); failed = 1; break; }
            alpha = rho_1 / sigma;
            FN(add)(N, alpha, u[0], x);
            for (int i = 0; i <= j; i++) FN(add)(N, -alpha, u[i + 1], r[i]);
            FN(precond)(precond, N, diag, r[j], q);
            FN(mlt)(A, q, r[j + 1]);
            it->nb_iter++;
        }
        if (failed) break;
        for (int j = 1; j <= l; j++) {
            for (int i = 1; i < j; i++)
                if (gamma[i] != (T)0) {
                    tau[i * ld + j] = FN(dot)(N, r[j], r[i]) / gamma[i];
                    FN(add)(N, -tau[i * ld + j], r[i], r[j]);
                }
            gamma[j] = FN(dot)(N, r[j], r[j]);
            if (gamma[j] != (T)0) gp[j] = FN(dot)(N, r[0], r[j]) / gamma[j];
        }
        gamma[l] = gp[l];
        omega = gamma[l];
        for (int j = l - 1; j >= 1; j--) {
            sigma = 0;
            for (int i = j + 1; i <= l; i++) sigma += tau[j * ld + i] * gamma[i];
            gamma[j] = gp[j] - sigma;
        }
        for (int j = 1; j <= l - 1; j++) {
            sigma = 0;
            for (int i = j + 1; i <= l - 1; i++) sigma += tau[j * ld + i] * gamma[i + 1];
            gt[j] = gamma[j + 1] + sigma;
        }
        FN(add)(N, gamma[1], r[0], x);
        FN(add)(N, -gp[l], r[l], r[0]);
        FN(add)(N, -gamma[l], u[l], u[0]);
        for (int j = 1; j <= l - 1; j++) {
            FN(add)(N, -gamma[j], u[j], u[0]);
            FN(add)(N, gt[j], r[j], x);
            FN(add)(N, -gp[j], r[j], r[0]);
        }
    }
    FN(copy)(N, x, q);
    FN(precond)(precond, N, diag, q, x);
done:
    for (int i = 0; i <= l; i++) { free(r[i]); free(u[i]); }
    free(r); free(u); free(q); free(r0); free(gamma); free(gp); free(gt); free(tau);
    return orc_error_code(it);
}

/* ---- Gmres(m): computation/solver/iterative/Gmres.cxx:45-182 ------- */
int FN(gmres)(const orc_mat *A, T *x, const T *b, int precond, const T *diag,
              orc_iter *outer)
{
    const int N = A->m;
    if (N <= 0) return 0;
    const int m = outer->parameter_restart;
    const int ld = m + 1;
    T **V = (T **)malloc(sizeof(T *) * (size_t)(m + 1));
    for (int i = 0; i < m + 1; i++) { V[i] = FN(dup)(N, b); FN(fill)(N, (T)0, V[i]); }
    /* H: (m+1) x (m+1), column-major, H(k,i) = H[i*ld + k] (Gmres.cxx:63) */
    double *H = (double *)calloc((size_t)ld * ld, sizeof(double));
    double *s = (double *)calloc((size_t)(m + 1), sizeof(double));
    double *rc = (double *)calloc((size_t)(m + 1), sizeof(double));
    double *rs = (double *)calloc((size_t)(m + 1), sizeof(double));
    T *w = FN(dup)(N, b), *r = FN(dup)(N, b), *u = FN(dup)(N, b);
    FN(fill)(N, (T)0, w); FN(fill)(N, (T)0, r); FN(fill)(N, (T)0, u);

    /* residual w = b - A x, then r = M^{-1} w  (Gmres.cxx:82-90) */
    FN(copy)(N, b, w);
    if (!outer->init_guess_null) FN(mltadd)((T)-1, A, x, (T)1, w);
    else FN(fill)(N, (T)0, x);
    FN(precond)(precond, N, diag, w, r);
    double beta = (double)FN(norm2)(N, r);
    if (FN(iter_init)(outer, N, r) != 0) goto done;

    outer->nb_iter = 0;
    while (!orc_finished(outer, beta, outer->nb_iter, outer->max_iter)) {
        /* V0 = r / beta, s = beta e1   (Gmres.cxx:105-108) */
        FN(copy)(N, r, V[0]);
        FN(scal)(N, (T)(1.0 / beta), V[0]);
        for (int k = 0; k < m + 1; k++) s[k] = 0.0;
        s[0] = beta;
        int i = 0, k;
        /* inner iteration copies outer (Iterative.cxx:113-124) with
         * nb = outer nb and max = outer nb + m (Gmres.cxx:114-116) */
        int inner_nb = outer->nb_iter, inner_max = outer->nb_iter + m;
        /* test-only record of the first cycle (the Arnoldi invariant) */
        const int record = outer->arnoldi_h != NULL && outer->arnoldi_v != NULL && outer->nb_iter == 0;
        if (record) FN(copy)(N, V[0], (T *)outer->arnoldi_v);
        do {
            FN(mlt)(A, V[i], u);
            FN(precond)(precond, N, diag, u, w);
            /* modified Gram-Schmidt (Gmres.cxx:127-132) */
            for (k = 0; k <= i; k++) {
                H[i * ld + k] = (double)FN(dot)(N, V[k], w);
                FN(add)(N, (T)(-H[i * ld + k]), V[k], w);
            }
            double hi_ip1 = (double)FN(norm2)(N, w);
            FN(copy)(N, w, V[i + 1]);
            if (hi_ip1 != 0.0) FN(scal)(N, (T)(1.0 / hi_ip1), V[i + 1]);
            if (record) {
                for (k = 0; k <= i; k++) outer->arnoldi_h[(size_t)i * (m + 1) + k] = H[i * ld + k];
                outer->arnoldi_h[(size_t)i * (m + 1) + i + 1] = hi_ip1;
                FN(copy)(N, V[i + 1], (T *)outer->arnoldi_v + (size_t)(i + 1) * N);
                outer->arnoldi_steps = i + 1;
            }
            for (k = 0; k < i; k++)
                orc_applyrot(&H[i * ld + k], &H[i * ld + k + 1], rc[k], rs[k]);
            if (hi_ip1 != 0.0) {
                orc_genrot(&H[i * ld + i], &hi_ip1, &rc[i], &rs[i]);
                orc_applyrot(&s[i], &s[i + 1], rc[i], rs[i]);
            }
            inner_nb++; outer->nb_iter++; i++;
        } while (!orc_finished(outer, fabs(s[i]), inner_nb, inner_max));

        /* H.Resize(i,i); s.Resize(i); Solve(H, s)  (Gmres.cxx:165-166) */
        orc_trsv_upper(i, H, ld, s);
        /* x += sum_k s(k) V(k)  (Gmres.cxx:169-170) */
        for (k = 0; k < i; k++)
            if (s[k] != 0.0) FN(add)(N, (T)s[k], V[k], x);
        /* true preconditioned residual (Gmres.cxx:173-178) */
        FN(copy)(N, b, w);
        FN(mltadd)((T)-1, A, x, (T)1, w);
        FN(precond)(precond, N, diag, w, r);
        beta = (double)FN(norm2)(N, r);
    }
done:
    for (int i = 0; i < m + 1; i++) free(V[i]);
    free(V); free(H); free(s); free(rc); free(rs); free(w); free(r); free(u);
    return orc_error_code(outer);
}

/* ---- SDDMM on the pattern: g_j = -(lam_i * x_col(j)) -------------- */
void FN(sddmm)(int m, const int *ptr, const int *ind, const T *lam, const T *x,
               T *g)
{
    for (int i = 0; i < m; i++)
        for (int j = ptr[i]; j < ptr[i + 1]; j++) g[j] = -(lam[i] * x[ind[j]]);
}

#undef FN
#undef CAT
#undef CAT_

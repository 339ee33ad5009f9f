/*
 * seldon_oracle.c -- TEST INFRASTRUCTURE ONLY (see seldon_oracle.h).
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#include "seldon_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Iteration(int, tol) constructor: Iterative.cxx:97-109. */
/* The dimensions ConvertMatrix_from_Coordinates gives the matrix
 * (matrix_sparse/Matrix_Conversions.cxx:1310-1316): m = GetNormInf(IndRow)
 * + 1 (the largest absolute shifted index, 0 without entries) but at least
 * A.GetM(), n likewise. */
void orc_coo_dims(int nnz, const int *row, const int *col, int index, int m_min, int n_min, int *m, int *n)
{
    int rm = 0, cm = 0;
    for (int i = 0; i < nnz; i++) {
        int r = abs(row[i] - index), c = abs(col[i] - index);
        if (r > rm) rm = r;
        if (c > cm) cm = c;
    }
    *m = rm + 1 > m_min ? rm + 1 : m_min;
    *n = cm + 1 > n_min ? cm + 1 : n_min;
}

void orc_iter_init(orc_iter *it, int max_iter, double tol)
{
    memset(it, 0, sizeof(*it));
    it->max_iter = max_iter;
    it->tolerance = tol;
    it->print_level = 1;
    it->init_guess_null = 1;
    it->parameter_restart = 10;
}

static void orc_hist_push(orc_iter *it, int nb, double reste)
{
    if (it->hist && it->hist_len < it->hist_cap) {
        it->hist[it->hist_len] = reste;
        if (it->hist_it) it->hist_it[it->hist_len] = nb;
    }
    it->hist_len++;
}

/* Iteration::Finished(vector) / Finished(scalar), Iterative.cxx:272-319:
 * reste = facteur_reste * |r|; stop when reste < tolerance or nb >= max.
 * `nb` and `max` are passed so that Gmres's inner copy can use its own. */
static int orc_finished(orc_iter *it, double norm, int nb, int max)
{
    double reste = it->facteur_reste * norm;
    orc_hist_push(it, nb, reste);
    return (reste < it->tolerance) || (nb >= max);
}

/* Iteration::Fail, Iterative.cxx:323-331. */
static void orc_fail(orc_iter *it, int code)
{
    it->fail_convergence = 1;
    it->error_code = code;
}

/* Iteration::ErrorCode, Iterative.cxx:344-354. */
static int orc_error_code(const orc_iter *it)
{
    if (it->nb_iter >= it->max_iter) return -2;
    if (it->fail_convergence) return it->error_code;
    return 0;
}

/* GenRot, "old BLAS version": Functions_Vector.cxx:612-649. */
void orc_genrot(double *a_in, double *b_in, double *c_, double *s_)
{
    double roe = fabs(*a_in) > fabs(*b_in) ? *a_in : *b_in;
    double scal = fabs(*a_in) + fabs(*b_in);
    double r, z;
    if (scal != 0.0) {
        double a_scl = *a_in / scal;
        double b_scl = *b_in / scal;
        r = scal * sqrt(a_scl * a_scl + b_scl * b_scl);
        if (roe < 0.0) r *= -1.0;
        *c_ = *a_in / r;
        *s_ = *b_in / r;
        z = 1.0;
        if (fabs(*a_in) > fabs(*b_in)) z = *s_;
        else if (fabs(*b_in) >= fabs(*a_in) && *c_ != 0.0) z = 1.0 / *c_;
    } else {
        *c_ = 1.0;
        *s_ = 0.0;
        r = 0.0;
        z = 0.0;
    }
    *a_in = r;
    *b_in = z;
}

/* ApplyRot: Functions_Vector.cxx:681-687. */
void orc_applyrot(double *x, double *y, double c_, double s_)
{
    double temp = c_ * *x + s_ * *y;
    *y = c_ * *y - s_ * *x;
    *x = temp;
}

/* Column-oriented back substitution, the reference BLAS dtrsv loop for
 * (ColMajor, Upper, NoTrans, NonUnit, incx = 1) that Blas_2.cxx:3236-3237
 * delegates to. */
void orc_trsv_upper(int n, const double *H, int ld, double *s)
{
    for (int j = n - 1; j >= 0; j--) {
        if (s[j] != 0.0) {
            s[j] = s[j] / H[j * ld + j];
            double temp = s[j];
            for (int i = j - 1; i >= 0; i--) s[i] -= temp * H[j * ld + i];
        }
    }
}

#define T double
#define SFX d
#include "seldon_oracle_body.h"
#undef T
#undef SFX

#define T float
#define SFX s
#include "seldon_oracle_body.h"
#undef T
#undef SFX

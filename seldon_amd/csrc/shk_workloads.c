/*
 * shk_workloads.c -- seeded synthetic inputs for the sparse Krylov path.
 *
 * Host-only C (no HIP): builds the CSR matrices and vectors that the
 * benchmarks, the parity tests and the oracle all consume, so every side
 * sees bit-identical inputs.  Every generator is counter-based: element or
 * row i draws from its own splitmix64 stream seeded by (seed, i), so the
 * output does not depend on the thread count and rows can be built in
 * parallel.
 *
 * CSR layout follows Seldon's Matrix<T,General,RowSparse>
 * (matrix_sparse/Matrix_Sparse.hxx:42-142): 0-based int ptr[m+1], int
 * ind[nnz] sorted within each row, values[nnz].
 *
 * Operators (all "recorded choices" where the workload text leaves room):
 *   laplace2d   5-point Dirichlet Laplacian, diag 4, neighbours -1.
 *   aniso2d     -eps*u_xx - u_yy: diag 2(1+eps), x-neighbours -eps,
 *               y-neighbours -1.
 *   convdiff3d  -nu*Lap(u) + v.grad(u), first-order upwind (v > 0 uses
 *               backward differences), rows scaled by h^2, h = 1/(n+1):
 *               a_ii = 6 nu + h (vx+vy+vz); a_west = -nu - h vx;
 *               a_east = -nu; likewise in y and z.
 *   diffusion3d -div(exp(theta) grad u), 7-point, face coefficient = the
 *               harmonic mean of the two cells, Dirichlet boundary faces use
 *               the cell value.
 *   powerlaw    unstructured-FEM-like: row length k ~ P(k) ~ k^-a on
 *               [kmin,kmax]; the diagonal plus k-1 distinct columns drawn
 *               uniformly within [i-W, i+W]; off-diagonals N(0,1); the
 *               diagonal = dd * sum_{j != i} |a_ij|.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SHK_WL_API __attribute__((visibility("default")))

static inline uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Stream for item i of a generator seeded with `seed`. */
static inline uint64_t stream_init(uint64_t seed, uint64_t i)
{
    uint64_t s = seed * 0xD1B54A32D192ED03ull ^ (i + 1) * 0x9E3779B97F4A7C15ull;
    (void)sm64(&s);
    return s;
}

static inline double u01_open_closed(uint64_t *s) /* (0,1] */
{
    return (double)((sm64(s) >> 11) + 1) * 0x1.0p-53;
}

static inline double u01(uint64_t *s) /* [0,1) */
{
    return (double)(sm64(s) >> 11) * 0x1.0p-53;
}

static inline double normal01(uint64_t *s)
{
    double u1 = u01_open_closed(s), u2 = u01(s);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* x_i ~ U[lo,hi), one draw from stream (seed, i). */
SHK_WL_API void shk_wl_uniform(int64_t n, uint64_t seed, double lo, double hi,
                               double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint64_t s = stream_init(seed, (uint64_t)i);
        out[i] = lo + (hi - lo) * u01(&s);
    }
}

/* x_i ~ N(mean, sd^2). */
SHK_WL_API void shk_wl_normal(int64_t n, uint64_t seed, double mean, double sd,
                              double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint64_t s = stream_init(seed, (uint64_t)i);
        out[i] = mean + sd * normal01(&s);
    }
}

/* ---------------------------------------------------------------- 2-D -- */

SHK_WL_API int64_t shk_wl_aniso2d_nnz(int nx, int ny)
{
    return 5LL * nx * ny - 2LL * nx - 2LL * ny;
}

/* Rows in natural order, x fastest: row = iy*nx + ix.  Columns sorted. */
SHK_WL_API void shk_wl_aniso2d(int nx, int ny, double eps, int *ptr, int *ind,
                               double *val)
{
    const double diag = 2.0 * (1.0 + eps);
    int64_t k = 0;
    ptr[0] = 0;
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
            int row = iy * nx + ix;
            if (iy > 0) { ind[k] = row - nx; val[k++] = -1.0; }
            if (ix > 0) { ind[k] = row - 1; val[k++] = -eps; }
            ind[k] = row; val[k++] = diag;
            if (ix < nx - 1) { ind[k] = row + 1; val[k++] = -eps; }
            if (iy < ny - 1) { ind[k] = row + nx; val[k++] = -1.0; }
            ptr[row + 1] = (int)k;
        }
}

/* 5-point Laplacian = the eps = 1 anisotropic operator (diag 4, -1). */
SHK_WL_API void shk_wl_laplace2d(int nx, int ny, int *ptr, int *ind, double *val)
{
    shk_wl_aniso2d(nx, ny, 1.0, ptr, ind, val);
}

/* ---------------------------------------------------------------- 3-D -- */

SHK_WL_API int64_t shk_wl_stencil3d_nnz(int n)
{
    int64_t n3 = (int64_t)n * n * n;
    return 7 * n3 - 6LL * n * n;
}

SHK_WL_API void shk_wl_convdiff3d(int n, double vx, double vy, double vz,
                                  double nu, int *ptr, int *ind, double *val)
{
    const double h = 1.0 / (n + 1);
    const double diag = 6.0 * nu + h * (vx + vy + vz);
    const double west = -nu - h * vx, east = -nu;
    const double south = -nu - h * vy, north = -nu;
    const double bottom = -nu - h * vz, top = -nu;
    const int64_t nn = (int64_t)n * n;
    /* row r = (iz*n + iy)*n + ix; every row's length is known up front, so
     * the rows are filled in parallel from a prefix over plane counts. */
    ptr[0] = 0;
    for (int64_t r = 0; r < nn * n; r++) {
        int ix = (int)(r % n), iy = (int)((r / n) % n), iz = (int)(r / nn);
        int len = 1 + (ix > 0) + (ix < n - 1) + (iy > 0) + (iy < n - 1) +
                  (iz > 0) + (iz < n - 1);
        ptr[r + 1] = ptr[r] + len;
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nn * n; r++) {
        int ix = (int)(r % n), iy = (int)((r / n) % n), iz = (int)(r / nn);
        int64_t k = ptr[r];
        if (iz > 0)     { ind[k] = (int)(r - nn); val[k++] = bottom; }
        if (iy > 0)     { ind[k] = (int)(r - n);  val[k++] = south; }
        if (ix > 0)     { ind[k] = (int)(r - 1);  val[k++] = west; }
        ind[k] = (int)r; val[k++] = diag;
        if (ix < n - 1) { ind[k] = (int)(r + 1);  val[k++] = east; }
        if (iy < n - 1) { ind[k] = (int)(r + n);  val[k++] = north; }
        if (iz < n - 1) { ind[k] = (int)(r + nn); val[k++] = top; }
    }
}

static inline double harm(double a, double b) { return 2.0 * a * b / (a + b); }

/* kappa[c] = exp(theta[c]) per cell; SPD 7-point matrix. */
SHK_WL_API void shk_wl_diffusion3d(int n, const double *kappa, int *ptr,
                                   int *ind, double *val)
{
    const int64_t nn = (int64_t)n * n;
    ptr[0] = 0;
    for (int64_t r = 0; r < nn * n; r++) {
        int ix = (int)(r % n), iy = (int)((r / n) % n), iz = (int)(r / nn);
        int len = 1 + (ix > 0) + (ix < n - 1) + (iy > 0) + (iy < n - 1) +
                  (iz > 0) + (iz < n - 1);
        ptr[r + 1] = ptr[r] + len;
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nn * n; r++) {
        int ix = (int)(r % n), iy = (int)((r / n) % n), iz = (int)(r / nn);
        const double kc = kappa[r];
        double diag = 0.0;
        int64_t k = ptr[r], kd;
        /* neighbours in increasing column order; the diagonal is summed
         * over the six faces in the fixed order -z,-y,-x,+x,+y,+z. */
        double f[6];
        f[0] = iz > 0 ? harm(kc, kappa[r - nn]) : kc;
        f[1] = iy > 0 ? harm(kc, kappa[r - n]) : kc;
        f[2] = ix > 0 ? harm(kc, kappa[r - 1]) : kc;
        f[3] = ix < n - 1 ? harm(kc, kappa[r + 1]) : kc;
        f[4] = iy < n - 1 ? harm(kc, kappa[r + n]) : kc;
        f[5] = iz < n - 1 ? harm(kc, kappa[r + nn]) : kc;
        for (int q = 0; q < 6; q++) diag += f[q];
        if (iz > 0)     { ind[k] = (int)(r - nn); val[k++] = -f[0]; }
        if (iy > 0)     { ind[k] = (int)(r - n);  val[k++] = -f[1]; }
        if (ix > 0)     { ind[k] = (int)(r - 1);  val[k++] = -f[2]; }
        kd = k;
        ind[k] = (int)r; val[k++] = diag;
        if (ix < n - 1) { ind[k] = (int)(r + 1);  val[k++] = -f[3]; }
        if (iy < n - 1) { ind[k] = (int)(r + n);  val[k++] = -f[4]; }
        if (iz < n - 1) { ind[k] = (int)(r + nn); val[k++] = -f[5]; }
        (void)kd;
    }
}

/* ------------------------------------------------------ power-law rows -- */

typedef struct {
    int kmin, kmax;
    double *cdf; /* cdf[k - kmin] */
} plaw_t;

static void plaw_build(plaw_t *p, int kmin, int kmax, double expo)
{
    p->kmin = kmin; p->kmax = kmax;
    int nk = kmax - kmin + 1;
    p->cdf = (double *)malloc(sizeof(double) * nk);
    double acc = 0.0;
    for (int q = 0; q < nk; q++) { acc += pow((double)(kmin + q), -expo); p->cdf[q] = acc; }
    for (int q = 0; q < nk; q++) p->cdf[q] /= acc;
    p->cdf[nk - 1] = 1.0;
}

static int plaw_draw(const plaw_t *p, uint64_t *s)
{
    double u = u01(s);
    int lo = 0, hi = p->kmax - p->kmin;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (p->cdf[mid] > u) hi = mid; else lo = mid + 1; }
    return p->kmin + lo;
}

static int row_len(const plaw_t *p, uint64_t seed, int64_t n, int64_t i, int W)
{
    uint64_t s = stream_init(seed, (uint64_t)i);
    int k = plaw_draw(p, &s);
    int64_t lo = i - W < 0 ? 0 : i - W, hi = i + W > n - 1 ? n - 1 : i + W;
    int64_t avail = hi - lo + 1;
    if (k > avail) k = (int)avail;
    return k;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Pass 1: row lengths into ptr (prefix sums); returns nnz. */
SHK_WL_API int64_t shk_wl_powerlaw_ptr(int64_t n, uint64_t seed, int kmin,
                                       int kmax, double expo, int W, int *ptr)
{
    plaw_t p; plaw_build(&p, kmin, kmax, expo);
    ptr[0] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) ptr[i + 1] = row_len(&p, seed, n, i, W);
    int64_t acc = 0;
    for (int64_t i = 0; i < n; i++) { acc += ptr[i + 1]; ptr[i + 1] = (int)acc; }
    free(p.cdf);
    return acc;
}

/* Pass 2: columns and values (ptr from pass 1). */
SHK_WL_API void shk_wl_powerlaw_fill(int64_t n, uint64_t seed, int kmin,
                                     int kmax, double expo, int W, double dd,
                                     const int *ptr, int *ind, double *val)
{
    plaw_t p; plaw_build(&p, kmin, kmax, expo);
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * 2 * (kmax + 1));
#pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < n; i++) {
            uint64_t s = stream_init(seed, (uint64_t)i);
            (void)plaw_draw(&p, &s); /* same first draw as row_len */
            int k = ptr[i + 1] - ptr[i];
            int64_t lo = i - W < 0 ? 0 : i - W, hi = i + W > n - 1 ? n - 1 : i + W;
            int64_t span = hi - lo + 1;
            /* draw k-1 distinct off-diagonal columns: draw, sort, dedupe,
             * redraw the shortfall until complete. */
            int have = 0;
            while (have < k - 1) {
                while (have < k - 1) {
                    int64_t c = lo + (int64_t)(u01(&s) * (double)span);
                    if (c > hi) c = hi;
                    if (c == i) continue;
                    buf[have++] = (int)c;
                }
                qsort(buf, have, sizeof(int), cmp_int);
                int u = 0;
                for (int q = 0; q < have; q++)
                    if (u == 0 || buf[q] != buf[u - 1]) buf[u++] = buf[q];
                have = u;
            }
            buf[have++] = (int)i;
            qsort(buf, have, sizeof(int), cmp_int);
            int64_t base = ptr[i];
            double asum = 0.0;
            int64_t kd = -1;
            for (int q = 0; q < have; q++) {
                ind[base + q] = buf[q];
                if (buf[q] == i) { kd = base + q; val[base + q] = 0.0; continue; }
                double v = normal01(&s);
                val[base + q] = v;
                asum += fabs(v);
            }
            val[kd] = dd * asum;
            if (asum == 0.0) val[kd] = 1.0; /* a lone diagonal */
        }
        free(buf);
    }
    free(p.cdf);
}

/* ------------------------------------------------------------- helpers -- */

/* Strict upper triangle plus diagonal of a sorted CSR (RowSymSparse input:
 * Seldon keeps only the upper part, matrix_sparse/Matrix_Conversions.cxx:
 * 1394-1467).  Pass ptr_u[m+1] to size it first with ind_u == NULL. */
SHK_WL_API int64_t shk_wl_upper(int m, const int *ptr, const int *ind,
                                const double *val, int *ptr_u, int *ind_u,
                                double *val_u)
{
    int64_t k = 0;
    ptr_u[0] = 0;
    for (int i = 0; i < m; i++) {
        for (int j = ptr[i]; j < ptr[i + 1]; j++)
            if (ind[j] >= i) {
                if (ind_u) { ind_u[k] = ind[j]; val_u[k] = val[j]; }
                k++;
            }
        ptr_u[i + 1] = (int)k;
    }
    return k;
}

/* y = A x for a CSR in double, plain sequential per-row sums (used to form
 * right-hand sides b = A x_true). */
SHK_WL_API void shk_wl_csr_apply(int64_t m, const int *ptr, const int *ind,
                                 const double *val, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) {
        double t = 0.0;
        for (int j = ptr[i]; j < ptr[i + 1]; j++) t += val[j] * x[ind[j]];
        y[i] = t;
    }
}

/*
 * pxo_ops.c -- oracle (test infrastructure): PDDO multi-indices, moment
 * matrix, full-pivot LU solve and per-bond g-weights.
 *
 * The solve restates MatrixXd::solveMatrixChecked (MatrixXd.cpp:26-48), i.e.
 * Eigen 3.5.0's FullPivLU as vendored by the reference:
 *   computeInPlace   external/eigen/Eigen/src/LU/FullPivLU.h:360-439
 *   _solve_impl      FullPivLU.h:594-628
 *   rank()           FullPivLU.h (threshold = n*eps*maxPivot)
 * and the blocked triangular solve it dispatches to for the row-major
 * right-hand side (Core/products/TriangularSolverMatrix.h), with the SSE2
 * register blocking of an x86-64 build without FMA (gebp mr = nr = 4, so
 * panels of 4 rows; GeneralBlockPanelKernel.h:440-460) -- see flu_solve.
 */
#include "pxo.h"
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PXO_MAXNOP 20
#define PXO_PANEL 4

/* PDOperators.cpp:51-75 */
void pxo_multi_indices(int dim, int order, int *qp, int *qq, int *qr)
{
    int k = 0;
    for (int o = 1; o <= order; o++) {
        if (dim == 3) {
            for (int p = o; p >= 0; --p)
                for (int q = o - p; q >= 0; --q) { qp[k] = p; qq[k] = q; qr[k] = o - p - q; k++; }
        } else {
            for (int p = o; p >= 0; --p) { qp[k] = p; qq[k] = o - p; qr[k] = 0; k++; }
        }
    }
}

int pxo_nop(int dim, int order)
{
    int k = 0;
    for (int o = 1; o <= order; o++) k += (dim == 3) ? (o + 1) * (o + 2) / 2 : (o + 1);
    return k;
}

static int factorial(int n) { int r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }

/* ---- Eigen FullPivLU restatement (column-major semantics) --------------- */
typedef struct {
    double lu[PXO_MAXNOP][PXO_MAXNOP];   /* lu[row][col] */
    int rowT[PXO_MAXNOP], colT[PXO_MAXNOP];
    int nonzero, n;
    double maxpivot;
} flu;

static void flu_compute(flu *f, const double a[PXO_MAXNOP][PXO_MAXNOP], int n)
{
    f->n = n;
    memcpy(f->lu, a, sizeof(f->lu));
    f->nonzero = n;
    f->maxpivot = 0.0;
    for (int k = 0; k < n; ++k) {
        /* biggest |a| in the corner, first in column-major order */
        double best = -1.0; int br = k, bc = k;
        for (int c = k; c < n; ++c)
            for (int r = k; r < n; ++r) {
                const double s = fabs(f->lu[r][c]);
                if (s > best) { best = s; br = r; bc = c; }
            }
        if (best == 0.0) {
            f->nonzero = k;
            for (int i = k; i < n; ++i) { f->rowT[i] = i; f->colT[i] = i; }
            break;
        }
        if (best > f->maxpivot) f->maxpivot = best;
        f->rowT[k] = br; f->colT[k] = bc;
        if (k != br)
            for (int c = 0; c < n; ++c) { double t = f->lu[k][c]; f->lu[k][c] = f->lu[br][c]; f->lu[br][c] = t; }
        if (k != bc)
            for (int r = 0; r < n; ++r) { double t = f->lu[r][k]; f->lu[r][k] = f->lu[r][bc]; f->lu[r][bc] = t; }
        if (k < n - 1) {
            const double piv = f->lu[k][k];
            for (int r = k + 1; r < n; ++r) f->lu[r][k] /= piv;
            for (int c = k + 1; c < n; ++c)
                for (int r = k + 1; r < n; ++r) f->lu[r][c] -= f->lu[r][k] * f->lu[k][c];
        }
    }
}

static int flu_rank(const flu *f)
{
    const double thr = fabs(f->maxpivot) * ((double)f->n * DBL_EPSILON);
    int r = 0;
    for (int i = 0; i < f->nonzero; ++i) r += (fabs(f->lu[i][i]) > thr);
    return r;
}

/* Solve A X = B (n x n rhs) exactly as FullPivLU::_solve_impl.  The rhs of
 * solveMatrixChecked is a row-major Map, so Eigen solves the transposed
 * problem with the triangle on the right (TriangularSolverMatrix.h:301-313
 * forwards to the OnTheRight driver :327-438 and trsmKernelR :83-170).  In
 * terms of the left problem that is: panels of PXO_PANEL rows; each panel
 * first subtracts one sequential dot product over all rows already solved
 * (GEBP, alpha = -1), then solves its own rows left-looking. */
static void flu_solve(const flu *f, const double b[PXO_MAXNOP][PXO_MAXNOP],
                      double x[PXO_MAXNOP][PXO_MAXNOP])
{
    const int n = f->n, rank = flu_rank(f);
    double c[PXO_MAXNOP][PXO_MAXNOP];
    if (rank == 0) { memset(x, 0, sizeof(double) * PXO_MAXNOP * PXO_MAXNOP); return; }
    /* step 1: c = P b (row transpositions in factorisation order) */
    memcpy(c, b, sizeof(c));
    for (int k = 0; k < n; ++k) {
        const int r = f->rowT[k];
        if (r != k) for (int j = 0; j < n; ++j) { double t = c[k][j]; c[k][j] = c[r][j]; c[r][j] = t; }
    }
    /* step 2: unit-lower solve, panels j2 = 0, 4, 8, ... */
    for (int j2 = 0; j2 < n; j2 += PXO_PANEL) {
        const int w = (n - j2 < PXO_PANEL) ? n - j2 : PXO_PANEL;
        if (j2 > 0)
            for (int r = j2; r < j2 + w; ++r)
                for (int j = 0; j < n; ++j) {
                    double acc = 0.0;
                    for (int k = 0; k < j2; ++k) acc += c[k][j] * f->lu[r][k];
                    c[r][j] = c[r][j] + (-1.0 * acc);
                }
        for (int r = j2; r < j2 + w; ++r)
            for (int k = j2; k < r; ++k)
                for (int j = 0; j < n; ++j) c[r][j] -= c[k][j] * f->lu[r][k];
    }
    /* step 3: upper solve on the rank x rank corner; panels aligned to the
     * start, visited from the last (partial) one down to 0 */
    const int sz = rank;
    int j2 = sz - ((sz % PXO_PANEL) ? (sz % PXO_PANEL) : PXO_PANEL);
    for (; j2 >= 0; j2 -= PXO_PANEL) {
        const int w = (sz - j2 < PXO_PANEL) ? sz - j2 : PXO_PANEL;
        const int po = j2 + w;                      /* solved rows po .. sz-1 */
        if (po < sz)
            for (int r = j2; r < po; ++r)
                for (int j = 0; j < n; ++j) {
                    double acc = 0.0;
                    for (int k = po; k < sz; ++k) acc += c[k][j] * f->lu[r][k];
                    c[r][j] = c[r][j] + (-1.0 * acc);
                }
        for (int r = po - 1; r >= j2; --r) {
            for (int k = r + 1; k < po; ++k)
                for (int j = 0; j < n; ++j) c[r][j] -= c[k][j] * f->lu[r][k];
            const double inv = 1.0 / f->lu[r][r];
            for (int j = 0; j < n; ++j) c[r][j] *= inv;
        }
    }
    /* step 4: undo the column permutation q (transpositions in order) */
    int q[PXO_MAXNOP];
    for (int i = 0; i < n; ++i) q[i] = i;
    for (int k = 0; k < n; ++k) { const int t = q[k]; q[k] = q[f->colT[k]]; q[f->colT[k]] = t; }
    for (int i = 0; i < rank; ++i) for (int j = 0; j < n; ++j) x[q[i]][j] = c[i][j];
    for (int i = rank; i < n; ++i) for (int j = 0; j < n; ++j) x[q[i]][j] = 0.0;
}

/* ---- CalcPDOperators.cpp:48-60 ----------------------------------------- */
static double weight(double xsi_norm, double delta)
{
    const double r = 2.0 * xsi_norm / delta;
    return exp(-r * r);
}

static double vol_correction(double xsi_norm, double delta, double dx_char)
{
    if (dx_char <= 0.0) return 1.0;
    if (xsi_norm <= delta - 0.5 * dx_char) return 1.0;
    const double vc = (delta - xsi_norm + 0.5 * dx_char) / dx_char;
    return vc < 0.0 ? 0.0 : (vc > 1.0 ? 1.0 : vc);
}

/* PDOperators::computeXsiMonomials (CalcPDOperators.cpp:76-87) */
static void monomials(int nop, const int *qp, const int *qq, const int *qr,
                      double ex, double ey, double ez, double *m)
{
    for (int k = 0; k < nop; ++k) {
        double mono = 1.0;
        for (int t = 0; t < qp[k]; ++t) mono *= ex;
        for (int t = 0; t < qq[k]; ++t) mono *= ey;
        for (int t = 0; t < qr[k]; ++t) mono *= ez;
        m[k] = mono;
    }
}

/* calcAMatrixCompute (CalcPDOperators.cpp:108-226) followed by
 * calcOperatorsCompute (:307-354) for every bond of node i (0-based). */
static unsigned char node_operators(int i, int dim, int nop, const int *qp,
                                    const int *qq, const int *qr,
                                    const double *coords, const double *vols,
                                    const int *off, const int *nbr, double delta,
                                    double dx_char, int vol_corr, double *ops)
{
    (void)dim;
    double A[PXO_MAXNOP][PXO_MAXNOP], B[PXO_MAXNOP][PXO_MAXNOP], X[PXO_MAXNOP][PXO_MAXNOP];
    double mono[PXO_MAXNOP], idp[PXO_MAXNOP];
    memset(A, 0, sizeof(A)); memset(B, 0, sizeof(B));
    const double inv_delta = 1.0 / delta;
    for (int k = 0; k < nop; ++k) {
        const int tot = qp[k] + qq[k] + qr[k];
        double v = 1.0;
        for (int t = 0; t < tot; ++t) v *= inv_delta;
        idp[k] = v;
        B[k][k] = (double)(factorial(qp[k]) * factorial(qq[k]) * factorial(qr[k]));
    }
    const double x1 = coords[3 * i], y1 = coords[3 * i + 1], z1 = coords[3 * i + 2];
    int usable = 0;
    for (int t = off[i]; t < off[i + 1]; ++t) {
        const int j = nbr[t] - 1;
        const double xx = coords[3 * j] - x1, xy = coords[3 * j + 1] - y1, xz = coords[3 * j + 2] - z1;
        const double xn = sqrt(xx * xx + xy * xy + xz * xz);
        if (xn <= 0.0) continue;
        const double w = weight(xn, delta);
        const double vc = vol_corr ? vol_correction(xn, delta, dx_char) : 1.0;
        const double bw = w * vc * vols[j];
        if (bw <= 0.0) continue;
        ++usable;
        monomials(nop, qp, qq, qr, xx * inv_delta, xy * inv_delta, xz * inv_delta, mono);
        for (int a = 0; a < nop; ++a) {
            const double xi = mono[a];
            for (int b = 0; b < nop; ++b) A[a][b] += bw * xi * mono[b];
        }
    }
    flu f;
    flu_compute(&f, (const double (*)[PXO_MAXNOP])A, nop);
    flu_solve(&f, (const double (*)[PXO_MAXNOP])B, X);
    /* consistency check (MatrixXd.cpp:37-47) */
    double resid = 0.0, bmax = 0.0, amax = 0.0, xmax = 0.0;
    for (int r = 0; r < nop; ++r)
        for (int c = 0; c < nop; ++c) {
            double s = 0.0;
            for (int k = 0; k < nop; ++k) s += A[r][k] * X[k][c];
            resid = fmax(resid, fabs(s - B[r][c]));
            bmax = fmax(bmax, fabs(B[r][c]));
            amax = fmax(amax, fabs(A[r][c]));
            xmax = fmax(xmax, fabs(X[r][c]));
        }
    const double scale = fmax(bmax, amax * xmax);
    const int consistent = resid <= 1.0e-8 * fmax(scale, DBL_MIN);
    unsigned char flag = 0;
    if (usable < nop) flag |= 2;
    if (usable < nop || !consistent) flag |= 1;
    /* per-bond operators */
    for (int t = off[i]; t < off[i + 1]; ++t) {
        double *op = ops + (size_t)t * nop;
        const int j = nbr[t] - 1;
        const double xx = coords[3 * j] - x1, xy = coords[3 * j + 1] - y1, xz = coords[3 * j + 2] - z1;
        const double xn = sqrt(xx * xx + xy * xy + xz * xz);
        for (int k = 0; k < nop; ++k) op[k] = 0.0;
        if (xn <= 0.0) continue;
        const double w = weight(xn, delta);
        const double vc = vol_corr ? vol_correction(xn, delta, dx_char) : 1.0;
        const double bw = w * vc * vols[j];
        if (bw <= 0.0) continue;
        monomials(nop, qp, qq, qr, xx * inv_delta, xy * inv_delta, xz * inv_delta, mono);
        for (int k = 0; k < nop; ++k) {
            double s = 0.0;
            for (int a = 0; a < nop; ++a) s += X[a][k] * mono[a];
            op[k] = s * bw * idp[k];
        }
    }
    return flag;
}

int pxo_operators(int n, int dim, int order, const double *coords,
                  const double *vols, const int *off, const int *nbr,
                  double delta, double dx_char, int vol_corr,
                  double *ops, unsigned char *flags, int nthreads)
{
    const int nop = pxo_nop(dim, order);
    if (nop > PXO_MAXNOP) return -1;
    int qp[PXO_MAXNOP], qq[PXO_MAXNOP], qr[PXO_MAXNOP];
    pxo_multi_indices(dim, order, qp, qq, qr);
    int ndeg = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(guided) num_threads(nthreads) reduction(+:ndeg)
#endif
    for (int i = 0; i < n; ++i) {
        const unsigned char fl = node_operators(i, dim, nop, qp, qq, qr, coords, vols,
                                                off, nbr, delta, dx_char, vol_corr, ops);
        if (flags) flags[i] = fl;
        ndeg += (fl & 1);
    }
    (void)nthreads;
    return ndeg;
}

/*
 * pxo_solve.cpp -- CPU oracle of the reference's default linear solver
 * (TEST INFRASTRUCTURE ONLY; see pxo.h).
 *
 * The reference's DefaultSolver (src/LinearSolver/DefaultSolver.cpp) solves
 * the implicit Newton system K dU = RHS directly:
 *   1. a symmetric reverse Cuthill-McKee ordering of the CSR pattern
 *      (computeOrdering, :52-150);
 *   2. one envelope per equation j, the first index k < j with a pattern
 *      entry in row j or column j (buildProfile, :164-194);
 *   3. A = L D U with unit L and U and no pivoting, column by column in the
 *      envelope (factorize, :287-370), failing on an exactly zero pivot;
 *   4. forward, diagonal and backward substitution (solveFactored, :372-424).
 *
 * This file restates those steps.  C++ (not C) because step 1 orders each
 * vertex's new neighbours with std::sort on their degree, and only the same
 * library sort reproduces the reference's order among equal degrees.  The
 * sums run in the reference's order, so that with -ffp-contract=off the
 * factors and the solution match the reference build bit for bit.
 */
#include <algorithm>
#include <climits>
#include <queue>
#include <vector>
#include "pxo.h"

namespace {

/* DefaultSolver.cpp:52-150.  new -> old in perm, old -> new in inv. */
void order(int n, const int *rp, const int *cols, int use_rcm, int *perm, int *inv)
{
    for (int i = 0; i < n; i++) perm[i] = i;
    if (use_rcm && n > 0) {
        std::vector<int> deg(n);
        for (int v = 0; v < n; v++) {
            int d = 0;
            for (int p = rp[v]; p < rp[v + 1]; p++) d += cols[p] != v;
            deg[v] = d;
        }
        std::vector<char> seen(n, 0);
        std::vector<int> cm;
        cm.reserve(n);
        std::vector<int> fresh;
        int from = 0;
        for (;;) {
            // next component: the first unvisited vertex of least degree at
            // or after the previous start (:91-102)
            int s = -1, sd = INT_MAX;
            for (int v = from; v < n && sd > 0; v++)
                if (!seen[v] && deg[v] < sd) { s = v; sd = deg[v]; }
            if (s < 0) break;
            from = s + 1;
            std::queue<int> q;
            seen[s] = 1;
            q.push(s);
            while (!q.empty()) {
                const int v = q.front();
                q.pop();
                cm.push_back(v);
                fresh.clear();
                for (int p = rp[v]; p < rp[v + 1]; p++) {
                    const int w = cols[p];
                    if (w != v && !seen[w]) { seen[w] = 1; fresh.push_back(w); }
                }
                std::sort(fresh.begin(), fresh.end(), [&deg](int a, int b) { return deg[a] < deg[b]; });
                for (int w : fresh) q.push(w);
            }
        }
        if ((int)cm.size() == n)
            for (int k = 0; k < n; k++) perm[k] = cm[n - 1 - k];
    }
    for (int k = 0; k < n; k++) inv[perm[k]] = k;
}

}  // namespace

extern "C" long long pxo_ldu_profile(int n, const int *rowptr, const int *cols, int use_rcm, int *perm,
                                     long long *jp)
{
    std::vector<int> inv(n);
    order(n, rowptr, cols, use_rcm, perm, inv.data());
    // :177-194 -- first[j]: least k < j with (j, k) or (k, j) in the pattern
    std::vector<int> first(n);
    for (int j = 0; j < n; j++) first[j] = j;
    for (int io = 0; io < n; io++) {
        const int i = inv[io];
        for (int p = rowptr[io]; p < rowptr[io + 1]; p++) {
            const int j = inv[cols[p]];
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            if (lo < first[hi]) first[hi] = lo;
        }
    }
    jp[0] = 0;
    for (int j = 0; j < n; j++) jp[j + 1] = jp[j] + (j - first[j]);
    return jp[n];
}

/* The whole solve: returns 0, or 1 + the 0-based equation (in the new
 * order) of the first zero pivot. */
extern "C" int pxo_ldu_solve(int n, const int *rowptr, const int *cols, const double *vals, const double *b,
                             int use_rcm, double *x, int *perm, long long *jp)
{
    pxo_ldu_profile(n, rowptr, cols, use_rcm, perm, jp);
    std::vector<int> inv(n);
    for (int k = 0; k < n; k++) inv[perm[k]] = k;
    const long long prof = jp[n];
    // column j of the upper part and row j of the lower part share the
    // envelope [j - K_j, j): up[jp[j] + r] = A(j-K_j+r, j), lo[jp[j] + r] = A(j, j-K_j+r)
    std::vector<double> up((size_t)prof, 0.0), lo((size_t)prof, 0.0), dg(n, 0.0);
    for (int io = 0; io < n; io++) {                       // copyValuesFromCSR, :248-285
        const int i = inv[io];
        for (int p = rowptr[io]; p < rowptr[io + 1]; p++) {
            const int j = inv[cols[p]];
            if (i == j) dg[i] = vals[p];
            else if (i < j) up[(size_t)(jp[j] + (i - (j - (jp[j + 1] - jp[j]))))] = vals[p];
            else lo[(size_t)(jp[i] + (j - (i - (jp[i + 1] - jp[i]))))] = vals[p];
        }
    }
    // factorize, :287-370.  While column j is reduced its entries still carry
    // the factor D(k) of their row (V = D U, V' = L D); earlier columns are final.
    for (int j = 0; j < n; j++) {
        const int kj = (int)(jp[j + 1] - jp[j]);
        const int s = j - kj;
        double *uj = up.data() + jp[j], *lj = lo.data() + jp[j];
        for (int r = 0; r < kj; r++) {
            const int i = s + r;
            const int ki = (int)(jp[i + 1] - jp[i]);
            const int m = r < ki ? r : ki;                  // overlap k in [i-m, i)
            if (m <= 0) continue;
            const double *ui = up.data() + jp[i + 1] - m, *li = lo.data() + jp[i + 1] - m;
            const double *ujk = uj + r - m, *ljk = lj + r - m;
            double su = 0.0, sl = 0.0;
            for (int t = 0; t < m; t++) {
                su += ujk[t] * li[t];
                sl += ljk[t] * ui[t];
            }
            uj[r] -= su;
            lj[r] -= sl;
        }
        double dsub = 0.0;
        for (int r = 0; r < kj; r++) {
            const double dinv = dg[s + r];
            const double u = uj[r] * dinv;
            dsub += lj[r] * u;
            uj[r] = u;
            lj[r] = lj[r] * dinv;
        }
        dg[j] -= dsub;
        if (dg[j] == 0.0) return 1 + j;
        dg[j] = 1.0 / dg[j];
    }
    // solveFactored, :372-424
    std::vector<double> y(n);
    for (int k = 0; k < n; k++) y[k] = b[perm[k]];
    for (int j = 0; j < n; j++) {
        const int kj = (int)(jp[j + 1] - jp[j]);
        if (kj <= 0) continue;
        const double *lj = lo.data() + jp[j], *yk = y.data() + (j - kj);
        double s = 0.0;
        for (int r = 0; r < kj; r++) s += lj[r] * yk[r];
        y[j] -= s;
    }
    for (int j = 0; j < n; j++) y[j] *= dg[j];
    for (int j = n - 1; j > 0; j--) {
        const int kj = (int)(jp[j + 1] - jp[j]);
        if (kj <= 0) continue;
        const double xj = y[j];
        const double *uj = up.data() + jp[j];
        double *yk = y.data() + (j - kj);
        for (int r = 0; r < kj; r++) yk[r] -= uj[r] * xj;
    }
    for (int io = 0; io < n; io++) x[io] = y[inv[io]];
    return 0;
}

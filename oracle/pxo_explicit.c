/*
 * pxo_explicit.c -- oracle (test infrastructure): explicit strong-form PDDO
 * fracture step, BC presets, central-difference driver, RandomIC and the
 * FP64 adjoint of the explicit step.
 */
#include "pxo.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ExplicitPDDOFracElement::recalcElasticConstants (ExplicitPDDOFracElement.cpp:37-52) */
void pxo_elastic_constants(int dim, int plane_stress, double E, double nu,
                           double *C11, double *C12, double *C66)
{
    if (E <= 0.0) { *C11 = *C12 = *C66 = 0.0; return; }
    if (dim >= 3 || !plane_stress) {
        const double d = (1.0 + nu) * (1.0 - 2.0 * nu);
        *C11 = E * (1.0 - nu) / d; *C12 = E * nu / d; *C66 = E / (2.0 * (1.0 + nu));
        return;
    }
    const double d = 1.0 - nu * nu;
    *C11 = E / d; *C12 = E * nu / d; *C66 = E / (2.0 * (1.0 + nu));
}

/* FractureCriterion.h:30-43 */
double pxo_critical_stretch(double G0, double E, double delta, int dim, int plane_stress)
{
    double s2;
    if (G0 <= 0.0 || E <= 0.0 || delta <= 0.0) s2 = 0.0;
    else if (dim >= 3) s2 = 5.0 * G0 / (6.0 * E * delta);
    else if (plane_stress) s2 = 4.0 * M_PI * G0 / (9.0 * E * delta);
    else s2 = 5.0 * M_PI * G0 / (12.0 * E * delta);
    return sqrt(s2);
}

/* ExplicitPDDOFracElement::estimateStableDt (:54-77), uniform spacing */
double pxo_stable_dt(int dim, int plane_stress, double E, double nu, double rho,
                     const double *dxyz)
{
    if (E <= 0.0 || rho <= 0.0) return -1.0;
    double dx = -1.0;
    for (int k = 0; k < 3; k++) if (dxyz[k] > 0.0 && (dx <= 0.0 || dxyz[k] < dx)) dx = dxyz[k];
    if (dx <= 0.0) return -1.0;
    double c11;
    if (dim >= 3 || !plane_stress) {
        const double d = (1.0 + nu) * (1.0 - 2.0 * nu);
        if (d <= 0.0) return -1.0;
        c11 = E * (1.0 - nu) / d;
    } else {
        const double d = 1.0 - nu * nu;
        if (d <= 0.0) return -1.0;
        c11 = E / d;
    }
    return dx / sqrt(c11 / rho);
}

/* op cache layout of ExplicitPDDOFracElement.cpp:157-165:
 * 2D [Gx,Gy,Gxx,Gyy,Gxy], 3D [Gx,Gy,Gz,Gxx,Gyy,Gzz,Gxy,Gxz,Gyz]
 * taken from the full slot order of PDOperators.cpp:51-75. */
void pxo_frac_opcache(int n, int dim, const int *off, const double *ops, double *opc)
{
    static const int s2[5] = {0, 1, 2, 4, 3};
    static const int s3[9] = {0, 1, 2, 3, 6, 8, 4, 5, 7};
    const int nop = (dim >= 3) ? 9 : 5;
    const int *s = (dim >= 3) ? s3 : s2;
    const long long nb = off[n];
    for (long long t = 0; t < nb; t++)
        for (int k = 0; k < nop; k++) opc[t * nop + k] = ops[t * nop + s[k]];
}

static void node_consts(const pxo_frac *p, const double *node_E, int i,
                        double *C11, double *C12, double *C66)
{
    if (node_E) pxo_elastic_constants(p->dim, p->plane_stress, node_E[i], p->nu, C11, C12, C66);
    else { *C11 = p->C11; *C12 = p->C12; *C66 = p->C66; }
}

/* ExplicitPDDOFracElement::preprocessIteration per-step loop (:237-299) */
void pxo_frac_accel(int n, const pxo_frac *p, const double *coords,
                    const int *off, const int *nbr, const double *opc,
                    double *health, const double *u, const double *node_E,
                    double *accel, int nthreads)
{
    const int dim = p->dim;
    const int nops = (dim >= 3) ? 9 : 5;
    const int damage = p->damage_on && p->sc > 0.0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(guided) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++) {
        const int rowI = i * dim;
        const double xi0 = coords[3 * i], yi0 = coords[3 * i + 1];
        const double zi0 = (dim >= 3) ? coords[3 * i + 2] : 0.0;
        if (damage) {
            for (int t = off[i]; t < off[i + 1]; t++) {
                if (health[t] <= 0.0) continue;
                const int j = nbr[t] - 1, rowJ = j * dim;
                double xi[3], eta[3];
                xi[0] = coords[3 * j] - xi0; xi[1] = coords[3 * j + 1] - yi0;
                xi[2] = (dim >= 3) ? coords[3 * j + 2] - zi0 : 0.0;
                eta[0] = u[rowJ] - u[rowI]; eta[1] = u[rowJ + 1] - u[rowI + 1];
                eta[2] = (dim >= 3) ? u[rowJ + 2] - u[rowI + 2] : 0.0;
                const double r = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
                if (r <= 0.0) continue;
                const double yx = xi[0] + eta[0], yy = xi[1] + eta[1], yz = xi[2] + eta[2];
                const double s = (sqrt(yx * yx + yy * yy + yz * yz) - r) / r;
                if (p->sc > 0.0 && (p->tension_only ? (s > p->sc) : (fabs(s) > p->sc)))
                    health[t] = 0.0;
            }
        }
        double C11, C12, C66;
        node_consts(p, node_E, i, &C11, &C12, &C66);
        double L[3] = {0.0, 0.0, 0.0};
        for (int t = off[i]; t < off[i + 1]; t++) {
            const double mu = health[t];
            if (mu <= 0.0) continue;
            const int j = nbr[t] - 1, rowJ = j * dim;
            const double *op = opc + (size_t)t * nops;
            const double dux = u[rowJ] - u[rowI], duy = u[rowJ + 1] - u[rowI + 1];
            if (dim >= 3) {
                const double duz = u[rowJ + 2] - u[rowI + 2];
                const double Gxx = op[3], Gyy = op[4], Gzz = op[5], Gxy = op[6], Gxz = op[7], Gyz = op[8];
                const double p11 = C11 * Gxx + C66 * (Gyy + Gzz);
                const double p22 = C11 * Gyy + C66 * (Gxx + Gzz);
                const double p33 = C11 * Gzz + C66 * (Gxx + Gyy);
                const double a = C12 + C66;
                const double p12 = a * Gxy, p13 = a * Gxz, p23 = a * Gyz;
                L[0] += mu * (p11 * dux + p12 * duy + p13 * duz);
                L[1] += mu * (p12 * dux + p22 * duy + p23 * duz);
                L[2] += mu * (p13 * dux + p23 * duy + p33 * duz);
            } else {
                const double Gxx = op[2], Gyy = op[3], Gxy = op[4];
                const double p11 = C11 * Gxx + C66 * Gyy;
                const double p22 = C66 * Gxx + C11 * Gyy;
                const double pxy = (C12 + C66) * Gxy;
                L[0] += mu * (p11 * dux + pxy * duy);
                L[1] += mu * (pxy * dux + p22 * duy);
            }
        }
        const double b[3] = {p->bx, p->by, p->bz};
        for (int a = 0; a < dim; a++) accel[(size_t)i * dim + a] = (L[a] + b[a]) / p->rho;
    }
    (void)nthreads;
}

/* BCSystem::presetSolution (BCSystem.cpp:114-129) -> DirichletBC::
 * presetSolutionAtTime (DirichletBC.cpp:107-147) / NeumannBC::presetSolution
 * (NeumannBC.cpp:88-105).  BCs apply in list order; the last writer wins. */
static int inside_box(const pxo_bc *bc, const double *coords, int id1)
{
    for (int ax = 0; ax < 3; ++ax) {
        const double lo = bc->box[2 * ax], hi = bc->box[2 * ax + 1];
        if (lo > hi) continue;
        const double x = coords[3 * (id1 - 1) + ax];
        if (x < lo || x > hi) return 0;
    }
    return 1;
}

void pxo_bc_preset(int n, int ndof, const double *coords, const int *mirror,
                   int nbc, const pxo_bc *bcs, double *u, double t)
{
    (void)n;
    for (int e = 0; e < nbc; e++) {
        const pxo_bc *bc = bcs + e;
        for (int q = 0; q < bc->nnodes; q++) {
            const int id = bc->nodes[q];
            const int g = (id - 1) * ndof;
            const int b1 = mirror ? mirror[id - 1] : 0;
            if (bc->type == 0) {
                if (b1 < 1) continue;           /* reference aborts; tests avoid it */
                for (int k = 0; k < bc->ncomp; k++)
                    u[g + bc->comp[k]] = u[(b1 - 1) * ndof + bc->comp[k]];
            } else {
                if (bc->has_box && !inside_box(bc, coords, id)) continue;
                for (int k = 0; k < bc->ncomp; k++) {
                    const double value = bc->value[k] + bc->velocity[k] * t;
                    u[g + bc->comp[k]] = (b1 >= 1 && !bc->direct)
                                             ? 2.0 * value - u[(b1 - 1) * ndof + bc->comp[k]]
                                             : value;
                }
            }
        }
    }
}

/* ExplicitPDDOFracElement::computeNodalProjection (ExplicitPDDOFracElement.cpp:
 * 324-405): damage = 1 - (sum of bond health)/|H_i|; otherwise the
 * displacement gradient gu[a][c] = sum_j (u_j - u_i)_a G_c(i,j) over the whole
 * family in family order (health is not consulted), eps = sym(gu),
 * sigma = C12 tr(eps) I + 2 C66 eps with sigma_zz = C12 tr(eps) under 2D plane
 * strain and 0 under plane stress, von Mises from sigma. */
void pxo_frac_project(int n, int dim, int plane_stress, double C12, double C66,
                      const int *off, const int *nbr, const double *opc,
                      const double *health, const double *u, int which, double *out)
{
    const int nops = (dim == 3) ? 9 : 5;
    const int nc = (which >= 2) ? 9 : 1;
    for (int i = 0; i < n; i++) {
        double *o = out + (size_t)i * nc;
        const int b = off[i], nb = off[i + 1] - off[i];
        if (which == 0) {
            if (nb <= 0) { o[0] = 0.0; continue; }
            double live = 0.0;
            for (int t = 0; t < nb; t++) live += health[b + t];
            o[0] = 1.0 - live / (double)nb;
            continue;
        }
        const int rowI = i * dim;
        double gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < nb; t++) {
            const int rowJ = (nbr[b + t] - 1) * dim;
            const double *op = opc + (size_t)(b + t) * nops;
            const double du0 = u[rowJ] - u[rowI], du1 = u[rowJ + 1] - u[rowI + 1];
            if (dim >= 3) {
                const double du2 = u[rowJ + 2] - u[rowI + 2];
                const double Gx = op[0], Gy = op[1], Gz = op[2];
                gu[0] += du0 * Gx; gu[1] += du0 * Gy; gu[2] += du0 * Gz;
                gu[3] += du1 * Gx; gu[4] += du1 * Gy; gu[5] += du1 * Gz;
                gu[6] += du2 * Gx; gu[7] += du2 * Gy; gu[8] += du2 * Gz;
            } else {
                const double Gx = op[0], Gy = op[1];
                gu[0] += du0 * Gx; gu[1] += du0 * Gy; gu[3] += du1 * Gx; gu[4] += du1 * Gy;
            }
        }
        double eps[6] = {0, 0, 0, 0, 0, 0};
        if (dim >= 3) {
            eps[0] = gu[0]; eps[1] = gu[4]; eps[2] = gu[8];
            eps[3] = 0.5 * (gu[1] + gu[3]); eps[4] = 0.5 * (gu[2] + gu[6]); eps[5] = 0.5 * (gu[5] + gu[7]);
        } else {
            eps[0] = gu[0]; eps[1] = gu[4]; eps[3] = 0.5 * (gu[1] + gu[3]);
        }
        const double lam = C12, mu = C66;
        const double tr = eps[0] + eps[1] + eps[2];
        double sig[6];
        sig[0] = lam * tr + 2.0 * mu * eps[0]; sig[1] = lam * tr + 2.0 * mu * eps[1];
        sig[2] = (dim >= 3) ? (lam * tr + 2.0 * mu * eps[2]) : (!plane_stress ? lam * tr : 0.0);
        sig[3] = 2.0 * mu * eps[3]; sig[4] = 2.0 * mu * eps[4]; sig[5] = 2.0 * mu * eps[5];
        if (which == 1) {
            const double sm = (sig[0] + sig[1] + sig[2]) / 3.0;
            const double a = sig[0] - sm, c = sig[1] - sm, d = sig[2] - sm;
            o[0] = sqrt(1.5 * (a * a + c * c + d * d + 2.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5])));
            continue;
        }
        const double *v = (which == 2) ? sig : eps;
        o[0] = v[0]; o[1] = v[3]; o[2] = v[4];
        o[3] = v[3]; o[4] = v[1]; o[5] = v[5];
        o[6] = v[4]; o[7] = v[5]; o[8] = v[2];
    }
}

/* CSR slot of (row, col) in a sorted row, -1 if not stored
 * (SparseMatrix::findEntry, SparseMatrix.cpp:187-214) */
static int csr_find(const int *rowptr, const int *cols, int row, int col)
{
    int lo = rowptr[row], hi = rowptr[row + 1] - 1;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (cols[mid] == col) return mid;
        if (cols[mid] < col) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* bcRowScale (BCBase.h:33-39): the stored diagonal, else the row's largest
 * |entry| (SparseMatrix::getRowMaxAbs, SparseMatrix.cpp:317-331), else 1 */
static double bc_row_scale(const int *rowptr, const double *K, int row, int dslot)
{
    double scale = K[dslot];
    if (fabs(scale) < 1.0e-100) {
        double m = 0.0;
        for (int k = rowptr[row]; k < rowptr[row + 1]; k++) {
            const double a = fabs(K[k]);
            if (a > m) m = a;
        }
        scale = m;
    }
    if (fabs(scale) < 1.0e-100) scale = 1.0;
    return scale;
}

/* BCSystem::applyBCs (BCSystem.cpp:212-233) in list order:
 *   Dirichlet (DirichletBC::applyAtTime, DirichletBC.cpp:176-230): box
 *     filter; row g zeroed with diagonal s; with a mirror bulk node b and not
 *     direct, K(g,b) = s and R(g) = s (2 g(t) - U(g) - U(b)); otherwise
 *     R(g) = s (g(t) - U(g)); g(t) = value + velocity t.
 *   Neumann (NeumannBC::apply, NeumannBC.cpp:127-151): no box; row g zeroed
 *     with diagonal s, K(g,b) = -s, R(g) = s (U(b) - U(g)).
 * s = bcRowScale of the row as it stands when the entry reaches it
 * (SparseMatrix::zeroRowAndSetDiagonal SparseMatrix.cpp:333-350, insertValue
 * SparseMatrix.cpp:216-233). */
int pxo_bc_rows(int n, int ndof, const double *coords, const int *mirror,
                int nbc, const pxo_bc *bcs, double t, const double *U,
                const int *rowptr, const int *cols, double *K, double *R)
{
    for (int e = 0; e < nbc; e++) {
        const pxo_bc *bc = bcs + e;
        for (int q = 0; q < bc->nnodes; q++) {
            const int id = bc->nodes[q];
            if (id < 1 || id > n) return -4;
            const int b1 = mirror ? mirror[id - 1] : 0;
            if (bc->type == 0 && (b1 < 1 || b1 > n)) return -3;
            if (bc->type == 1 && bc->has_box && !inside_box(bc, coords, id)) continue;
            for (int k = 0; k < bc->ncomp; k++) {
                const int c = bc->comp[k];
                if (c < 0 || c >= ndof) return -4;
                const int g = (id - 1) * ndof + c;
                const int dslot = csr_find(rowptr, cols, g, g);
                if (dslot < 0) return -1;
                const double scale = bc_row_scale(rowptr, K, g, dslot);
                for (int s = rowptr[g]; s < rowptr[g + 1]; s++) K[s] = 0.0;
                K[dslot] = scale;
                if (bc->type == 0) {
                    const int bd = (b1 - 1) * ndof + c;
                    const int ps = csr_find(rowptr, cols, g, bd);
                    if (ps < 0) return -2;
                    K[ps] = -scale;
                    R[g] = scale * (U[bd] - U[g]);
                } else {
                    const double value = bc->value[k] + bc->velocity[k] * t;
                    if (b1 >= 1 && !bc->direct) {
                        const int bd = (b1 - 1) * ndof + c;
                        const int ps = csr_find(rowptr, cols, g, bd);
                        if (ps < 0) return -2;
                        K[ps] = scale;
                        R[g] = scale * (2.0 * value - U[g] - U[bd]);
                    } else {
                        R[g] = scale * (value - U[g]);
                    }
                }
            }
        }
    }
    return 0;
}

/* Step schedule of TimeStepping::solveExplicitDynamics (TimeStepping.cpp:497-510) */
int pxo_schedule(double dt, double total, int cap, double *dts, double *ts)
{
    const double tEps = 1.0e-12 * total;
    double t = 0.0;
    int step = 0;
    while (t < total - tEps) {
        const double dtA = fmin(dt, total - t);
        const double tNew = t + dtA;
        if (step < cap) { if (dts) dts[step] = dtA; if (ts) ts[step] = tNew; }
        step++;
        t = tNew;
    }
    return step;
}

static void mark_breaks(long long nb, const double *health, int *break_step,
                        int step)
{
    if (!break_step) return;
    for (long long t = 0; t < nb; t++)
        if (health[t] <= 0.0 && break_step[t] == INT32_MAX) break_step[t] = step;
}

/* TimeStepping::solveExplicitDynamics (TimeStepping.cpp:417-599) */
int pxo_explicit_run(int n, const pxo_frac *p, const double *coords,
                     const int *off, const int *nbr, const double *opc,
                     const int *mirror, int nbc, const pxo_bc *bcs,
                     double dt, double total, double *u, double *uold,
                     const double *v0, double *health, int *break_step,
                     const double *node_E, int nthreads)
{
    const int dim = p->dim;
    const long long nd = (long long)n * dim;
    const long long nb = off[n];
    double *rate = malloc(sizeof(double) * (size_t)nd);
    if (break_step)
        for (long long t = 0; t < nb; t++) break_step[t] = (health[t] <= 0.0) ? -1 : INT32_MAX;
    /* t = 0 rate and u^{-1} = u^0 - dt v0 + dt^2/2 a0 (:479-487) */
    pxo_frac_accel(n, p, coords, off, nbr, opc, health, u, node_E, rate, nthreads);
    mark_breaks(nb, health, break_step, 0);
    {
        const double half = 0.5 * dt * dt;
        for (long long k = 0; k < nd; k++) uold[k] = u[k] - dt * (v0 ? v0[k] : 0.0) + half * rate[k];
    }
    const double tEps = 1.0e-12 * total;
    double umaxHist[50] = {0.0};
    double dtPrev = dt, t = 0.0;
    int step = 0, status = 0;
    while (t < total - tEps) {
        step++;
        const double dtA = fmin(dt, total - t);
        const double tNew = t + dtA;
        const double dt2 = dtA * dtA;
        pxo_frac_accel(n, p, coords, off, nbr, opc, health, u, node_E, rate, nthreads);
        mark_breaks(nb, health, break_step, step);
        if (dtA == dtPrev) {
            for (long long k = 0; k < nd; k++) {
                const double unew = 2.0 * u[k] - uold[k] + dt2 * rate[k];
                uold[k] = u[k]; u[k] = unew;
            }
        } else {
            const double rho = dtA / dtPrev;
            const double w = 0.5 * dtA * (dtA + dtPrev);
            for (long long k = 0; k < nd; k++) {
                const double unew = u[k] + rho * (u[k] - uold[k]) + w * rate[k];
                uold[k] = u[k]; u[k] = unew;
            }
        }
        pxo_bc_preset(n, dim, coords, mirror, nbc, bcs, u, tNew);
        /* blow-up guard (:550-573) */
        const int slot = (step - 1) % 50;
        const double past = (step > 50) ? umaxHist[slot] : -1.0;
        const double cap = (past >= 0.0) ? 1.0e3 * fmax(1.0, past) : INFINITY;
        double umax = 0.0; int finite = 1;
        for (long long k = 0; k < nd; k++) {
            const double av = fabs(u[k]);
            if (!(av <= DBL_MAX)) finite = 0;
            else if (av > umax) umax = av;
        }
        if (!finite || umax > cap) { status = -step; break; }
        umaxHist[slot] = umax;
        t = tNew; dtPrev = dtA;
    }
    free(rate);
    return status ? status : step;
}

/* ---- RandomIC::apply (RandomIC.cpp:13-42): std::mt19937_64 ------------- */
typedef struct { uint64_t mt[312]; int idx; } mt64;

static void mt64_seed(mt64 *m, uint64_t s)
{
    m->mt[0] = s;
    for (int i = 1; i < 312; i++)
        m->mt[i] = 6364136223846793005ULL * (m->mt[i - 1] ^ (m->mt[i - 1] >> 62)) + (uint64_t)i;
    m->idx = 312;
}

static uint64_t mt64_next(mt64 *m)
{
    if (m->idx >= 312) {
        for (int i = 0; i < 312; i++) {
            const uint64_t x = (m->mt[i] & 0xFFFFFFFF80000000ULL) | (m->mt[(i + 1) % 312] & 0x7FFFFFFFULL);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
            m->mt[i] = m->mt[(i + 156) % 312] ^ xa;
        }
        m->idx = 0;
    }
    uint64_t y = m->mt[m->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

void pxo_random_ic(uint64_t seed, double lo, double hi, int nnodes,
                   const int *nodes, int ndof, int ndofs_set, const int *dofs,
                   double *values)
{
    mt64 g; mt64_seed(&g, seed);
    const double inv53 = 1.0 / 9007199254740992.0;
    const double width = hi - lo;
    for (int q = 0; q < nnodes; q++) {
        const int base = (nodes[q] - 1) * ndof;
        for (int k = 0; k < ndofs_set; k++) {
            const uint64_t bits = mt64_next(&g) >> 11U;
            values[base + dofs[k] - 1] = lo + width * ((double)bits * inv53);
        }
    }
}

/* ---- a22: adjoint of the explicit step (new; the reference has none) ----
 * Frozen-health forward map, uniform steps (dt, nsteps = total/dt schedule):
 *   u^{-1} = u^0 - dt v0 + dt^2/2 a(u^0; h^0)
 *   w^k    = 2u^{k-1} - u^{k-2} + dt^2 a(u^{k-1}; h^k),  u^k = S w^k + g(t_k)
 *   a_i(u; h) = (E_i * (Khat_h u)_i + b) / rho,
 *   (Khat_h u)_i = sum_j h_ij Phat_ij (u_j - u_i) with Phat from C(E=1).
 * h^k_t = 1 iff break_step[t] > k (break_step -1 = seeded, 0 = t=0 check).
 * The forward uses C(E_i) exactly as pxo_frac_accel; the derivative uses
 * dC/dE = C(E=1) because C is linear in E. */

static void khat_apply(int n, int dim, const int *off, const int *nbr,
                       const double *opc, const int *break_step, int k,
                       const double *Cu, const double *u, double *out,
                       int nthreads)
{
    /* out_i = sum_j h P(Cu) (u_j - u_i) with constants Cu = {C11,C12,C66} */
    const int nops = (dim >= 3) ? 9 : 5;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++) {
        double L[3] = {0.0, 0.0, 0.0};
        for (int t = off[i]; t < off[i + 1]; t++) {
            if (!(break_step[t] > k)) continue;
            const int j = nbr[t] - 1;
            const double *op = opc + (size_t)t * nops;
            const double dux = u[j * dim] - u[i * dim], duy = u[j * dim + 1] - u[i * dim + 1];
            if (dim >= 3) {
                const double duz = u[j * dim + 2] - u[i * dim + 2];
                const double Gxx = op[3], Gyy = op[4], Gzz = op[5], Gxy = op[6], Gxz = op[7], Gyz = op[8];
                const double p11 = Cu[0] * Gxx + Cu[2] * (Gyy + Gzz);
                const double p22 = Cu[0] * Gyy + Cu[2] * (Gxx + Gzz);
                const double p33 = Cu[0] * Gzz + Cu[2] * (Gxx + Gyy);
                const double a = Cu[1] + Cu[2];
                const double p12 = a * Gxy, p13 = a * Gxz, p23 = a * Gyz;
                L[0] += p11 * dux + p12 * duy + p13 * duz;
                L[1] += p12 * dux + p22 * duy + p23 * duz;
                L[2] += p13 * dux + p23 * duy + p33 * duz;
            } else {
                const double Gxx = op[2], Gyy = op[3], Gxy = op[4];
                const double p11 = Cu[0] * Gxx + Cu[2] * Gyy;
                const double p22 = Cu[2] * Gxx + Cu[0] * Gyy;
                const double pxy = (Cu[1] + Cu[2]) * Gxy;
                L[0] += p11 * dux + pxy * duy;
                L[1] += pxy * dux + p22 * duy;
            }
        }
        for (int a = 0; a < dim; a++) out[(size_t)i * dim + a] = L[a];
    }
    (void)nthreads;
}

/* z = Khat^T y (serial scatter; Phat is symmetric per bond) */
static void khat_transpose(int n, int dim, const int *off, const int *nbr,
                           const double *opc, const int *break_step, int k,
                           const double *Cu, const double *y, double *z)
{
    const int nops = (dim >= 3) ? 9 : 5;
    memset(z, 0, sizeof(double) * (size_t)n * dim);
    for (int i = 0; i < n; i++) {
        for (int t = off[i]; t < off[i + 1]; t++) {
            if (!(break_step[t] > k)) continue;
            const int j = nbr[t] - 1;
            const double *op = opc + (size_t)t * nops;
            double P[3][3] = {{0}};
            if (dim >= 3) {
                const double Gxx = op[3], Gyy = op[4], Gzz = op[5], Gxy = op[6], Gxz = op[7], Gyz = op[8];
                const double a = Cu[1] + Cu[2];
                P[0][0] = Cu[0] * Gxx + Cu[2] * (Gyy + Gzz);
                P[1][1] = Cu[0] * Gyy + Cu[2] * (Gxx + Gzz);
                P[2][2] = Cu[0] * Gzz + Cu[2] * (Gxx + Gyy);
                P[0][1] = P[1][0] = a * Gxy; P[0][2] = P[2][0] = a * Gxz; P[1][2] = P[2][1] = a * Gyz;
            } else {
                const double Gxx = op[2], Gyy = op[3], Gxy = op[4];
                P[0][0] = Cu[0] * Gxx + Cu[2] * Gyy;
                P[1][1] = Cu[2] * Gxx + Cu[0] * Gyy;
                P[0][1] = P[1][0] = (Cu[1] + Cu[2]) * Gxy;
            }
            for (int a = 0; a < dim; a++) {
                double s = 0.0;
                for (int b = 0; b < dim; b++) s += P[b][a] * y[(size_t)i * dim + b];
                z[(size_t)j * dim + a] += s;
                z[(size_t)i * dim + a] -= s;
            }
        }
    }
}

/* S^T applied in place: reverse BC order, reverse node order. */
static void preset_transpose(int ndof, const double *coords, const int *mirror,
                             int nbc, const pxo_bc *bcs, double *lam)
{
    for (int e = nbc - 1; e >= 0; e--) {
        const pxo_bc *bc = bcs + e;
        for (int q = bc->nnodes - 1; q >= 0; q--) {
            const int id = bc->nodes[q];
            const int g = (id - 1) * ndof;
            const int b1 = mirror ? mirror[id - 1] : 0;
            if (bc->type == 1 && bc->has_box && !inside_box(bc, coords, id)) continue;
            if (bc->type == 0 && b1 < 1) continue;
            for (int k = bc->ncomp - 1; k >= 0; k--) {
                const int c = bc->comp[k];
                double alpha;
                if (bc->type == 0) alpha = 1.0;
                else alpha = (b1 >= 1 && !bc->direct) ? -1.0 : 0.0;
                const double lg = lam[g + c];
                lam[g + c] = 0.0;
                if (alpha != 0.0) lam[(b1 - 1) * ndof + c] += alpha * lg;
            }
        }
    }
}

static void accel_frozen(int n, const pxo_frac *p, const int *off, const int *nbr,
                         const double *opc, const int *break_step, int k,
                         const double *node_E, const double *u, double *tmp,
                         double *acc, int nthreads)
{
    /* a_i = (L_i + b)/rho with node constants C(E_i) (same as pxo_frac_accel
     * with frozen health) */
    const int dim = p->dim;
    const int nops = (dim >= 3) ? 9 : 5;
    (void)tmp;
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++) {
        double C11, C12, C66;
        node_consts(p, node_E, i, &C11, &C12, &C66);
        double L[3] = {0.0, 0.0, 0.0};
        for (int t = off[i]; t < off[i + 1]; t++) {
            if (!(break_step[t] > k)) continue;
            const int j = nbr[t] - 1;
            const double *op = opc + (size_t)t * nops;
            const double dux = u[j * dim] - u[i * dim], duy = u[j * dim + 1] - u[i * dim + 1];
            if (dim >= 3) {
                const double duz = u[j * dim + 2] - u[i * dim + 2];
                const double Gxx = op[3], Gyy = op[4], Gzz = op[5], Gxy = op[6], Gxz = op[7], Gyz = op[8];
                const double p11 = C11 * Gxx + C66 * (Gyy + Gzz);
                const double p22 = C11 * Gyy + C66 * (Gxx + Gzz);
                const double p33 = C11 * Gzz + C66 * (Gxx + Gyy);
                const double a = C12 + C66;
                const double p12 = a * Gxy, p13 = a * Gxz, p23 = a * Gyz;
                L[0] += 1.0 * (p11 * dux + p12 * duy + p13 * duz);
                L[1] += 1.0 * (p12 * dux + p22 * duy + p23 * duz);
                L[2] += 1.0 * (p13 * dux + p23 * duy + p33 * duz);
            } else {
                const double Gxx = op[2], Gyy = op[3], Gxy = op[4];
                const double p11 = C11 * Gxx + C66 * Gyy;
                const double p22 = C66 * Gxx + C11 * Gyy;
                const double pxy = (C12 + C66) * Gxy;
                L[0] += 1.0 * (p11 * dux + pxy * duy);
                L[1] += 1.0 * (pxy * dux + p22 * duy);
            }
        }
        const double b[3] = {p->bx, p->by, p->bz};
        for (int a = 0; a < dim; a++) acc[(size_t)i * dim + a] = (L[a] + b[a]) / p->rho;
    }
    (void)nthreads;
}

static double frozen_forward(int n, const pxo_frac *p, const double *coords,
                             const int *off, const int *nbr, const double *opc,
                             const int *mirror, int nbc, const pxo_bc *bcs,
                             double dt, int nsteps, const double *u0,
                             const double *v0, const int *break_step,
                             const double *node_E, int nsens, const int *sens,
                             int comp, double *hist, int nthreads)
{
    /* hist (optional): (nsteps+2) x n*dim, row 0 = u^{-1}, row k+1 = u^k */
    const int dim = p->dim;
    const size_t nd = (size_t)n * dim;
    double *um = malloc(sizeof(double) * nd), *uc = malloc(sizeof(double) * nd);
    double *acc = malloc(sizeof(double) * nd);
    memcpy(uc, u0, sizeof(double) * nd);
    accel_frozen(n, p, off, nbr, opc, break_step, 0, node_E, uc, NULL, acc, nthreads);
    const double half = 0.5 * dt * dt;
    for (size_t k = 0; k < nd; k++) um[k] = uc[k] - dt * (v0 ? v0[k] : 0.0) + half * acc[k];
    if (hist) { memcpy(hist, um, sizeof(double) * nd); memcpy(hist + nd, uc, sizeof(double) * nd); }
    double t = 0.0;
    const double dt2 = dt * dt;
    for (int s = 1; s <= nsteps; s++) {
        const double tNew = t + dt;
        accel_frozen(n, p, off, nbr, opc, break_step, s, node_E, uc, NULL, acc, nthreads);
        for (size_t k = 0; k < nd; k++) {
            const double unew = 2.0 * uc[k] - um[k] + dt2 * acc[k];
            um[k] = uc[k]; uc[k] = unew;
        }
        pxo_bc_preset(n, dim, coords, mirror, nbc, bcs, uc, tNew);
        if (hist) memcpy(hist + (size_t)(s + 1) * nd, uc, sizeof(double) * nd);
        t = tNew;
    }
    double J = 0.0;
    for (int q = 0; q < nsens; q++) { const double v = uc[(size_t)sens[q] * dim + comp]; J += v * v; }
    free(um); free(uc); free(acc);
    return J;
}

double pxo_frac_loss_frozen(int n, const pxo_frac *p, const double *coords,
                            const int *off, const int *nbr, const double *opc,
                            const int *mirror, int nbc, const pxo_bc *bcs,
                            double dt, int nsteps, const double *u0,
                            const double *v0, const int *break_step,
                            const double *node_E, int nsens, const int *sens,
                            int comp, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    if (!node_E) return NAN;
    return frozen_forward(n, p, coords, off, nbr, opc, mirror, nbc, bcs, dt,
                          nsteps, u0, v0, break_step, node_E,
                          nsens, sens, comp, NULL, nthreads);
}

double pxo_frac_adjoint(int n, const pxo_frac *p, const double *coords,
                        const int *off, const int *nbr, const double *opc,
                        const int *mirror, int nbc, const pxo_bc *bcs,
                        double dt, int nsteps, const double *u0,
                        const double *v0, const int *break_step,
                        const double *node_E, int nsens, const int *sens,
                        int comp, double *grad_e, double *grad_b, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    const int dim = p->dim;
    const size_t nd = (size_t)n * dim;
    const int K = nsteps;
    if (!node_E) return NAN;
    double *hist = malloc(sizeof(double) * nd * (size_t)(K + 2));
    const double J = frozen_forward(n, p, coords, off, nbr, opc, mirror, nbc, bcs, dt, K,
                                    u0, v0, break_step, node_E, nsens, sens, comp, hist,
                                    nthreads);
    double Cu[3];
    pxo_elastic_constants(dim, p->plane_stress, 1.0, p->nu, &Cu[0], &Cu[1], &Cu[2]);
    double *lc = calloc(nd, sizeof(double)), *lp = calloc(nd, sizeof(double));
    double *wb = malloc(sizeof(double) * nd), *ku = malloc(sizeof(double) * nd);
    double *z = malloc(sizeof(double) * nd), *ew = malloc(sizeof(double) * nd);
    memset(grad_e, 0, sizeof(double) * (size_t)n);
    double gb[3] = {0.0, 0.0, 0.0};
    const double *uK = hist + (size_t)(K + 1) * nd;
    for (int q = 0; q < nsens; q++) {
        const size_t idx = (size_t)sens[q] * dim + comp;
        lc[idx] += 2.0 * uK[idx];
    }
    const double c = dt * dt / p->rho;
    for (int k = K; k >= 1; k--) {
        const double *ukm1 = hist + (size_t)k * nd;      /* u^{k-1} */
        memcpy(wb, lc, sizeof(double) * nd);
        preset_transpose(dim, coords, mirror, nbc, bcs, wb);
        khat_apply(n, dim, off, nbr, opc, break_step, k, Cu, ukm1, ku, nthreads);
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int a = 0; a < dim; a++) s += ku[(size_t)i * dim + a] * wb[(size_t)i * dim + a];
            grad_e[i] += c * s;
            for (int a = 0; a < dim; a++) {
                ew[(size_t)i * dim + a] = node_E[i] * wb[(size_t)i * dim + a];
                gb[a] += c * wb[(size_t)i * dim + a];   /* u^k = ... + c (L + b) */
            }
        }
        khat_transpose(n, dim, off, nbr, opc, break_step, k, Cu, ew, z);
        for (size_t q = 0; q < nd; q++) {
            const double newprev = lp[q] + 2.0 * wb[q] + c * z[q];
            lp[q] = -wb[q];
            lc[q] = newprev;
        }
    }
    /* u^{-1} term: Ebar += (c/2) (Khat_0 u^0) . ubar^{-1} */
    khat_apply(n, dim, off, nbr, opc, break_step, 0, Cu, hist + nd, ku, nthreads);
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int a = 0; a < dim; a++) {
            s += ku[(size_t)i * dim + a] * lp[(size_t)i * dim + a];
            gb[a] += 0.5 * c * lp[(size_t)i * dim + a];   /* u^{-1} = ... + (c/2)(L + b) */
        }
        grad_e[i] += 0.5 * c * s;
    }
    if (grad_b)
        for (int a = 0; a < dim; a++) grad_b[a] = gb[a];
    free(hist); free(lc); free(lp); free(wb); free(ku); free(z); free(ew);
    return J;
}

/*
 * pxo_elements.c -- oracle (test infrastructure): Poisson, diffusion,
 * dynamic fracture, Cahn-Hilliard and stress-coupled Cahn-Hilliard CSR
 * assembly, restated from the reference's serial assembler.
 */
#include "pxo.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* SparseMatrix::findEntry (SparseMatrix.cpp:187-197): binary search */
static int find_slot(const int *rowptr, const int *cols, int row, int col)
{
    int lo = rowptr[row], hi = rowptr[row + 1];
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (cols[mid] < col) lo = mid + 1; else hi = mid;
    }
    return (lo < rowptr[row + 1] && cols[lo] == col) ? lo : -1;
}

/* ElementBase::pdLaplacian (ElementBase.h:317-321) on the full op vector */
static double lap_of(const double *op, int dim)
{
    if (dim >= 3) return op[3] + op[6] + op[8];   /* d2/dx2 + d2/dy2 + d2/dz2 */
    return op[2] + op[4];                        /* d2/dx2 + d2/dy2 */
}

/* PDSystem::formResidualAndJacobian (FormResidualAndJacobian.cpp:25-132)
 * with PoissonElement callbacks (PoissonElement.cpp:26-58). ndof = 1.
 * f_node (optional) generalises the reference's scalar source m_F. */
void pxo_assemble_poisson(int n, int dim, int nop, const int *off, const int *nbr,
                          const double *ops, const int *rowptr, const int *cols,
                          double sigma, const double *f_node, double f_const,
                          const double *u, double *kvals, double *rhs)
{
    memset(kvals, 0, sizeof(double) * (size_t)rowptr[n]);
    memset(rhs, 0, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        const int d = find_slot(rowptr, cols, i, i);
        rhs[i] += f_node ? f_node[i] : f_const;
        kvals[d] += -0.0;
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            const double coeff = sigma * lap_of(ops + (size_t)t * nop, dim);
            rhs[i] += coeff * (u[j] - u[i]);
            kvals[d] += -(-coeff);
            kvals[find_slot(rowptr, cols, i, j)] += -coeff;
        }
    }
}

/* DiffusionElement::computeBondResidualAndJacobian / computeNodalResidual-
 * AndJacobian (DiffusionElement.cpp:28-62) in the assembler of
 * FormResidualAndJacobian.cpp:25-132: bond coeff = -D * lap_ij,
 * R_i += coeff (u_j - u_i), K_II = -coeff, K_IJ = coeff; nodal
 * R = (u_i - uold_i)/dt - f, K = 1/dt; K = -dF/dU. */
void pxo_assemble_diffusion(int n, int dim, int nop, const int *off, const int *nbr,
                            const double *ops, const int *rowptr, const int *cols,
                            double D, double f, double dt, const double *u,
                            const double *uold, double *kvals, double *rhs)
{
    memset(kvals, 0, sizeof(double) * (size_t)rowptr[n]);
    memset(rhs, 0, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        const int d = find_slot(rowptr, cols, i, i);
        rhs[i] += (u[i] - uold[i]) / dt - f;
        kvals[d] += -(1.0 / dt);
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            const double coeff = -D * lap_of(ops + (size_t)t * nop, dim);
            rhs[i] += coeff * (u[j] - u[i]);
            kvals[d] += -(-coeff);
            kvals[find_slot(rowptr, cols, i, j)] += -coeff;
        }
    }
}

/* PDDODynamicFracElement::preprocessIteration, per-step commit
 * (PDDODynamicFracElement.cpp:148-204) */
void pxo_dynfrac_commit(int n, int dim, const double *coords, const int *off, const int *nbr,
                        double sc, int tension_only, int damage_on, double dt_prev,
                        const double *uold, double *uprev, double *vel, double *health)
{
    const size_t nd = (size_t)n * dim;
    for (size_t k = 0; k < nd; k++) {
        vel[k] = (uold[k] - uprev[k]) / dt_prev;
        uprev[k] = uold[k];
    }
    if (!(damage_on && sc > 0.0)) return;
    for (int i = 0; i < n; i++) {
        const double xi0 = coords[3 * i], yi0 = coords[3 * i + 1];
        const double zi0 = (dim >= 3) ? coords[3 * i + 2] : 0.0;
        for (int t = off[i]; t < off[i + 1]; t++) {
            if (health[t] <= 0.0) continue;
            const int j = nbr[t] - 1;
            double xi[3], eta[3];
            xi[0] = coords[3 * j] - xi0;
            xi[1] = coords[3 * j + 1] - yi0;
            xi[2] = (dim >= 3) ? coords[3 * j + 2] - zi0 : 0.0;
            eta[0] = uold[j * dim] - uold[i * dim];
            eta[1] = uold[j * dim + 1] - uold[i * dim + 1];
            eta[2] = (dim >= 3) ? uold[j * dim + 2] - uold[i * dim + 2] : 0.0;
            const double r = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
            if (r <= 0.0) continue;
            const double yx = xi[0] + eta[0], yy = xi[1] + eta[1], yz = xi[2] + eta[2];
            const double yr = sqrt(yx * yx + yy * yy + yz * yz);
            const double s = (yr - r) / r;
            if (tension_only ? (s > sc) : (fabs(s) > sc)) health[t] = 0.0;
        }
    }
}

/* PDDODynamicFracElement::computeBondResidualAndJacobian / computeNodal-
 * ResidualAndJacobian (PDDODynamicFracElement.cpp:207-297) in the assembler
 * of FormResidualAndJacobian.cpp:25-132: nodal R = b - c0 (u - uold - v dt),
 * K = -c0 I with c0 = rho/dt^2; per intact bond the Navier block
 * P = mu C:G(i,j), R = P (u_j - u_i), K_II = -P, K_IJ = P; K = -dF/dU. */
void pxo_assemble_dynfrac(int n, int dim, int nop, const int *off, const int *nbr,
                          const double *ops, const int *rowptr, const int *cols,
                          double C11, double C12, double C66, double rho, const double *b,
                          double dt, const double *vel, const double *health,
                          const double *u, const double *uold, double *kvals, double *rhs)
{
    memset(kvals, 0, sizeof(double) * (size_t)rowptr[n * dim]);
    memset(rhs, 0, sizeof(double) * (size_t)n * dim);
    const double c0 = rho / (dt * dt);
    for (int i = 0; i < n; i++) {
        const int ri = i * dim;
        for (int a = 0; a < dim; a++) {
            const double v = vel[ri + a];
            rhs[ri + a] += b[a] - c0 * (u[ri + a] - uold[ri + a] - v * dt);
            for (int bb = 0; bb < dim; bb++)
                kvals[find_slot(rowptr, cols, ri + a, ri + bb)] += -((a == bb) ? -c0 : 0.0);
        }
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1, rj = j * dim;
            const double *op = ops + (size_t)t * nop;
            const double mu = health[t];
            double R[3] = {0, 0, 0}, P[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            if (mu > 0.0) {
                if (dim >= 3) {
                    const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
                    const double aa = C12 + C66;
                    const double p11 = mu * (C11 * Gxx + C66 * (Gyy + Gzz));
                    const double p22 = mu * (C11 * Gyy + C66 * (Gxx + Gzz));
                    const double p33 = mu * (C11 * Gzz + C66 * (Gxx + Gyy));
                    const double p12 = mu * aa * Gxy, p13 = mu * aa * Gxz, p23 = mu * aa * Gyz;
                    const double dux = u[rj] - u[ri], duy = u[rj + 1] - u[ri + 1], duz = u[rj + 2] - u[ri + 2];
                    R[0] = p11 * dux + p12 * duy + p13 * duz;
                    R[1] = p12 * dux + p22 * duy + p23 * duz;
                    R[2] = p13 * dux + p23 * duy + p33 * duz;
                    P[0][0] = p11; P[0][1] = p12; P[0][2] = p13;
                    P[1][0] = p12; P[1][1] = p22; P[1][2] = p23;
                    P[2][0] = p13; P[2][1] = p23; P[2][2] = p33;
                } else {
                    const double Gxx = op[2], Gxy = op[3], Gyy = op[4];
                    const double p11 = mu * (C11 * Gxx + C66 * Gyy);
                    const double p22 = mu * (C66 * Gxx + C11 * Gyy);
                    const double pxy = mu * (C12 + C66) * Gxy;
                    const double dux = u[rj] - u[ri], duy = u[rj + 1] - u[ri + 1];
                    R[0] = p11 * dux + pxy * duy;
                    R[1] = pxy * dux + p22 * duy;
                    P[0][0] = p11; P[0][1] = pxy; P[1][0] = pxy; P[1][1] = p22;
                }
            }
            for (int a = 0; a < dim; a++) {
                rhs[ri + a] += R[a];
                for (int bb = 0; bb < dim; bb++) {
                    kvals[find_slot(rowptr, cols, ri + a, ri + bb)] += -(-P[a][bb]);
                    kvals[find_slot(rowptr, cols, ri + a, rj + bb)] += -P[a][bb];
                }
            }
        }
    }
}

/* CahnHilliardElement::preprocessIteration one-off geometry
 * (CahnHilliardElement.cpp:93-153). */
void pxo_ch_geometry(int n, int dim, int nop, const double *coords,
                     const int *off, const int *nbr, const double *ops,
                     const int *mirror, double *lapbond, int *rev,
                     unsigned char *isghost)
{
    for (int i = 0; i < n; i++) isghost[i] = (mirror && mirror[i] > 0) ? 1 : 0;
    for (int i = 0; i < n; i++) {
        const double xi0 = coords[3 * i], yi0 = coords[3 * i + 1];
        const double zi0 = (dim >= 3) ? coords[3 * i + 2] : 0.0;
        const double qi = xi0 * xi0 + yi0 * yi0 + zi0 * zi0;
        double lapq = 0.0;
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            const double w = lap_of(ops + (size_t)t * nop, dim);
            lapbond[t] = w;
            if (!isghost[j]) {
                const double xj = coords[3 * j], yj = coords[3 * j + 1];
                const double zj = (dim >= 3) ? coords[3 * j + 2] : 0.0;
                lapq += w * (xj * xj + yj * yj + zj * zj - qi);
            }
        }
        const double lam = (lapq > 1.0e-12) ? (2.0 * (double)dim / lapq) : 1.0;
        for (int t = off[i]; t < off[i + 1]; t++)
            if (!isghost[nbr[t] - 1]) lapbond[t] *= lam;
    }
    for (int i = 0; i < n; i++)
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            int r = -1;
            for (int s = off[j]; s < off[j + 1]; s++)
                if (nbr[s] == i + 1) { r = s; break; }
            rev[t] = r;
        }
}

static double clampc(double c)
{
    const double e = 1.0e-8;
    return fmax(e, fmin(1.0 - e, c));
}

/* mobility M(c)=M0 c(1-c) (CahnHilliardElement.cpp:46-49), per node */
void pxo_ch_mobility(int n, const double *u, double M0, double *mnode)
{
    for (int i = 0; i < n; i++) {
        const double cc = clampc(u[2 * i]);
        mnode[i] = M0 * cc * (1.0 - cc);
    }
}

/* Serial assembler with the CH callbacks (CahnHilliardElement.cpp:178-292).
 * ndof = 2 (c, mu). */
void pxo_assemble_ch(int n, const int *off, const int *nbr, const int *rev,
                     const double *lapbond, const unsigned char *isghost,
                     const double *vol, const double *mnode,
                     double kappa, double chi, double dt,
                     const int *rowptr, const int *cols,
                     const double *u, const double *uold,
                     double *kvals, double *rhs)
{
    memset(kvals, 0, sizeof(double) * (size_t)rowptr[2 * n]);
    memset(rhs, 0, sizeof(double) * (size_t)n * 2);
    for (int i = 0; i < n; i++) {
        const int r0 = 2 * i, r1 = 2 * i + 1;
        const int d00 = find_slot(rowptr, cols, r0, r0), d01 = find_slot(rowptr, cols, r0, r1);
        const int d10 = find_slot(rowptr, cols, r1, r0), d11 = find_slot(rowptr, cols, r1, r1);
        const double c = u[r0], mu = u[r1], cold = uold[r0];
        const double cc = clampc(c);
        const double fp = log(cc / (1.0 - cc)) + chi * (1.0 - 2.0 * cc);
        const double fpp = 1.0 / (cc * (1.0 - cc)) - 2.0 * chi;
        /* nodal block: R=[(c-cold)/dt, mu-f'], K_node=[[1/dt,0],[-f'',1]] */
        rhs[r0] += (c - cold) / dt;
        rhs[r1] += mu - fp;
        kvals[d00] += -(1.0 / dt);
        kvals[d01] += -0.0;
        kvals[d10] += -(-fpp);
        kvals[d11] += -1.0;
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            const int ghostJ = isghost[j];
            const double g2ij = lapbond[t];
            double coeff_c = 0.0;
            if (!ghostJ) {
                const double Mi = mnode[i], Mj = mnode[j], Vi = vol[i], Vj = vol[j];
                const int rv = rev[t];
                double wij;
                if (rv >= 0 && Vi > 0.0) wij = 0.5 * (Mi * g2ij + (Vj / Vi) * Mj * lapbond[rv]);
                else wij = Mi * g2ij;
                coeff_c = -wij;
            }
            const double coeff_mu = ghostJ ? 0.0 : (kappa * g2ij);
            const double dC = u[2 * j] - u[r0];
            const double dMu = u[2 * j + 1] - u[r1];
            /* LocalR = [coeff_c dMu, coeff_mu dC];
             * K_II = [[0,-coeff_c],[-coeff_mu,0]], K_IJ = -K_II */
            rhs[r0] += coeff_c * dMu;
            rhs[r1] += coeff_mu * dC;
            const int j0 = 2 * j, j1 = 2 * j + 1;
            kvals[d00] += -0.0;
            kvals[d01] += -(-coeff_c);
            kvals[d10] += -(-coeff_mu);
            kvals[d11] += -0.0;
            kvals[find_slot(rowptr, cols, r0, j0)] += -0.0;
            kvals[find_slot(rowptr, cols, r0, j1)] += -coeff_c;
            kvals[find_slot(rowptr, cols, r1, j0)] += -coeff_mu;
            kvals[find_slot(rowptr, cols, r1, j1)] += -0.0;
        }
    }
}

/* StressCahnHilliardElement::recalcElasticConstants
 * (StressCahnHilliardElement.cpp:80-128): out = C11, C12, C66, A, Kh, Ch.
 * Returns 0, or -1 for a Poisson ratio the reference rejects. */
int pxo_stressch_consts(int dim, int plane_stress, double E, double nu, double Omega, double *out)
{
    for (int k = 0; k < 6; k++) out[k] = 0.0;
    if (E <= 0.0) return 0;
    if (dim < 3 && plane_stress) {
        const double d = 1.0 - nu * nu;
        if (d <= 0.0) return -1;
        out[0] = E / d;
        out[1] = E * nu / d;
        out[2] = E / (2.0 * (1.0 + nu));
        out[3] = E * Omega / (3.0 * (1.0 - nu));
        out[4] = (out[0] + out[1]) / 3.0;
        out[5] = -2.0 * out[3] / 3.0;
        return 0;
    }
    const double d = (1.0 + nu) * (1.0 - 2.0 * nu);
    if (d <= 0.0) return -1;
    out[0] = E * (1.0 - nu) / d;
    out[1] = E * nu / d;
    out[2] = E / (2.0 * (1.0 + nu));
    out[3] = E * Omega / (3.0 * (1.0 - 2.0 * nu));
    out[4] = (out[0] + 2.0 * out[1]) / 3.0;
    out[5] = -out[3];
    return 0;
}

/* regLog (StressCahnHilliardElement.cpp:42-46): ln(c/(1-c)) and its
 * derivative inside [delta, 1-delta], continued by the tangent line outside. */
static void reg_log(double c, double delta, double *L, double *Lp)
{
    if (c < delta) {
        const double Lpd = 1.0 / (delta * (1.0 - delta));
        *L = log(delta / (1.0 - delta)) + Lpd * (c - delta);
        *Lp = Lpd;
    } else if (c > 1.0 - delta) {
        const double cb = 1.0 - delta, Lpb = 1.0 / (cb * (1.0 - cb));
        *L = log(cb / (1.0 - cb)) + Lpb * (c - cb);
        *Lp = Lpb;
    } else {
        *L = log(c / (1.0 - c));
        *Lp = 1.0 / (c * (1.0 - c));
    }
}

/* StressCahnHilliardElement through the serial assembler
 * (FormResidualAndJacobian.cpp:25-132): Picard mobility per node
 * (StressCahnHilliardElement.cpp:253-278, only M enters the bond term), bond
 * term (:282-428) with the lambda-corrected, ghost-excluded Laplacian of
 * pxo_ch_geometry, nodal term (:431-467). ndof = dim + 2 (c, mu, u).
 * k = C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref. */
void pxo_assemble_stressch(int n, int dim, int nop, const int *off, const int *nbr,
                           const double *ops, const int *rev, const double *lapbond,
                           const unsigned char *isghost, const double *vol, const double *k,
                           double dt, const int *rowptr, const int *cols,
                           const double *u, const double *uold, double *kvals, double *rhs)
{
    const int nd = dim + 2;
    const double C11 = k[0], C12 = k[1], C66 = k[2], A = k[3], Kh = k[4], Ch = k[5];
    const double D = k[6], kappa = k[7], chi = k[8], Omega = k[9], cref = k[10];
    double *mn = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        const double cc = clampc(u[(size_t)nd * i]);
        mn[i] = D * cc * (1.0 - cc);
    }
    memset(kvals, 0, sizeof(double) * (size_t)rowptr[(size_t)nd * n]);
    memset(rhs, 0, sizeof(double) * (size_t)n * nd);
    double KL[5][5], R[5];
    for (int i = 0; i < n; i++) {
        const size_t ri = (size_t)nd * i;
        /* nodal term */
        const double c = u[ri], mu = u[ri + 1], cold = uold[ri];
        double L, Lp;
        reg_log(c, 1.0e-3, &L, &Lp);
        const double fp = L + chi * (1.0 - 2.0 * c);
        const double fpp = Lp - 2.0 * chi;
        memset(KL, 0, sizeof KL);
        memset(R, 0, sizeof R);
        R[0] = (c - cold) / dt;
        R[1] = mu - fp + Omega * Ch * (c - cref);
        KL[0][0] = 1.0 / dt;
        KL[1][0] = -fpp + Omega * Ch;
        KL[1][1] = 1.0;
        for (int a = 0; a < nd; a++) {
            rhs[ri + a] += R[a];
            for (int b = 0; b < nd; b++)
                kvals[find_slot(rowptr, cols, (int)ri + a, (int)ri + b)] += -KL[a][b];
        }
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            const size_t rj = (size_t)nd * j;
            const double *op = ops + (size_t)t * nop;
            const int ghostJ = isghost[j];
            const double g2ij = lapbond[t];
            double coeff_c = 0.0;
            if (!ghostJ) {
                const double Mi = mn[i], Mj = mn[j], Vi = vol[i], Vj = vol[j];
                const int rv = rev[t];
                double wij;
                if (rv >= 0 && Vi > 0.0) wij = 0.5 * (Mi * g2ij + (Vj / Vi) * Mj * lapbond[rv]);
                else wij = Mi * g2ij;
                coeff_c = -wij;
            }
            const double kl = ghostJ ? 0.0 : (kappa * g2ij);
            double d[5];
            for (int a = 0; a < nd; a++) d[a] = u[rj + a] - u[ri + a];
            /* P = the bond's dR/du_j; K_II = -P, K_IJ = P */
            memset(KL, 0, sizeof KL);
            if (dim >= 3) {
                const double Gx = op[0], Gy = op[1], Gz = op[2];
                const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
                const double aa = C12 + C66;
                const double cu[3] = {Omega * Kh * Gx, Omega * Kh * Gy, Omega * Kh * Gz};
                const double p11 = C11 * Gxx + C66 * (Gyy + Gzz);
                const double p22 = C11 * Gyy + C66 * (Gxx + Gzz);
                const double p33 = C11 * Gzz + C66 * (Gxx + Gyy);
                const double p12 = aa * Gxy, p13 = aa * Gxz, p23 = aa * Gyz;
                const double cx = -A * Gx, cy = -A * Gy, cz = -A * Gz;
                R[0] = coeff_c * d[1];
                R[1] = kl * d[0] + cu[0] * d[2] + cu[1] * d[3] + cu[2] * d[4];
                R[2] = cx * d[0] + p11 * d[2] + p12 * d[3] + p13 * d[4];
                R[3] = cy * d[0] + p12 * d[2] + p22 * d[3] + p23 * d[4];
                R[4] = cz * d[0] + p13 * d[2] + p23 * d[3] + p33 * d[4];
                KL[0][1] = coeff_c;
                KL[1][0] = kl; KL[1][2] = cu[0]; KL[1][3] = cu[1]; KL[1][4] = cu[2];
                KL[2][0] = cx; KL[2][2] = p11; KL[2][3] = p12; KL[2][4] = p13;
                KL[3][0] = cy; KL[3][2] = p12; KL[3][3] = p22; KL[3][4] = p23;
                KL[4][0] = cz; KL[4][2] = p13; KL[4][3] = p23; KL[4][4] = p33;
            } else {
                const double Gx = op[0], Gy = op[1], Gxx = op[2], Gxy = op[3], Gyy = op[4];
                const double cux = Omega * Kh * Gx, cuy = Omega * Kh * Gy;
                const double p11 = C11 * Gxx + C66 * Gyy;
                const double p22 = C66 * Gxx + C11 * Gyy;
                const double pxy = (C12 + C66) * Gxy;
                const double cx = -A * Gx, cy = -A * Gy;
                R[0] = coeff_c * d[1];
                R[1] = kl * d[0] + cux * d[2] + cuy * d[3];
                R[2] = cx * d[0] + p11 * d[2] + pxy * d[3];
                R[3] = cy * d[0] + pxy * d[2] + p22 * d[3];
                KL[0][1] = coeff_c;
                KL[1][0] = kl; KL[1][2] = cux; KL[1][3] = cuy;
                KL[2][0] = cx; KL[2][2] = p11; KL[2][3] = pxy;
                KL[3][0] = cy; KL[3][2] = pxy; KL[3][3] = p22;
            }
            for (int a = 0; a < nd; a++) {
                rhs[ri + a] += R[a];
                for (int b = 0; b < nd; b++) {
                    kvals[find_slot(rowptr, cols, (int)(ri + a), (int)(ri + b))] += -(-KL[a][b]);
                    kvals[find_slot(rowptr, cols, (int)(ri + a), (int)(rj + b))] += -KL[a][b];
                }
            }
        }
    }
    free(mn);
}

/* FracStressCahnHilliardElement species Laplacian (FracStressCahnHilliard-
 * Element.cpp:222-268): the full-family pdLaplacian per bond, zero on bonds
 * to boundary ghosts, no consistency rescaling.  The reverse-bond slots and
 * ghost flags are those of pxo_ch_geometry. */
void pxo_fsch_lapc(int n, int dim, int nop, const int *off, const int *nbr, const double *ops,
                   const int *mirror, double *lapc)
{
    for (int i = 0; i < n; i++)
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            lapc[t] = (mirror && mirror[j] > 0) ? 0.0 : lap_of(ops + (size_t)t * nop, dim);
        }
}

/* Per-step history commit on the accepted u_old (FracStressCahnHilliard-
 * Element.cpp:303-398): with rho > 0 the velocity (u_old - u_prev)/dt_prev
 * and u_prev = u_old of the displacement dofs; then, with damage on and
 * sc > 0, irreversible failure of every intact bond whose elastic stretch
 * (geometric stretch minus the swelling (Omega/3)(c_bar - cref)) reaches sc,
 * positive part only when tension_only.  Both ends of a bond see the same
 * stretch, so the reciprocal break of the reference is the per-node test
 * applied at the other end. */
void pxo_fsch_commit(int n, int dim, const double *coords, const int *off, const int *nbr,
                     double sc, int tension_only, int damage_on, double Omega, double cref,
                     double rho, double dt_prev, const double *uold, double *uprev, double *vel,
                     double *health)
{
    const int nd = dim + 2;
    if (rho > 0.0)
        for (int i = 0; i < n; i++)
            for (int a = 0; a < dim; a++) {
                const double ua = uold[(size_t)nd * i + 2 + a];
                vel[(size_t)dim * i + a] = (ua - uprev[(size_t)dim * i + a]) / dt_prev;
                uprev[(size_t)dim * i + a] = ua;
            }
    if (!(damage_on && sc > 0.0)) return;
    for (int i = 0; i < n; i++) {
        const size_t ri = (size_t)nd * i;
        const double ci = uold[ri];
        const double xi0 = coords[3 * i], yi0 = coords[3 * i + 1];
        const double zi0 = (dim >= 3) ? coords[3 * i + 2] : 0.0;
        for (int t = off[i]; t < off[i + 1]; t++) {
            if (health[t] <= 0.0) continue;
            const int j = nbr[t] - 1;
            const size_t rj = (size_t)nd * j;
            double xi[3], eta[3];
            xi[0] = coords[3 * j] - xi0;
            xi[1] = coords[3 * j + 1] - yi0;
            xi[2] = (dim >= 3) ? coords[3 * j + 2] - zi0 : 0.0;
            eta[0] = uold[rj + 2] - uold[ri + 2];
            eta[1] = uold[rj + 3] - uold[ri + 3];
            eta[2] = (dim >= 3) ? uold[rj + 4] - uold[ri + 4] : 0.0;
            const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
            if (r2 <= 0.0) continue;
            const double r = sqrt(r2);
            const double rinv = 1.0 / r;
            const double yx = xi[0] + eta[0], yy = xi[1] + eta[1], yz = xi[2] + eta[2];
            const double sn = (sqrt(yx * yx + yy * yy + yz * yz) - r) * rinv;
            const double cbar = 0.5 * (ci + uold[rj]);
            const double s_el = sn - (Omega / 3.0) * (cbar - cref);
            const double so = tension_only ? fmax(s_el, 0.0) : fabs(s_el);
            if (so >= sc) health[t] = 0.0;
        }
    }
}

/* FracStressCahnHilliardElement through the serial assembler
 * (FormResidualAndJacobian.cpp:25-132): bond term (FracStressCahnHilliard-
 * Element.cpp:411-526: species flux and kappa Laplacian ungated, mechanics
 * and the sigma_h coupling gated by 1 or the residual stiffness), nodal term
 * (:528-561: clamped logarithm, optional inertia with the committed
 * velocity).  k = C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref, rho,
 * residual stiffness. */
void pxo_assemble_fsch(int n, int dim, int nop, const int *off, const int *nbr,
                       const double *ops, const int *rev, const double *lapc,
                       const unsigned char *isghost, const double *vol, const double *k,
                       double dt, const double *vel, const double *health,
                       const int *rowptr, const int *cols, const double *u, const double *uold,
                       double *kvals, double *rhs)
{
    const int nd = dim + 2;
    const double C11 = k[0], C12 = k[1], C66 = k[2], A = k[3], Kh = k[4], Ch = k[5];
    const double D = k[6], kappa = k[7], chi = k[8], Omega = k[9], cref = k[10];
    const double rho = k[11], resid = k[12];
    double *mn = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        const double cc = clampc(u[(size_t)nd * i]);
        mn[i] = D * cc * (1.0 - cc);
    }
    memset(kvals, 0, sizeof(double) * (size_t)rowptr[(size_t)nd * n]);
    memset(rhs, 0, sizeof(double) * (size_t)n * nd);
    double KL[5][5], R[5];
    for (int i = 0; i < n; i++) {
        const size_t ri = (size_t)nd * i;
        const double c = u[ri], mu = u[ri + 1], cold = uold[ri];
        const double cc = clampc(c);
        const double fp = log(cc / (1.0 - cc)) + chi * (1.0 - 2.0 * cc);
        const double fpp = 1.0 / (cc * (1.0 - cc)) - 2.0 * chi;
        memset(KL, 0, sizeof KL);
        memset(R, 0, sizeof R);
        R[0] = (c - cold) / dt;
        R[1] = mu - fp + Omega * Ch * (c - cref);
        KL[0][0] = 1.0 / dt;
        KL[1][0] = -fpp + Omega * Ch;
        KL[1][1] = 1.0;
        if (rho > 0.0) {
            const double c0 = rho / (dt * dt);
            for (int a = 0; a < dim; a++) {
                const double v = vel[(size_t)dim * i + a];
                R[2 + a] += -c0 * (u[ri + 2 + a] - uold[ri + 2 + a] - v * dt);
                KL[2 + a][2 + a] += -c0;
            }
        }
        for (int a = 0; a < nd; a++) {
            rhs[ri + a] += R[a];
            for (int b = 0; b < nd; b++)
                kvals[find_slot(rowptr, cols, (int)ri + a, (int)ri + b)] += -KL[a][b];
        }
        for (int t = off[i]; t < off[i + 1]; t++) {
            const int j = nbr[t] - 1;
            const size_t rj = (size_t)nd * j;
            const double *op = ops + (size_t)t * nop;
            const double g = (health[t] > 0.0) ? 1.0 : resid;
            double wf = 0.0, g2c = 0.0;
            if (!isghost[j]) {
                g2c = lapc[t];
                const int rv = rev[t];
                if (rv >= 0) wf = 0.5 * (mn[i] * g2c + (vol[j] / vol[i]) * mn[j] * lapc[rv]);
                else wf = mn[i] * g2c;
            }
            const double coeff_c = -wf, kl = kappa * g2c;
            double d[5];
            for (int a = 0; a < nd; a++) d[a] = u[rj + a] - u[ri + a];
            memset(KL, 0, sizeof KL);
            if (dim >= 3) {
                const double Gx = op[0], Gy = op[1], Gz = op[2];
                const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
                const double aa = C12 + C66;
                const double cux = g * Omega * Kh * Gx, cuy = g * Omega * Kh * Gy, cuz = g * Omega * Kh * Gz;
                const double p11 = g * (C11 * Gxx + C66 * (Gyy + Gzz));
                const double p22 = g * (C11 * Gyy + C66 * (Gxx + Gzz));
                const double p33 = g * (C11 * Gzz + C66 * (Gxx + Gyy));
                const double p12 = g * aa * Gxy, p13 = g * aa * Gxz, p23 = g * aa * Gyz;
                const double cx = -g * A * Gx, cy = -g * A * Gy, cz = -g * A * Gz;
                R[0] = coeff_c * d[1];
                R[1] = kl * d[0] + cux * d[2] + cuy * d[3] + cuz * d[4];
                R[2] = cx * d[0] + p11 * d[2] + p12 * d[3] + p13 * d[4];
                R[3] = cy * d[0] + p12 * d[2] + p22 * d[3] + p23 * d[4];
                R[4] = cz * d[0] + p13 * d[2] + p23 * d[3] + p33 * d[4];
                KL[0][1] = coeff_c;
                KL[1][0] = kl; KL[1][2] = cux; KL[1][3] = cuy; KL[1][4] = cuz;
                KL[2][0] = cx; KL[2][2] = p11; KL[2][3] = p12; KL[2][4] = p13;
                KL[3][0] = cy; KL[3][2] = p12; KL[3][3] = p22; KL[3][4] = p23;
                KL[4][0] = cz; KL[4][2] = p13; KL[4][3] = p23; KL[4][4] = p33;
            } else {
                const double Gx = op[0], Gy = op[1], Gxx = op[2], Gxy = op[3], Gyy = op[4];
                const double cux = g * Omega * Kh * Gx, cuy = g * Omega * Kh * Gy;
                const double p11 = g * (C11 * Gxx + C66 * Gyy);
                const double p22 = g * (C66 * Gxx + C11 * Gyy);
                const double pxy = g * (C12 + C66) * Gxy;
                const double cx = -g * A * Gx, cy = -g * A * Gy;
                R[0] = coeff_c * d[1];
                R[1] = kl * d[0] + cux * d[2] + cuy * d[3];
                R[2] = cx * d[0] + p11 * d[2] + pxy * d[3];
                R[3] = cy * d[0] + pxy * d[2] + p22 * d[3];
                KL[0][1] = coeff_c;
                KL[1][0] = kl; KL[1][2] = cux; KL[1][3] = cuy;
                KL[2][0] = cx; KL[2][2] = p11; KL[2][3] = pxy;
                KL[3][0] = cy; KL[3][2] = pxy; KL[3][3] = p22;
            }
            for (int a = 0; a < nd; a++) {
                rhs[ri + a] += R[a];
                for (int b = 0; b < nd; b++) {
                    kvals[find_slot(rowptr, cols, (int)(ri + a), (int)(ri + b))] += -(-KL[a][b]);
                    kvals[find_slot(rowptr, cols, (int)(ri + a), (int)(rj + b))] += -KL[a][b];
                }
            }
        }
    }
    free(mn);
}

/* Output projections of the stress-coupled Cahn-Hilliard elements
 * (StressCahnHilliardElement.cpp:481-658, FracStressCahnHilliardElement.cpp:
 * 47-110, 575-672).  which: 0 strain, 1 stress, 2 vonMises, 3 sigmaH,
 * 4 fPrime, 5 fDoublePrime, 6 gradC, 7 damage, 8 DamageI; out holds 9
 * (tensors), 3 (gradC) or 1 value per node.  k = C11, C12, C66, A, cref,
 * chi.  frac selects the fracture element (clamped logarithm; damage from
 * health, DamageI from the bonds broken after the crack seed). */
void pxo_sch_project(int n, int dim, int nop, const int *off, const int *nbr, const double *ops,
                     const double *k, int frac, int plane_strain, const double *health,
                     const double *seed, const double *u, int which, double *out)
{
    const int nd = dim + 2;
    const double C11 = k[0], C12 = k[1], C66 = k[2], A = k[3], cref = k[4], chi = k[5];
    const int nc = (which <= 1) ? 9 : (which == 6) ? 3 : 1;
    for (int i = 0; i < n; i++) {
        double *o = out + (size_t)nc * i;
        for (int q = 0; q < nc; q++) o[q] = 0.0;
        const size_t ri = (size_t)nd * i;
        const double c = u[ri];
        const int deg = off[i + 1] - off[i];
        if (which == 4 || which == 5) {
            double L, Lp;
            if (frac) {
                const double cc = clampc(c);
                L = log(cc / (1.0 - cc));
                Lp = 1.0 / (cc * (1.0 - cc));
                o[0] = (which == 4) ? L + chi * (1.0 - 2.0 * cc) : Lp - 2.0 * chi;
            } else {
                reg_log(c, 1.0e-3, &L, &Lp);
                o[0] = (which == 4) ? L + chi * (1.0 - 2.0 * c) : Lp - 2.0 * chi;
            }
            continue;
        }
        if (which == 7 || which == 8) {
            int cnt = 0;
            for (int t = off[i]; t < off[i + 1]; t++)
                if (health[t] <= 0.0 && (which == 7 || seed[t] > 0.0)) cnt++;
            double phi = deg > 0 ? (double)cnt / (double)deg : 0.0;
            if (which == 8) phi = fmin(fmax(phi, 0.0), 1.0);
            o[0] = phi;
            continue;
        }
        if (which == 6) {
            double gx = 0.0, gy = 0.0, gz = 0.0;
            for (int t = off[i]; t < off[i + 1]; t++) {
                const double *op = ops + (size_t)t * nop;
                const double dc = u[(size_t)nd * (nbr[t] - 1)] - c;
                gx += dc * op[0];
                gy += dc * op[1];
                if (dim >= 3) gz += dc * op[2];
            }
            o[0] = gx; o[1] = gy; o[2] = gz;
            continue;
        }
        const double dc = c - cref;
        if (dim >= 3) {
            double g[3][3] = {{0}};   /* g[a][b] = sum (u_a,j - u_a,i) G_b */
            for (int t = off[i]; t < off[i + 1]; t++) {
                const double *op = ops + (size_t)t * nop;
                const size_t rj = (size_t)nd * (nbr[t] - 1);
                for (int a = 0; a < 3; a++) {
                    const double du = u[rj + 2 + a] - u[ri + 2 + a];
                    for (int b = 0; b < 3; b++) g[a][b] += du * op[b];
                }
            }
            const double exx = g[0][0], eyy = g[1][1], ezz = g[2][2];
            const double exy = 0.5 * (g[0][1] + g[1][0]), exz = 0.5 * (g[0][2] + g[2][0]);
            const double eyz = 0.5 * (g[1][2] + g[2][1]);
            const double sxx = C11 * exx + C12 * (eyy + ezz) - A * dc;
            const double syy = C11 * eyy + C12 * (exx + ezz) - A * dc;
            const double szz = C11 * ezz + C12 * (exx + eyy) - A * dc;
            const double sxy = 2.0 * C66 * exy, sxz = 2.0 * C66 * exz, syz = 2.0 * C66 * eyz;
            if (which == 0) {
                const double e9[9] = {exx, exy, exz, exy, eyy, eyz, exz, eyz, ezz};
                for (int q = 0; q < 9; q++) o[q] = e9[q];
            } else if (which == 1) {
                const double s9[9] = {sxx, sxy, sxz, sxy, syy, syz, sxz, syz, szz};
                for (int q = 0; q < 9; q++) o[q] = s9[q];
            } else if (which == 3) {
                o[0] = (sxx + syy + szz) / 3.0;
            } else if (which == 2) {
                const double a = sxx - syy, b = syy - szz, cc = szz - sxx;
                o[0] = sqrt(0.5 * (a * a + b * b + cc * cc + 6.0 * (sxy * sxy + sxz * sxz + syz * syz)));
            }
            continue;
        }
        double dxx = 0.0, dxy = 0.0, dyx = 0.0, dyy = 0.0;
        for (int t = off[i]; t < off[i + 1]; t++) {
            const double *op = ops + (size_t)t * nop;
            const size_t rj = (size_t)nd * (nbr[t] - 1);
            const double dux = u[rj + 2] - u[ri + 2], duy = u[rj + 3] - u[ri + 3];
            dxx += dux * op[0];
            dxy += dux * op[1];
            dyx += duy * op[0];
            dyy += duy * op[1];
        }
        const double exx = dxx, eyy = dyy, exy = 0.5 * (dxy + dyx);
        const double sxx = C11 * exx + C12 * eyy - A * dc;
        const double syy = C12 * exx + C11 * eyy - A * dc;
        const double sxy = 2.0 * C66 * exy;
        const double szz = plane_strain ? (C12 * (exx + eyy) - A * dc) : 0.0;
        if (which == 0) {
            o[0] = exx; o[1] = exy; o[3] = exy; o[4] = eyy;
        } else if (which == 1) {
            o[0] = sxx; o[1] = sxy; o[3] = sxy; o[4] = syy; o[8] = szz;
        } else if (which == 3) {
            o[0] = (sxx + syy + szz) / 3.0;
        } else if (which == 2) {
            const double d1 = sxx - syy, d2 = syy - szz, d3 = szz - sxx;
            o[0] = sqrt(0.5 * (d1 * d1 + d2 * d2 + d3 * d3 + 6.0 * sxy * sxy));
        }
    }
}

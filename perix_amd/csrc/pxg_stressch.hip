// pxg_stressch.hip -- the stress-coupled Cahn-Hilliard elements
// (StressCahnHilliardElement, FracStressCahnHilliardElement) on the device,
// and the block-element assembly kernels they share with the Cahn-Hilliard
// and dynamic fracture elements: CSR assembly of the
// (c, mu, u) system with the conservative degenerate-mobility species flux,
// the lambda-corrected ghost-excluded gradient-energy Laplacian, the
// chemo-mechanical coupling and the PDDO Navier operator
// (StressCahnHilliardElement.cpp:139-467 through the assembler of
// FormResidualAndJacobian.cpp:25-132).
//
// One wavefront per node, lanes over its bonds (see k_sch): every sum of the
// node's (dim+2)x(dim+2) diagonal block and residual is taken in the
// reference's order (nodal term first, then the bonds in family order), so K
// comes out bit for bit; the residual's mu rows carry ln(c/(1-c)), whose
// device logarithm can differ from the host's in the last bit.  The
// Picard mobility M(c) = D c (1-c) is evaluated from u at both ends of each
// bond, as the reference's per-iteration cache holds it.
//
// The fracture variant adds irreversible bond failure on the elastic
// (swelling-corrected) stretch, committed once per step on u_old, a health
// gate with a residual-stiffness floor on the mechanical and sigma_h terms,
// optional inertia with the committed velocity, and its own species
// Laplacian (ghost bonds zeroed, no consistency rescaling).
#include <algorithm>
#include "pxg_internal.h"

using namespace pxg;

namespace {

unsigned grids(long long n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ double sch_mobility(double c, double D)
{
    // mobility (StressCahnHilliardElement.cpp:61-64) with the 1e-8 clamp
    const double cc = fmax(1.0e-8, fmin(1.0 - 1.0e-8, c));
    return D * cc * (1.0 - cc);
}

// regLog (StressCahnHilliardElement.cpp:42-46)
__device__ __forceinline__ void sch_reglog(double c, double &L, double &Lp)
{
    const double delta = 1.0e-3;
    if (c < delta) {
        const double Lpd = 1.0 / (delta * (1.0 - delta));
        L = log(delta / (1.0 - delta)) + Lpd * (c - delta);
        Lp = Lpd;
    } else if (c > 1.0 - delta) {
        const double cb = 1.0 - delta, Lpb = 1.0 / (cb * (1.0 - cb));
        L = log(cb / (1.0 - cb)) + Lpb * (c - cb);
        Lp = Lpb;
    } else {
        L = log(c / (1.0 - c));
        Lp = 1.0 / (c * (1.0 - c));
    }
}

// ---- fracture variant (FracStressCahnHilliardElement) ----------------------

// species Laplacian: the full-family pdLaplacian, zero on bonds to ghosts
// (FracStressCahnHilliardElement.cpp:238-268; no consistency rescaling)
__global__ void k_fsch_lapc(long long nb, int dim, int nop, const int *nbr, const double *ops,
                            const int *mirror, double *lapc)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const double *op = ops + (size_t)t * nop;
    const double w = (dim >= 3) ? op[3] + op[6] + op[8] : op[2] + op[4];
    lapc[t] = (mirror[nbr[t] - 1] > 0) ? 0.0 : w;
}

__global__ void k_fsch_init(long long nb, int n, int dim, const unsigned char *seed, const double *uold,
                            unsigned char *health, double *uprev, double *vel)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb) health[t] = seed[t];
    if (t < (long long)n * dim) {
        vel[t] = 0.0;
        if (uold) uprev[t] = uold[(t / dim) * (dim + 2) + 2 + t % dim];
    }
}

// per-step commit on the accepted u_old (FracStressCahnHilliardElement.cpp:303-398)
template <int DIM>
__global__ void k_fsch_commit(int n, const double *coords, const int *off, const int *nbr, double sc,
                              int tension_only, int damage_on, double Omega, double cref, double rho,
                              double dt_prev, const double *uold, double *uprev, double *vel,
                              unsigned char *health)
{
    constexpr int ND = DIM + 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t ri = (size_t)i * ND;
    if (rho > 0.0) {
        for (int a = 0; a < DIM; a++) {
            const double ua = uold[ri + 2 + a];
            vel[(size_t)i * DIM + a] = (ua - uprev[(size_t)i * DIM + a]) / dt_prev;
            uprev[(size_t)i * DIM + a] = ua;
        }
    }
    if (!(damage_on && sc > 0.0)) return;
    // both ends of a bond compute the same stretch (negated xi and eta), so
    // testing every bond from its own end reproduces the reference's
    // reciprocal break
    const double ci = uold[ri];
    const double xi0 = coords[3 * (size_t)i], yi0 = coords[3 * (size_t)i + 1];
    const double zi0 = (DIM >= 3) ? coords[3 * (size_t)i + 2] : 0.0;
    for (int t = off[i]; t < off[i + 1]; t++) {
        if (!health[t]) continue;
        const int j = nbr[t] - 1;
        const size_t rj = (size_t)j * ND;
        double xi[3], eta[3];
        xi[0] = coords[3 * (size_t)j] - xi0;
        xi[1] = coords[3 * (size_t)j + 1] - yi0;
        xi[2] = (DIM >= 3) ? coords[3 * (size_t)j + 2] - zi0 : 0.0;
        eta[0] = uold[rj + 2] - uold[ri + 2];
        eta[1] = uold[rj + 3] - uold[ri + 3];
        eta[2] = (DIM >= 3) ? uold[rj + 4] - uold[ri + 4] : 0.0;
        const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        if (r2 <= 0.0) continue;
        const double r = sqrt(r2);
        const double rinv = 1.0 / r;
        const double yx = xi[0] + eta[0], yy = xi[1] + eta[1], yz = xi[2] + eta[2];
        const double sn = (sqrt(yx * yx + yy * yy + yz * yz) - r) * rinv;
        const double cbar = 0.5 * (ci + uold[rj]);
        const double s_el = sn - (Omega / 3.0) * (cbar - cref);
        const double so = tension_only ? fmax(s_el, 0.0) : fabs(s_el);
        if (so >= sc) health[t] = 0;
    }
}

struct FschConst {
    double C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref, rho, resid;
};

// everything an assembly launch reads and writes
struct SchArgs {
    int n, nop;
    const int *off, *nbr;
    const double *ops;
    const int *rev;
    const double *lap;               // lambda-corrected (stress CH) or species Laplacian (fracture)
    const unsigned char *isghost;
    const double *vol;
    FschConst k;
    double dt;
    const double *vel;               // fracture / dynamic: committed velocity
    const unsigned char *health;     // fracture / dynamic: bond health
    const int *rowptr, *slot, *diag, *mslot;
    const double *u, *uold;
    double *kvals, *rhs;
    double b[3];                     // dynamic fracture: body force
    const double *mnode;             // Picard-frozen nodal mobility from the host (null: from u)
};

// Element kinds sharing the block assembly: the stress-coupled
// Cahn-Hilliard element, its fracture variant, and the plain Cahn-Hilliard
// element (the species/chemical-potential part alone, dofs c and mu).
enum { EL_SCH = 0, EL_FSCH = 1, EL_CH = 2, EL_DYN = 3 };

template <int DIM, int EL>
__device__ __forceinline__ constexpr int el_nd() { return EL == EL_CH ? 2 : EL == EL_DYN ? DIM : DIM + 2; }

// The structural nonzeros of a bond's dR/du_j block P, in staging order:
// (0,1), then for rows a = 1..ND-1 the columns 0, 2, 3, ..., ND-1.  They are
// the bond terms of the diagonal block; its other entries hold only the
// nodal term.
template <int ND>
__device__ __forceinline__ constexpr int sch_nk() { return 1 + (ND - 1) * (ND - 1); }

template <int ND>
__device__ __forceinline__ void sch_entry(int e, int &a, int &b)
{
    if (e == 0) { a = 0; b = 1; return; }
    const int q = (e - 1) % (ND - 1);
    a = 1 + (e - 1) / (ND - 1);
    b = (q == 0) ? 0 : q + 1;
}

template <int ND>
__device__ __forceinline__ bool sch_staged(int a, int b) { return (a == 0 && b == 1) || (a >= 1 && b != 1); }

template <int ND>
__device__ __forceinline__ int sch_index(int a, int b) { return (a == 0) ? 0 : 1 + (a - 1) * (ND - 1) + (b == 0 ? 0 : b - 1); }

// per element: the dynamic fracture element's bond blocks are full (the
// symmetric Navier block), so every diagonal-block entry is staged, row-major
template <int DIM, int EL>
__device__ __forceinline__ constexpr int el_nk()
{
    return EL == EL_DYN ? DIM * DIM : sch_nk<el_nd<DIM, EL>()>();
}

template <int DIM, int EL>
__device__ __forceinline__ void el_entry(int e, int &a, int &b)
{
    if constexpr (EL == EL_DYN) { a = e / DIM; b = e % DIM; }
    else sch_entry<el_nd<DIM, EL>()>(e, a, b);
}

template <int DIM, int EL>
__device__ __forceinline__ bool el_staged(int a, int b)
{
    if constexpr (EL == EL_DYN) return true;
    else return sch_staged<el_nd<DIM, EL>()>(a, b);
}

template <int DIM, int EL>
__device__ __forceinline__ int el_index(int a, int b)
{
    if constexpr (EL == EL_DYN) return a * DIM + b;
    else return sch_index<el_nd<DIM, EL>()>(a, b);
}

// Node-level quantities of node i (the row owner) used by its bonds.
template <int ND>
struct SchNode {
    double ui[ND];
    double Mi, Vi;
};

// One bond t of node i: P and the residual terms R
// (StressCahnHilliardElement.cpp:282-428, FracStressCahnHilliardElement.cpp:
// 400-526, CahnHilliardElement.cpp:178-254).  The fracture variant gates the
// mechanics and sigma_h coupling with g (1 or the residual stiffness); with
// g = 1 every product rounds as the reference's ungated forms.
template <int DIM, int EL>
__device__ __forceinline__ void sch_bond(const SchArgs &A, const SchNode<el_nd<DIM, EL>()> &nd, int t,
                                         double (&P)[el_nd<DIM, EL>()][el_nd<DIM, EL>()],
                                         double (&R)[el_nd<DIM, EL>()])
{
    constexpr int ND = el_nd<DIM, EL>();
    const FschConst &k = A.k;
    const int j = A.nbr[t] - 1;
    const size_t rj = (size_t)j * ND;
    if constexpr (EL == EL_DYN) {
        // PDDODynamicFracElement.cpp:207-274: the health-gated Navier block;
        // a broken bond's blocks are zero
#pragma unroll
        for (int a = 0; a < ND; a++) {
            R[a] = 0.0;
#pragma unroll
            for (int b = 0; b < ND; b++) P[a][b] = 0.0;
        }
        if (!A.health[t]) return;
        const double *op = A.ops + (size_t)t * A.nop;
        const double mu = 1.0;
        if constexpr (DIM == 3) {
            const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
            const double aa = k.C12 + k.C66;
            const double p11 = mu * (k.C11 * Gxx + k.C66 * (Gyy + Gzz));
            const double p22 = mu * (k.C11 * Gyy + k.C66 * (Gxx + Gzz));
            const double p33 = mu * (k.C11 * Gzz + k.C66 * (Gxx + Gyy));
            const double p12 = mu * aa * Gxy, p13 = mu * aa * Gxz, p23 = mu * aa * Gyz;
            const double dux = A.u[rj] - nd.ui[0], duy = A.u[rj + 1] - nd.ui[1], duz = A.u[rj + 2] - nd.ui[2];
            R[0] = p11 * dux + p12 * duy + p13 * duz;
            R[1] = p12 * dux + p22 * duy + p23 * duz;
            R[2] = p13 * dux + p23 * duy + p33 * duz;
            P[0][0] = p11; P[0][1] = p12; P[0][2] = p13;
            P[1][0] = p12; P[1][1] = p22; P[1][2] = p23;
            P[2][0] = p13; P[2][1] = p23; P[2][2] = p33;
        } else {
            const double Gxx = op[2], Gxy = op[3], Gyy = op[4];
            const double p11 = mu * (k.C11 * Gxx + k.C66 * Gyy);
            const double p22 = mu * (k.C66 * Gxx + k.C11 * Gyy);
            const double pxy = mu * (k.C12 + k.C66) * Gxy;
            const double dux = A.u[rj] - nd.ui[0], duy = A.u[rj + 1] - nd.ui[1];
            R[0] = p11 * dux + pxy * duy;
            R[1] = pxy * dux + p22 * duy;
            P[0][0] = p11; P[0][1] = pxy; P[1][0] = pxy; P[1][1] = p22;
        }
        return;
    } else {
    // Every load of the bond is issued before the first use: the reverse
    // bond, the bond's Laplacian and health hang off t alone, the
    // neighbour's flag, volume, mobility, state and the reverse bond's
    // Laplacian off j or rv.  Loaded only where a branch needs them, they
    // formed a chain of five dependent rounds (off, nbr, isghost, rev, lap).
    const int rv = A.rev[t];
    const double lapt = A.lap[t];
    const double lapr = (rv >= 0) ? A.lap[rv] : 0.0;
    const bool ghostJ = A.isghost[j] != 0;
    const double volj = A.vol[j];
    double uj[ND];
#pragma unroll
    for (int a = 0; a < ND; a++) uj[a] = A.u[rj + a];
    const double mj = A.mnode ? A.mnode[j] : sch_mobility(uj[0], k.D);
    double g = 1.0, coeff_c = 0.0, kl = 0.0;
    if constexpr (EL == EL_FSCH) {
        g = A.health[t] ? 1.0 : k.resid;
        double wf = 0.0, g2c = 0.0;
        if (!ghostJ) {
            g2c = lapt;
            if (rv >= 0) wf = 0.5 * (nd.Mi * g2c + (volj / nd.Vi) * mj * lapr);
            else wf = nd.Mi * g2c;
        }
        coeff_c = -wf;
        kl = k.kappa * g2c;
    } else {
        const double g2ij = lapt;
        if (!ghostJ) {
            double wij;
            if (rv >= 0 && nd.Vi > 0.0)
                wij = 0.5 * (nd.Mi * g2ij + (volj / nd.Vi) * mj * lapr);
            else wij = nd.Mi * g2ij;
            coeff_c = -wij;
            kl = k.kappa * g2ij;
        }
    }
    double d[ND];
#pragma unroll
    for (int a = 0; a < ND; a++) d[a] = uj[a] - nd.ui[a];
#pragma unroll
    for (int a = 0; a < ND; a++)
#pragma unroll
        for (int b = 0; b < ND; b++) P[a][b] = 0.0;
    if constexpr (EL == EL_CH) {
        R[0] = coeff_c * d[1];
        R[1] = kl * d[0];
        P[0][1] = coeff_c;
        P[1][0] = kl;
    } else {
        const double *op = A.ops + (size_t)t * A.nop;
        const double go = g * k.Omega, gA = -g * k.A;
        if constexpr (DIM == 3) {
            const double Gx = op[0], Gy = op[1], Gz = op[2];
            const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
            const double ga = g * (k.C12 + k.C66);
            const double cux = go * k.Kh * Gx, cuy = go * k.Kh * Gy, cuz = go * k.Kh * Gz;
            const double p11 = g * (k.C11 * Gxx + k.C66 * (Gyy + Gzz));
            const double p22 = g * (k.C11 * Gyy + k.C66 * (Gxx + Gzz));
            const double p33 = g * (k.C11 * Gzz + k.C66 * (Gxx + Gyy));
            const double p12 = ga * Gxy, p13 = ga * Gxz, p23 = ga * Gyz;
            const double cx = gA * Gx, cy = gA * Gy, cz = gA * Gz;
            R[0] = coeff_c * d[1];
            R[1] = kl * d[0] + cux * d[2] + cuy * d[3] + cuz * d[4];
            R[2] = cx * d[0] + p11 * d[2] + p12 * d[3] + p13 * d[4];
            R[3] = cy * d[0] + p12 * d[2] + p22 * d[3] + p23 * d[4];
            R[4] = cz * d[0] + p13 * d[2] + p23 * d[3] + p33 * d[4];
            P[0][1] = coeff_c;
            P[1][0] = kl; P[1][2] = cux; P[1][3] = cuy; P[1][4] = cuz;
            P[2][0] = cx; P[2][2] = p11; P[2][3] = p12; P[2][4] = p13;
            P[3][0] = cy; P[3][2] = p12; P[3][3] = p22; P[3][4] = p23;
            P[4][0] = cz; P[4][2] = p13; P[4][3] = p23; P[4][4] = p33;
        } else {
            const double Gx = op[0], Gy = op[1], Gxx = op[2], Gxy = op[3], Gyy = op[4];
            const double cux = go * k.Kh * Gx, cuy = go * k.Kh * Gy;
            const double p11 = g * (k.C11 * Gxx + k.C66 * Gyy);
            const double p22 = g * (k.C66 * Gxx + k.C11 * Gyy);
            const double pxy = g * (k.C12 + k.C66) * Gxy;
            const double cx = gA * Gx, cy = gA * Gy;
            R[0] = coeff_c * d[1];
            R[1] = kl * d[0] + cux * d[2] + cuy * d[3];
            R[2] = cx * d[0] + p11 * d[2] + pxy * d[3];
            R[3] = cy * d[0] + pxy * d[2] + p22 * d[3];
            P[0][1] = coeff_c;
            P[1][0] = kl; P[1][2] = cux; P[1][3] = cuy;
            P[2][0] = cx; P[2][2] = p11; P[2][3] = pxy;
            P[3][0] = cy; P[3][2] = pxy; P[3][3] = p22;
        }
    }
    }
}

// The nodal term (StressCahnHilliardElement.cpp:449-466,
// FracStressCahnHilliardElement.cpp:528-561, CahnHilliardElement.cpp:256-292)
// of sum e: staged entry e < NK of the diagonal block, residual component
// e - NK otherwise, as the assembler leaves it before the bonds:
// 0 + (-K_node) and 0 + R_node.
template <int DIM, int EL>
__device__ __forceinline__ double sch_nodal(const SchArgs &A, const SchNode<el_nd<DIM, EL>()> &nd, size_t ri,
                                            int ic, int e)
{
    constexpr int ND = el_nd<DIM, EL>();
    constexpr int NK = el_nk<DIM, EL>();
    const FschConst &k = A.k;
    const double dt = A.dt;
    if constexpr (EL == EL_DYN) {
        // R = b - c0 (u - uold - v dt), K = -c0 I (PDDODynamicFracElement.cpp:276-297)
        const double c0 = k.rho / (dt * dt);
        if (e < NK) {
            const int a = e / DIM, b = e % DIM;
            return 0.0 + (-((a == b) ? -c0 : 0.0));
        }
        const int a = e - NK;
        double ua = 0.0, bf = 0.0;        // nd.ui[a], b[a] without a dynamic index
#pragma unroll
        for (int q = 0; q < ND; q++)
            if (q == a) { ua = nd.ui[q]; bf = A.b[q]; }
        return 0.0 + (bf - c0 * (ua - A.uold[ri + a] - A.vel[ri + a] * dt));
    } else {
    const double c = nd.ui[0], mu = nd.ui[1], cold = A.uold[ri];
    double fp, fpp;
    if constexpr (EL == EL_SCH) {
        double L, Lp;
        sch_reglog(c, L, Lp);
        fp = L + k.chi * (1.0 - 2.0 * c);
        fpp = Lp - 2.0 * k.chi;
    } else {
        const double cc = fmax(1.0e-8, fmin(1.0 - 1.0e-8, c));
        fp = log(cc / (1.0 - cc)) + k.chi * (1.0 - 2.0 * cc);
        fpp = 1.0 / (cc * (1.0 - cc)) - 2.0 * k.chi;
    }
    const bool inertia = EL == EL_FSCH && k.rho > 0.0;
    const double c0 = inertia ? k.rho / (dt * dt) : 0.0;
    if (e < NK) {
        int a, b;
        el_entry<DIM, EL>(e, a, b);
        double kn = 0.0;
        if (a == 1 && b == 0) kn = (EL == EL_CH) ? -fpp : -fpp + k.Omega * k.Ch;
        else if (inertia && a >= 2 && a == b) kn = 0.0 + -c0;
        return 0.0 + (-kn);
    }
    const int a = e - NK;
    if (a == 0) return 0.0 + (c - cold) / dt;
    if (a == 1) {
        if constexpr (EL == EL_CH) return 0.0 + (mu - fp);
        else return 0.0 + (mu - fp + k.Omega * k.Ch * (c - k.cref));
    }
    if (inertia) {
        const double v = A.vel[(size_t)ic * DIM + a - 2];
        double ua = 0.0;                  // nd.ui[a] without a dynamic index
#pragma unroll
        for (int q = 2; q < ND; q++)
            if (q == a) ua = nd.ui[q];
        return 0.0 + (0.0 + -c0 * (ua - A.uold[ri + a] - v * dt));
    }
    return 0.0;
    }
}

// the diagonal-block entries without bond terms: 0 + (-K_node)
template <int ND>
__device__ __forceinline__ double sch_nodal_only(int a, int b, double dt)
{
    const double kn = (a == 0 && b == 0) ? 1.0 / dt : (a == 1 && b == 1) ? 1.0 : 0.0;
    return 0.0 + (-kn);
}

// ---- assembly: one node per LPN lanes of a wavefront, lanes over its bonds ----
//
// k_sch_rows (rows of at most MAXB node blocks): lane l of the node computes
// bonds l, l + LPN, ... and stages P's structural nonzeros and the residual
// terms in LDS, with a map from the row's block positions to bonds.  One
// lane per diagonal-block entry / residual component then adds the staged
// terms in family order onto the nodal term -- the reference's order, so
// every value comes out bit for bit -- and finally the node's lanes write its
// CSR rows front to back, LPN consecutive values per store, the K_IJ blocks
// as 0 + (-P) (FormResidualAndJacobian.cpp:119-128).  2D families (28 bonds
// at the usual horizon) take 32 lanes, two nodes per wavefront, and the
// Cahn-Hilliard element's four sums 16 lanes, four nodes per wavefront.
//
// k_sch (any row length): the same sums, in chunks of 64 bonds, with each
// lane writing its bond's K_IJ block straight into the rows.
constexpr int kSchWaves = 4;   // nodes per block of k_sch

template <int DIM, int EL, int MAXB, int NW, int LPN>
__global__ void __launch_bounds__(64 * NW) k_sch_rows(SchArgs A)
{
    constexpr int ND = el_nd<DIM, EL>();
    constexpr int NK = el_nk<DIM, EL>();
    constexpr int NS = NK + ND;
    constexpr int NPW = 64 / LPN;            // nodes per wavefront
    constexpr int NB = NW * NPW;             // nodes per block
    static_assert(NS + ND * ND <= LPN, "one lane per sum and per nodal-only entry");
    __shared__ double s_v[NB][NS][MAXB + 1];
    __shared__ double s_d[NB][ND * ND];
    __shared__ int s_inv[NB][MAXB];
    const int l = threadIdx.x % LPN, w = threadIdx.x / LPN;   // w: node slot in the block
    const int i = blockIdx.x * NB + w;
    const bool active = i < A.n;
    const int ic = active ? i : 0;   // reads of an idle slot stay in bounds
    const size_t ri = (size_t)ic * ND;
    const int r0 = A.rowptr[ri];
    const int rowlen = A.rowptr[ri + 1] - r0;
    int rowoff[ND];
#pragma unroll
    for (int a = 0; a < ND; a++) rowoff[a] = A.rowptr[ri + a] - r0;
    SchNode<ND> nd;
#pragma unroll
    for (int a = 0; a < ND; a++) nd.ui[a] = A.u[ri + a];
    nd.Mi = A.mnode ? A.mnode[ic] : sch_mobility(nd.ui[0], A.k.D);
    nd.Vi = A.vol[ic];
    const int t0 = active ? A.off[i] : 0, deg = active ? A.off[i + 1] - t0 : 0;
    for (int tl = l; tl < deg; tl += LPN) {
        const int t = t0 + tl;
        double P[ND][ND], R[ND];
        sch_bond<DIM, EL>(A, nd, t, P, R);
#pragma unroll
        for (int e = 0; e < NK; e++) {
            int a, b;
            el_entry<DIM, EL>(e, a, b);
            s_v[w][e][tl] = P[a][b];
        }
#pragma unroll
        for (int a = 0; a < ND; a++) s_v[w][NK + a][tl] = R[a];
        s_inv[w][(A.slot[t] - r0) / ND] = tl;
    }
    if (active && l == 0) {
        s_inv[w][(A.diag[i] - r0) / ND] = -1;
        if (A.mslot[i] >= 0) s_inv[w][(A.mslot[i] - r0) / ND] = -2;
    }
    __syncthreads();
    if (l < NS) {
        double acc = sch_nodal<DIM, EL>(A, nd, ri, ic, l);
#pragma unroll 8
        for (int q = 0; q < deg; q++) acc += s_v[w][l][q];   // in order; the loads pipeline
        if (l < NK) {
            int a, b;
            el_entry<DIM, EL>(l, a, b);
            s_d[w][a * ND + b] = acc;
        } else if (active) {
            A.rhs[ri + l - NK] = acc;
        }
    } else if (l < NS + ND * ND) {
        const int q = l - NS, a = q / ND, b = q % ND;
        if (!el_staged<DIM, EL>(a, b)) s_d[w][q] = sch_nodal_only<ND>(a, b, A.dt);
    }
    __syncthreads();
    if (!active) return;
#pragma unroll
    for (int a = 0; a < ND; a++) {
        double *row = A.kvals + r0 + rowoff[a];
        for (int q = l; q < rowlen; q += LPN) {
            const int kb = q / ND, b = q - kb * ND;
            const int src = s_inv[w][kb];
            double v = 0.0;                               // mirror block, structural zeros
            if (src >= 0) {
                if (el_staged<DIM, EL>(a, b)) v = 0.0 + (-s_v[w][el_index<DIM, EL>(a, b)][src]);
            } else if (src == -1) {
                v = s_d[w][a * ND + b];
            }
            row[q] = v;
        }
    }
}

template <int DIM, int EL>
__global__ void __launch_bounds__(64 * kSchWaves) k_sch(SchArgs A)
{
    constexpr int ND = el_nd<DIM, EL>();
    constexpr int NK = el_nk<DIM, EL>();
    constexpr int NS = NK + ND;
    __shared__ double s_v[kSchWaves][NS][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * kSchWaves;
    const int i = i0 + w;
    const bool active = i < A.n;
    const int ic = active ? i : 0;   // reads of an idle wave stay in bounds
    // chunks of 64 bonds: the block runs the largest count of its nodes
    int nch = 0;
#pragma unroll
    for (int q = 0; q < kSchWaves; q++)
        if (i0 + q < A.n) nch = max(nch, (A.off[i0 + q + 1] - A.off[i0 + q] + 63) >> 6);
    const size_t ri = (size_t)ic * ND;
    const int rb = A.rowptr[ri];
    int rowoff[ND];
#pragma unroll
    for (int a = 0; a < ND; a++) rowoff[a] = A.rowptr[ri + a] - rb;
    SchNode<ND> nd;
#pragma unroll
    for (int a = 0; a < ND; a++) nd.ui[a] = A.u[ri + a];
    nd.Mi = A.mnode ? A.mnode[ic] : sch_mobility(nd.ui[0], A.k.D);
    nd.Vi = A.vol[ic];
    const int t0 = active ? A.off[i] : 0, t1 = active ? A.off[i + 1] : 0;
    if (active && lane < ND * ND && A.mslot[i] >= 0) {
        // the ghost's mirror block is in the pattern but receives no term
        A.kvals[A.mslot[i] + rowoff[lane / ND] + lane % ND] = 0.0;
    }
    double acc = (lane < NS) ? sch_nodal<DIM, EL>(A, nd, ri, ic, lane) : 0.0;
    for (int ch = 0; ch < nch; ch++) {
        const int t = t0 + ch * 64 + lane;
        if (t < t1) {
            double P[ND][ND], R[ND];
            sch_bond<DIM, EL>(A, nd, t, P, R);
            const int s0 = A.slot[t];
#pragma unroll
            for (int a = 0; a < ND; a++)
#pragma unroll
                for (int b = 0; b < ND; b++) A.kvals[s0 + rowoff[a] + b] = 0.0 + (-P[a][b]);
#pragma unroll
            for (int e = 0; e < NK; e++) {
                int a, b;
                el_entry<DIM, EL>(e, a, b);
                s_v[w][e][lane] = P[a][b];
            }
#pragma unroll
            for (int a = 0; a < ND; a++) s_v[w][NK + a][lane] = R[a];
        }
        __syncthreads();
        if (lane < NS) {
            const int m = min(64, t1 - (t0 + ch * 64));
#pragma unroll 8
            for (int q = 0; q < m; q++) acc += s_v[w][lane][q];   // in order; the loads pipeline
        }
        __syncthreads();
    }
    if (!active) return;
    const int d0 = A.diag[i];
    if (lane < NK) {
        int a, b;
        el_entry<DIM, EL>(lane, a, b);
        A.kvals[d0 + rowoff[a] + b] = acc;
    } else if (lane < NS) {
        A.rhs[ri + lane - NK] = acc;
    }
    if (lane < ND * ND) {
        const int a = lane / ND, b = lane % ND;
        if (!el_staged<DIM, EL>(a, b)) A.kvals[d0 + rowoff[a] + b] = sch_nodal_only<ND>(a, b, A.dt);
    }
}

template <int DIM, int EL>
int launch_sch(hipStream_t st, int maxblk, const SchArgs &A)
{
    const int n = A.n;
    if constexpr (EL == EL_CH || (EL == EL_DYN && DIM == 2)) {
        // four / six sums per node: 16 lanes, four nodes per wavefront
        if (maxblk <= 32) {
            hipLaunchKernelGGL((k_sch_rows<DIM, EL, 32, 4, 16>), dim3(grids(n, 16)), dim3(256), 0, st, A);
            PXG_HIP(hipGetLastError());
            return PXG_OK;
        }
    } else if constexpr (DIM == 2) {
        if (maxblk <= 32) {
            hipLaunchKernelGGL((k_sch_rows<DIM, EL, 32, 4, 32>), dim3(grids(n, 8)), dim3(256), 0, st, A);
            PXG_HIP(hipGetLastError());
            return PXG_OK;
        }
    }
    if (maxblk <= 64)
        hipLaunchKernelGGL((k_sch_rows<DIM, EL, 64, 4, 64>), dim3(grids(n, 4)), dim3(256), 0, st, A);
    else if (maxblk <= 128)
        hipLaunchKernelGGL((k_sch_rows<DIM, EL, 128, 2, 64>), dim3(grids(n, 2)), dim3(128), 0, st, A);
    else
        hipLaunchKernelGGL((k_sch<DIM, EL>), dim3(grids(n, kSchWaves)), dim3(64 * kSchWaves), 0, st, A);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

// Output projections (StressCahnHilliardElement.cpp:481-658,
// FracStressCahnHilliardElement.cpp:47-110, 575-672), one thread per node,
// family sums in family order: which = PXG_SCH_* (0 strain, 1 stress,
// 2 vonMises, 3 sigmaH, 4 fPrime, 5 fDoublePrime, 6 gradC, 7 damage,
// 8 DamageI); 9, 3 or 1 values per node.
template <int DIM>
__global__ void k_sch_project(int n, int nop, const int *off, const int *nbr, const double *ops, double C11,
                              double C12, double C66, double A, double cref, double chi, int frac,
                              int plane_strain, const unsigned char *health, const unsigned char *seed,
                              const double *u, int which, double *out)
{
    constexpr int ND = DIM + 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nc = (which <= 1) ? 9 : (which == 6) ? 3 : 1;
    double *o = out + (size_t)nc * i;
    const size_t ri = (size_t)i * ND;
    const double c = u[ri];
    const int t0 = off[i], t1 = off[i + 1], deg = t1 - t0;
    double v[9];
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] = 0.0;
    if (which == 4 || which == 5) {
        if (frac) {
            const double cc = fmax(1.0e-8, fmin(1.0 - 1.0e-8, c));
            v[0] = (which == 4) ? log(cc / (1.0 - cc)) + chi * (1.0 - 2.0 * cc)
                                : 1.0 / (cc * (1.0 - cc)) - 2.0 * chi;
        } else {
            double L, Lp;
            sch_reglog(c, L, Lp);
            v[0] = (which == 4) ? L + chi * (1.0 - 2.0 * c) : Lp - 2.0 * chi;
        }
    } else if (which == 7 || which == 8) {
        // damage: broken share of the family; DamageI: share broken by the
        // per-step commits (intact in the crack seed)
        int cnt = 0;
        for (int t = t0; t < t1; t++)
            if (!health[t] && (which == 7 || seed[t])) cnt++;
        double phi = deg > 0 ? (double)cnt / (double)deg : 0.0;
        if (which == 8) phi = fmin(fmax(phi, 0.0), 1.0);
        v[0] = phi;
    } else if (which == 6) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int t = t0; t < t1; t++) {
            const double *op = ops + (size_t)t * nop;
            const double dc = u[(size_t)(nbr[t] - 1) * ND] - c;
            gx += dc * op[0];
            gy += dc * op[1];
            if (DIM == 3) gz += dc * op[2];
        }
        v[0] = gx; v[1] = gy; v[2] = gz;
    } else {
        const double dc = c - cref;
        if constexpr (DIM == 3) {
            double g[3][3];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) g[a][b] = 0.0;
            for (int t = t0; t < t1; t++) {
                const double *op = ops + (size_t)t * nop;
                const size_t rj = (size_t)(nbr[t] - 1) * ND;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const double du = u[rj + 2 + a] - u[ri + 2 + a];
#pragma unroll
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
                v[0] = exx; v[1] = exy; v[2] = exz; v[3] = exy; v[4] = eyy; v[5] = eyz;
                v[6] = exz; v[7] = eyz; v[8] = ezz;
            } else if (which == 1) {
                v[0] = sxx; v[1] = sxy; v[2] = sxz; v[3] = sxy; v[4] = syy; v[5] = syz;
                v[6] = sxz; v[7] = syz; v[8] = szz;
            } else if (which == 3) {
                v[0] = (sxx + syy + szz) / 3.0;
            } else {
                const double a = sxx - syy, b = syy - szz, cc = szz - sxx;
                v[0] = sqrt(0.5 * (a * a + b * b + cc * cc + 6.0 * (sxy * sxy + sxz * sxz + syz * syz)));
            }
        } else {
            double dxx = 0.0, dxy = 0.0, dyx = 0.0, dyy = 0.0;
            for (int t = t0; t < t1; t++) {
                const double *op = ops + (size_t)t * nop;
                const size_t rj = (size_t)(nbr[t] - 1) * ND;
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
                v[0] = exx; v[1] = exy; v[3] = exy; v[4] = eyy;
            } else if (which == 1) {
                v[0] = sxx; v[1] = sxy; v[3] = sxy; v[4] = syy; v[8] = szz;
            } else if (which == 3) {
                v[0] = (sxx + syy + szz) / 3.0;
            } else {
                const double d1 = sxx - syy, d2 = syy - szz, d3 = szz - sxx;
                v[0] = sqrt(0.5 * (d1 * d1 + d2 * d2 + d3 * d3 + 6.0 * sxy * sxy));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 9; q++)
        if (q < nc) o[q] = v[q];
}

} // namespace

// recalcElasticConstants (StressCahnHilliardElement.cpp:80-128)
static int sch_constants(int dim, int plane_stress, double E, double nu, double Omega, double out[6])
{
    for (int q = 0; q < 6; q++) out[q] = 0.0;
    if (E <= 0.0) return PXG_OK;
    if (dim < 3 && plane_stress) {
        const double d = 1.0 - nu * nu;
        PXG_CHECK(d > 0.0, PXG_ERR_ARG, "pxg_stressch_setup: plane_stress needs nu^2 < 1");
        out[0] = E / d;
        out[1] = E * nu / d;
        out[2] = E / (2.0 * (1.0 + nu));
        out[3] = E * Omega / (3.0 * (1.0 - nu));
        out[4] = (out[0] + out[1]) / 3.0;
        out[5] = -2.0 * out[3] / 3.0;
        return PXG_OK;
    }
    const double d = (1.0 + nu) * (1.0 - 2.0 * nu);
    PXG_CHECK(d > 0.0, PXG_ERR_ARG, "pxg_stressch_setup: needs (1+nu)(1-2nu) > 0");
    out[0] = E * (1.0 - nu) / d;
    out[1] = E * nu / d;
    out[2] = E / (2.0 * (1.0 + nu));
    out[3] = E * Omega / (3.0 * (1.0 - 2.0 * nu));
    out[4] = (out[0] + 2.0 * out[1]) / 3.0;
    out[5] = -out[3];
    return PXG_OK;
}

extern "C" int pxg_stressch_setup(pxg_ctx *ctx, const pxg_stressch_params *p)
{
    PXG_CHECK(ctx && p, PXG_ERR_ARG, "pxg_stressch_setup: bad arguments");
    PXG_CHECK(ctx->nop > 0 && ctx->order >= 2, PXG_ERR_STATE,
              "pxg_stressch_setup: needs order-2 operators (StressCahnHilliardElement.cpp:158-161)");
    PXG_CHECK(ctx->nnz > 0 && ctx->csr_ndof == ctx->dim + 2, PXG_ERR_STATE,
              "pxg_stressch_setup: needs a CSR pattern with dim+2 dofs per node (c, mu, u)");
    double e[6];
    const int rc = sch_constants(ctx->dim, p->plane_stress, p->E, p->nu, p->Omega, e);
    if (rc) return rc;
    const int g = ch_geometry(ctx);
    if (g) return g;
    StressChState &S = ctx->sch;
    const double v[11] = {e[0], e[1], e[2], e[3], e[4], e[5], p->D, p->kappa, p->chi, p->Omega, p->cref};
    for (int q = 0; q < 11; q++) S.k[q] = v[q];
    S.plane_strain = (ctx->dim >= 3 || !p->plane_stress) ? 1 : 0;
    S.ready = 1;
    return PXG_OK;
}

extern "C" int pxg_stressch_constants(pxg_ctx *ctx, double *out)
{
    PXG_CHECK(ctx && ctx->sch.ready && out, PXG_ERR_STATE, "pxg_stressch_constants: call pxg_stressch_setup first");
    for (int q = 0; q < 11; q++) out[q] = ctx->sch.k[q];
    return PXG_OK;
}

extern "C" int pxg_assemble_stressch_dev(pxg_ctx *ctx, double dt, const double *u_dev, const double *uold_dev,
                                         double *kvals_dev, double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->sch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_assemble_stressch: call pxg_stressch_setup first");
    PXG_CHECK(ctx->csr_ndof == ctx->dim + 2, PXG_ERR_STATE, "pxg_assemble_stressch: CSR pattern changed");
    PXG_CHECK(u_dev && uold_dev && kvals_dev && rhs_dev, PXG_ERR_ARG, "pxg_assemble_stressch: bad arguments");
    PXG_CHECK(dt > 0.0, PXG_ERR_ARG,
              "pxg_assemble_stressch: the element is transient-only (StressCahnHilliardElement.cpp:443-447)");
    PXG_FAULT(ctx, "pxg_assemble_stressch");
    PXG_HIP(hipSetDevice(ctx->device));
    const double *q = ctx->sch.k;
    const SchArgs A{ctx->n, ctx->nop, ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->rev.p, ctx->lap.p, ctx->isghost.p,
                    ctx->vols.p,
                    FschConst{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], 0.0, 1.0}, dt,
                    nullptr, nullptr, ctx->rowptr.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, u_dev, uold_dev,
                    kvals_dev, rhs_dev, {0.0, 0.0, 0.0}};
    return ctx->dim == 3 ? launch_sch<3, EL_SCH>(ctx->stream, ctx->csr_maxblk, A)
                         : launch_sch<2, EL_SCH>(ctx->stream, ctx->csr_maxblk, A);
}

extern "C" int pxg_assemble_stressch(pxg_ctx *ctx, double dt, const double *u, const double *uold, double *kvals,
                                     double *rhs)
{
    PXG_CHECK(ctx && ctx->sch.ready, PXG_ERR_STATE, "pxg_assemble_stressch: call pxg_stressch_setup first");
    PXG_CHECK(u && uold && kvals && rhs, PXG_ERR_ARG, "pxg_assemble_stressch: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * (ctx->dim + 2);
    PXG_HIP(ctx->su.alloc(nd));
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(ctx->sk.alloc(ctx->nnz));
    PXG_HIP(ctx->sr.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    const int rc = pxg_assemble_stressch_dev(ctx, dt, ctx->su.p, ctx->suo.p, ctx->sk.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

// ---- fracture variant entry points -----------------------------------------

extern "C" int pxg_fsch_setup(pxg_ctx *ctx, const pxg_fsch_params *p, const double *uold0)
{
    PXG_CHECK(ctx && p, PXG_ERR_ARG, "pxg_fsch_setup: bad arguments");
    PXG_CHECK(ctx->nop > 0 && ctx->order >= 2, PXG_ERR_STATE,
              "pxg_fsch_setup: needs order-2 operators (FracStressCahnHilliardElement.cpp:180-183)");
    PXG_CHECK(ctx->nnz > 0 && ctx->csr_ndof == ctx->dim + 2, PXG_ERR_STATE,
              "pxg_fsch_setup: needs a CSR pattern with dim+2 dofs per node (c, mu, u)");
    PXG_CHECK(!(p->rho > 0.0) || uold0, PXG_ERR_ARG, "pxg_fsch_setup: rho > 0 needs the initial u_old");
    double e[6];
    const int rc = sch_constants(ctx->dim, p->base.plane_stress, p->base.E, p->base.nu, p->base.Omega, e);
    if (rc) return rc;
    const int g = ch_geometry(ctx);   // reverse bonds and ghost flags
    if (g) return g;
    FracSchState &F = ctx->fsch;
    hipStream_t st = ctx->stream;
    const double v[13] = {e[0], e[1], e[2], e[3], e[4], e[5], p->base.D, p->base.kappa, p->base.chi,
                          p->base.Omega, p->base.cref, p->rho, p->residual_stiffness};
    for (int q = 0; q < 13; q++) F.k[q] = v[q];
    F.sc = p->sc;
    F.plane_strain = (ctx->dim >= 3 || !p->base.plane_stress) ? 1 : 0;
    F.tension_only = p->tension_only;
    F.damage_on = p->damage_on;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    PXG_HIP(F.health.alloc((size_t)ctx->nb));
    PXG_HIP(F.lapc.alloc((size_t)ctx->nb));
    PXG_HIP(F.vel.alloc(nd));
    PXG_HIP(F.uprev.alloc(nd));
    const double *uo = nullptr;
    if (p->rho > 0.0) {
        const size_t nu = (size_t)ctx->n * (ctx->dim + 2);
        PXG_HIP(ctx->suo.alloc(nu));
        PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold0, sizeof(double) * nu, hipMemcpyHostToDevice, st));
        uo = ctx->suo.p;
    }
    hipLaunchKernelGGL(k_fsch_lapc, dim3(grids(ctx->nb, 256)), dim3(256), 0, st, ctx->nb, ctx->dim, ctx->nop,
                       ctx->nbr.p, ctx->ops.p, ctx->mirror.p, F.lapc.p);
    const long long m = std::max((long long)ctx->nb, (long long)nd);
    hipLaunchKernelGGL(k_fsch_init, dim3(grids(m, 256)), dim3(256), 0, st, ctx->nb, ctx->n, ctx->dim,
                       ctx->seed.p, uo, F.health.p, F.uprev.p, F.vel.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipStreamSynchronize(st));
    F.ready = 1;
    return PXG_OK;
}

extern "C" int pxg_fsch_constants(pxg_ctx *ctx, double *out)
{
    PXG_CHECK(ctx && ctx->fsch.ready && out, PXG_ERR_STATE, "pxg_fsch_constants: call pxg_fsch_setup first");
    for (int q = 0; q < 13; q++) out[q] = ctx->fsch.k[q];
    out[13] = ctx->fsch.sc;
    return PXG_OK;
}

extern "C" int pxg_fsch_commit_dev(pxg_ctx *ctx, const double *uold_dev, double dt_prev)
{
    PXG_CHECK(ctx && ctx->fsch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_fsch_commit: call pxg_fsch_setup first (again after a CSR rebuild or a Poisson or diffusion assembly)");
    PXG_CHECK(ctx->csr_ndof == ctx->dim + 2, PXG_ERR_STATE, "pxg_fsch_commit: CSR pattern changed");
    PXG_CHECK(uold_dev && dt_prev > 0.0, PXG_ERR_ARG, "pxg_fsch_commit: bad arguments");
    PXG_FAULT(ctx, "pxg_fsch_commit");
    PXG_HIP(hipSetDevice(ctx->device));
    FracSchState &F = ctx->fsch;
    const unsigned g = grids(ctx->n, 128);
    if (ctx->dim == 3)
        hipLaunchKernelGGL(k_fsch_commit<3>, dim3(g), dim3(128), 0, ctx->stream, ctx->n, ctx->coords.p,
                           ctx->off.p, ctx->nbr.p, F.sc, F.tension_only, F.damage_on, F.k[9], F.k[10], F.k[11],
                           dt_prev, uold_dev, F.uprev.p, F.vel.p, F.health.p);
    else
        hipLaunchKernelGGL(k_fsch_commit<2>, dim3(g), dim3(128), 0, ctx->stream, ctx->n, ctx->coords.p,
                           ctx->off.p, ctx->nbr.p, F.sc, F.tension_only, F.damage_on, F.k[9], F.k[10], F.k[11],
                           dt_prev, uold_dev, F.uprev.p, F.vel.p, F.health.p);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_fsch_commit(pxg_ctx *ctx, const double *uold, double dt_prev)
{
    PXG_CHECK(ctx && ctx->fsch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_fsch_commit: call pxg_fsch_setup first (again after a CSR rebuild or a Poisson or diffusion assembly)");
    PXG_CHECK(uold, PXG_ERR_ARG, "pxg_fsch_commit: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    const size_t nu = (size_t)ctx->n * (ctx->dim + 2);
    PXG_HIP(ctx->suo.alloc(nu));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nu, hipMemcpyHostToDevice, ctx->stream));
    const int rc = pxg_fsch_commit_dev(ctx, ctx->suo.p, dt_prev);
    if (rc) return rc;
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_assemble_fsch_dev(pxg_ctx *ctx, double dt, const double *u_dev, const double *uold_dev,
                                     double *kvals_dev, double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->fsch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_assemble_fsch: call pxg_fsch_setup first (again after a CSR rebuild or a Poisson or diffusion assembly)");
    PXG_CHECK(ctx->csr_ndof == ctx->dim + 2, PXG_ERR_STATE, "pxg_assemble_fsch: CSR pattern changed");
    PXG_CHECK(u_dev && uold_dev && kvals_dev && rhs_dev, PXG_ERR_ARG, "pxg_assemble_fsch: bad arguments");
    PXG_CHECK(dt > 0.0, PXG_ERR_ARG,
              "pxg_assemble_fsch: the element is transient-only (FracStressCahnHilliardElement.cpp:536-540)");
    PXG_FAULT(ctx, "pxg_assemble_fsch");
    PXG_HIP(hipSetDevice(ctx->device));
    const FracSchState &F = ctx->fsch;
    const double *q = F.k;
    const SchArgs A{ctx->n, ctx->nop, ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->rev.p, F.lapc.p, ctx->isghost.p,
                    ctx->vols.p,
                    FschConst{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11], q[12]}, dt,
                    F.vel.p, F.health.p, ctx->rowptr.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, u_dev, uold_dev,
                    kvals_dev, rhs_dev, {0.0, 0.0, 0.0}};
    return ctx->dim == 3 ? launch_sch<3, EL_FSCH>(ctx->stream, ctx->csr_maxblk, A)
                         : launch_sch<2, EL_FSCH>(ctx->stream, ctx->csr_maxblk, A);
}

extern "C" int pxg_assemble_fsch(pxg_ctx *ctx, double dt, const double *u, const double *uold, double *kvals,
                                 double *rhs)
{
    PXG_CHECK(ctx && ctx->fsch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_assemble_fsch: call pxg_fsch_setup first (again after a CSR rebuild or a Poisson or diffusion assembly)");
    PXG_CHECK(u && uold && kvals && rhs, PXG_ERR_ARG, "pxg_assemble_fsch: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * (ctx->dim + 2);
    PXG_HIP(ctx->su.alloc(nd));
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(ctx->sk.alloc(ctx->nnz));
    PXG_HIP(ctx->sr.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    const int rc = pxg_assemble_fsch_dev(ctx, dt, ctx->su.p, ctx->suo.p, ctx->sk.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_fsch_get_state(pxg_ctx *ctx, unsigned char *health, double *vel)
{
    PXG_CHECK(ctx && ctx->fsch.ready, PXG_ERR_STATE, "pxg_fsch_get_state: call pxg_fsch_setup first");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const FracSchState &F = ctx->fsch;
    if (health) PXG_HIP(hipMemcpyAsync(health, F.health.p, (size_t)ctx->nb, hipMemcpyDeviceToHost, st));
    if (vel)
        PXG_HIP(hipMemcpyAsync(vel, F.vel.p, sizeof(double) * ctx->n * ctx->dim, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

// ---- output projections ----------------------------------------------------

static int sch_project(pxg_ctx *ctx, bool frac, int which, const double *u_dev, double *out_dev)
{
    PXG_HIP(hipSetDevice(ctx->device));
    const double *k = frac ? ctx->fsch.k : ctx->sch.k;
    const int ps = frac ? ctx->fsch.plane_strain : ctx->sch.plane_strain;
    const unsigned char *h = frac ? ctx->fsch.health.p : nullptr;
    const unsigned g = grids(ctx->n, 128);
    if (ctx->dim == 3)
        hipLaunchKernelGGL(k_sch_project<3>, dim3(g), dim3(128), 0, ctx->stream, ctx->n, ctx->nop, ctx->off.p,
                           ctx->nbr.p, ctx->ops.p, k[0], k[1], k[2], k[3], k[10], k[8], (int)frac, ps, h,
                           ctx->seed.p, u_dev, which, out_dev);
    else
        hipLaunchKernelGGL(k_sch_project<2>, dim3(g), dim3(128), 0, ctx->stream, ctx->n, ctx->nop, ctx->off.p,
                           ctx->nbr.p, ctx->ops.p, k[0], k[1], k[2], k[3], k[10], k[8], (int)frac, ps, h,
                           ctx->seed.p, u_dev, which, out_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

static int sch_project_host(pxg_ctx *ctx, bool frac, int which, const double *u, double *out)
{
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nu = (size_t)ctx->n * (ctx->dim + 2);
    const size_t no = (size_t)ctx->n * pxg_sch_components(which);
    PXG_HIP(ctx->su.alloc(nu));
    PXG_HIP(ctx->sr.alloc(no));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nu, hipMemcpyHostToDevice, st));
    const int rc = sch_project(ctx, frac, which, ctx->su.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(out, ctx->sr.p, sizeof(double) * no, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_sch_components(int which)
{
    return (which == PXG_SCH_STRAIN || which == PXG_SCH_STRESS) ? 9 : (which == PXG_SCH_GRADC) ? 3 : 1;
}

// the names each element lists (getAvailableProjections)
static bool stressch_has(int which) { return which >= PXG_SCH_STRAIN && which <= PXG_SCH_GRADC; }
static bool fsch_has(int which)
{
    return (which >= PXG_SCH_STRAIN && which <= PXG_SCH_SIGMAH) || which == PXG_SCH_GRADC ||
           which == PXG_SCH_DAMAGE || which == PXG_SCH_DAMAGEI;
}

extern "C" int pxg_stressch_project_dev(pxg_ctx *ctx, int which, const double *u_dev, double *out_dev)
{
    PXG_CHECK(ctx && ctx->sch.ready, PXG_ERR_STATE, "pxg_stressch_project: call pxg_stressch_setup first");
    PXG_CHECK(stressch_has(which) && u_dev && out_dev, PXG_ERR_ARG,
              "pxg_stressch_project: bad arguments (which: strain .. gradC)");
    PXG_FAULT(ctx, "pxg_stressch_project");
    return sch_project(ctx, false, which, u_dev, out_dev);
}

extern "C" int pxg_stressch_project(pxg_ctx *ctx, int which, const double *u, double *out)
{
    PXG_CHECK(ctx && ctx->sch.ready, PXG_ERR_STATE, "pxg_stressch_project: call pxg_stressch_setup first");
    PXG_CHECK(stressch_has(which) && u && out, PXG_ERR_ARG,
              "pxg_stressch_project: bad arguments (which: strain .. gradC)");
    PXG_FAULT(ctx, "pxg_stressch_project");
    return sch_project_host(ctx, false, which, u, out);
}

extern "C" int pxg_fsch_project_dev(pxg_ctx *ctx, int which, const double *u_dev, double *out_dev)
{
    PXG_CHECK(ctx && ctx->fsch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_fsch_project: call pxg_fsch_setup first (again after a CSR rebuild or a Poisson or diffusion assembly)");
    PXG_CHECK(fsch_has(which) && u_dev && out_dev, PXG_ERR_ARG,
              "pxg_fsch_project: bad arguments (which: strain, stress, vonMises, sigmaH, gradC, damage, DamageI)");
    PXG_FAULT(ctx, "pxg_fsch_project");
    return sch_project(ctx, true, which, u_dev, out_dev);
}

extern "C" int pxg_fsch_project(pxg_ctx *ctx, int which, const double *u, double *out)
{
    PXG_CHECK(ctx && ctx->fsch.ready && ctx->ch_ready, PXG_ERR_STATE,
              "pxg_fsch_project: call pxg_fsch_setup first (again after a CSR rebuild or a Poisson or diffusion assembly)");
    PXG_CHECK(fsch_has(which) && u && out, PXG_ERR_ARG,
              "pxg_fsch_project: bad arguments (which: strain, stress, vonMises, sigmaH, gradC, damage, DamageI)");
    PXG_FAULT(ctx, "pxg_fsch_project");
    return sch_project_host(ctx, true, which, u, out);
}

// The Cahn-Hilliard element's assembly on the same kernels (dofs c, mu):
// M(c) = M0 c (1-c) clamped, the clamped logarithm, lambda-corrected
// ghost-excluded Laplacian (CahnHilliardElement.cpp:93-292).
int pxg::launch_ch_rows(pxg_ctx *ctx, double kappa, double chi, double M0, double dt, const double *u_dev,
                        const double *uold_dev, double *kvals_dev, double *rhs_dev)
{
    const SchArgs A{ctx->n, ctx->nop, ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->rev.p, ctx->lap.p, ctx->isghost.p,
                    ctx->vols.p,
                    FschConst{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, M0, kappa, chi, 0.0, 0.0, 0.0, 1.0}, dt, nullptr,
                    nullptr, ctx->rowptr.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, u_dev, uold_dev, kvals_dev,
                    rhs_dev, {0.0, 0.0, 0.0}, ctx->mob_set ? ctx->mob.p : nullptr};
    return ctx->dim == 3 ? launch_sch<3, EL_CH>(ctx->stream, ctx->csr_maxblk, A)
                         : launch_sch<2, EL_CH>(ctx->stream, ctx->csr_maxblk, A);
}

// The dynamic fracture element's assembly on the same kernels (dofs u).
int pxg::launch_dyn_rows(pxg_ctx *ctx, double dt, const double *u_dev, const double *uold_dev, double *kvals_dev,
                         double *rhs_dev)
{
    const DynFracState &D = ctx->dyn;
    const SchArgs A{ctx->n, ctx->nop, ctx->off.p, ctx->nbr.p, ctx->ops.p, nullptr, nullptr, nullptr, ctx->vols.p,
                    FschConst{D.C11, D.C12, D.C66, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, D.rho, 1.0}, dt,
                    D.vel.p, D.health.p, ctx->rowptr.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, u_dev, uold_dev,
                    kvals_dev, rhs_dev, {D.b[0], D.b[1], D.b[2]}};
    return ctx->dim == 3 ? launch_sch<3, EL_DYN>(ctx->stream, ctx->csr_maxblk, A)
                         : launch_sch<2, EL_DYN>(ctx->stream, ctx->csr_maxblk, A);
}

// ---- element state handed in by the host ------------------------------------
// The reference's seam passes the element's own host-side state on every
// call (perix_cuda::assembleStressCahnHilliard, CudaAssemble.h:185-191, and
// perix_cuda::assembleCached, CudaAssemble.h:216-219, called from
// FormResidualAndJacobianCUDA.cpp:229-302): the committed bond health, the
// committed velocity and the Picard-frozen mobility.  These entry points
// upload that state and run the same block-element kernels as the
// device-resident elements.

static int upload_health(pxg_ctx *ctx, const double *health)
{
    // health > 0 intact (the reference keeps 0.0 / 1.0 per bond)
    const long long nb = ctx->nb;
    std::vector<unsigned char> h((size_t)(nb > 0 ? nb : 1), 1);
    if (health)
        for (long long t = 0; t < nb; t++) h[(size_t)t] = health[t] > 0.0 ? 1 : 0;
    PXG_HIP(ctx->xhealth.alloc(h.size()));
    PXG_HIP(hipMemcpyAsync(ctx->xhealth.p, h.data(), h.size(), hipMemcpyHostToDevice, ctx->stream));
    // the host vector must outlive the copy
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_assemble_sch_state(pxg_ctx *ctx, const pxg_sch_state *s, double dt, const double *u,
                                      const double *uold, double *kvals, double *rhs)
{
    PXG_CHECK(ctx && ctx->nop >= (ctx->dim == 3 ? 9 : 5) && ctx->nnz > 0, PXG_ERR_STATE,
              "pxg_assemble_sch_state: needs order-2 operators and a CSR pattern");
    PXG_CHECK(ctx->csr_ndof == ctx->dim + 2, PXG_ERR_STATE, "pxg_assemble_sch_state: needs dim+2 dofs per node");
    PXG_CHECK(s && s->rev && s->lap && s->mnode && u && uold && kvals && rhs && dt > 0.0, PXG_ERR_ARG,
              "pxg_assemble_sch_state: bad arguments (the element is transient-only, dt > 0)");
    PXG_CHECK(!(s->rho > 0.0) || s->vel, PXG_ERR_ARG, "pxg_assemble_sch_state: rho > 0 needs the velocity");
    PXG_FAULT(ctx, "pxg_assemble_sch_state");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int n = ctx->n, dim = ctx->dim;
    const long long nb = ctx->nb;
    const size_t nd = (size_t)n * (dim + 2);
    PXG_HIP(ctx->xrev.alloc((size_t)nb));
    PXG_HIP(ctx->xlap.alloc((size_t)nb));
    PXG_HIP(ctx->xmob.alloc((size_t)n));
    PXG_HIP(ctx->xghost.alloc((size_t)n));
    PXG_HIP(hipMemcpyAsync(ctx->xrev.p, s->rev, sizeof(int) * (size_t)nb, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->xlap.p, s->lap, sizeof(double) * (size_t)nb, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->xmob.p, s->mnode, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, st));
    if (s->isghost) PXG_HIP(hipMemcpyAsync(ctx->xghost.p, s->isghost, (size_t)n, hipMemcpyHostToDevice, st));
    else PXG_HIP(hipMemsetAsync(ctx->xghost.p, 0, (size_t)n, st));
    const bool inertia = s->rho > 0.0;
    if (inertia) {
        PXG_HIP(ctx->xvel.alloc((size_t)n * dim));
        PXG_HIP(hipMemcpyAsync(ctx->xvel.p, s->vel, sizeof(double) * (size_t)n * dim, hipMemcpyHostToDevice, st));
    }
    {
        const int rc = upload_health(ctx, s->health);
        if (rc) return rc;
    }
    PXG_HIP(ctx->su.alloc(nd));
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(ctx->sk.alloc((size_t)ctx->nnz));
    PXG_HIP(ctx->sr.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    const SchArgs A{n, ctx->nop, ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->xrev.p, ctx->xlap.p, ctx->xghost.p,
                    ctx->vols.p,
                    FschConst{s->C11, s->C12, s->C66, s->A, s->Kh, s->Ch, 0.0, s->kappa, s->chi, s->Omega, s->cref,
                              inertia ? s->rho : 0.0, s->kres},
                    dt, inertia ? ctx->xvel.p : nullptr, ctx->xhealth.p, ctx->rowptr.p, ctx->slot.p, ctx->diag.p,
                    ctx->mslot.p, ctx->su.p, ctx->suo.p, ctx->sk.p, ctx->sr.p, {0.0, 0.0, 0.0}, ctx->xmob.p};
    const int rc = dim == 3 ? launch_sch<3, EL_FSCH>(st, ctx->csr_maxblk, A) : launch_sch<2, EL_FSCH>(st, ctx->csr_maxblk, A);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * (size_t)ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_assemble_dynfrac_state(pxg_ctx *ctx, const pxg_dynfrac_state *s, double dt, const double *u,
                                          const double *uold, double *kvals, double *rhs)
{
    PXG_CHECK(ctx && ctx->nop >= (ctx->dim == 3 ? 9 : 5) && ctx->nnz > 0, PXG_ERR_STATE,
              "pxg_assemble_dynfrac_state: needs order-2 operators and a CSR pattern");
    PXG_CHECK(ctx->csr_ndof == ctx->dim, PXG_ERR_STATE, "pxg_assemble_dynfrac_state: needs dim dofs per node");
    PXG_CHECK(s && s->vel && u && uold && kvals && rhs && dt > 0.0 && s->rho > 0.0, PXG_ERR_ARG,
              "pxg_assemble_dynfrac_state: bad arguments (dt > 0 and rho > 0)");
    PXG_FAULT(ctx, "pxg_assemble_dynfrac_state");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int n = ctx->n, dim = ctx->dim;
    const size_t nd = (size_t)n * dim;
    PXG_HIP(ctx->xvel.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->xvel.p, s->vel, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    {
        const int rc = upload_health(ctx, s->health);
        if (rc) return rc;
    }
    PXG_HIP(ctx->su.alloc(nd));
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(ctx->sk.alloc((size_t)ctx->nnz));
    PXG_HIP(ctx->sr.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    const SchArgs A{n, ctx->nop, ctx->off.p, ctx->nbr.p, ctx->ops.p, nullptr, nullptr, nullptr, ctx->vols.p,
                    FschConst{s->C11, s->C12, s->C66, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, s->rho, 1.0}, dt,
                    ctx->xvel.p, ctx->xhealth.p, ctx->rowptr.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, ctx->su.p,
                    ctx->suo.p, ctx->sk.p, ctx->sr.p, {s->bx, s->by, s->bz}, nullptr};
    const int rc = dim == 3 ? launch_sch<3, EL_DYN>(st, ctx->csr_maxblk, A) : launch_sch<2, EL_DYN>(st, ctx->csr_maxblk, A);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * (size_t)ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

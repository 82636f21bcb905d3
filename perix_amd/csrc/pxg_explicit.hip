// pxg_explicit.hip -- matrix-free explicit PDDO fracture on the device.
//
// Reference: ExplicitPDDOFracElement::preprocessIteration (per-step loop,
// ExplicitPDDOFracElement.cpp:237-299), the central-difference driver
// TimeStepping::solveExplicitDynamics (TimeStepping.cpp:417-599) and the BC
// presets (BCSystem.cpp:114-129, DirichletBC.cpp:107-147, NeumannBC.cpp:88-105).
//
// pxg_frac_run takes one launch per step: k_frac_step (pxg_step.hip) with
// damage, the Navier family sum, the central-difference update, the merged
// BC program and the blow-up guard folded in.  The adjoint's frozen-health
// recompute runs the step without presets followed by k_bc_preset (one
// thread per preset target node, guard decided by the launch's last block).
// This file holds k_bc_preset, the per-bond operator build, the health
// bookkeeping and the host side of the explicit loop.
#include "pxg_internal.h"
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <unordered_map>
#include <vector>

namespace pxg {

// BC presets on the freshly updated u^{n+1}, then the blow-up guard
// (TimeStepping.cpp:550-573) decided by the last block of the launch.
__global__ void __launch_bounds__(256) k_bc_preset(int count, int dim, const int *node, const int *src,
                                                   const int *mode, const double *val0, const double *vel,
                                                   double t, double *u, unsigned long long *gmax,
                                                   double *ghist, int *gstate, int step)
{
    __shared__ unsigned long long s_max[4];
    __shared__ int s_last;
    if (gstate[0]) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long mymax = 0;
    if (q < count) {
        const int i = node[q], b = src[q];
        for (int c = 0; c < dim; c++) {
            const int m = mode[q * 3 + c];
            if (m == 1) u[(size_t)i * dim + c] = u[(size_t)b * dim + c];
            else if (m == 2) {
                const double g = val0[q * 3 + c] + vel[q * 3 + c] * t;
                u[(size_t)i * dim + c] = 2.0 * g - u[(size_t)b * dim + c];
            } else if (m == 3) {
                u[(size_t)i * dim + c] = val0[q * 3 + c] + vel[q * 3 + c] * t;
            }
        }
        for (int c = 0; c < dim; c++) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(u[(size_t)i * dim + c]));
            mymax = bits > mymax ? bits : mymax;
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long o = __shfl_xor(mymax, m);
        mymax = o > mymax ? o : mymax;
    }
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mymax;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = s_max[0];
        for (int w = 1; w < 4; w++) m = s_max[w] > m ? s_max[w] : m;
        if (m) atomicMax(&gmax[1], m);
        __threadfence();
        const unsigned long long prev = atomicAdd(&gmax[2], 1ULL);
        s_last = (prev == (unsigned long long)(gridDim.x - 1));
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();
        const unsigned long long g0 = atomicAdd(&gmax[0], 0ULL), g1 = atomicAdd(&gmax[1], 0ULL);
        const unsigned long long m = g0 > g1 ? g0 : g1;
        const bool finite = m <= (unsigned long long)__double_as_longlong(DBL_MAX);
        const double umax = __longlong_as_double((long long)m);
        const int slot = (step - 1) % 50;
        const double past = (step > 50) ? ghist[slot] : -1.0;
        const double cap = (past >= 0.0) ? 1.0e3 * fmax(1.0, past) : INFINITY;
        if (!finite || umax > cap) { gstate[0] = 1; gstate[1] = step; }
        else ghist[slot] = umax;
        gstate[2] = step;
        gmax[0] = 0; gmax[1] = 0; gmax[2] = 0;
    }
}

// per-bond Navier operators from the g-weights (full slot order of
// PDOperators.cpp:51-75); component-major layout P[c*ps + t], ps = nb
// rounded up to a multiple of 64 so that every component starts aligned.
__global__ void k_frac_ops(int n, int dim, int nop, const int *off, const double *ops,
                           const double *nodeE, double C11, double C12, double C66, double nu,
                           int plane_stress, long long ps, const int *bsh, long long pst, double *P64,
                           float *P32, double *Phat)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    double c11 = C11, c12 = C12, c66 = C66;
    if (nodeE) {
        // ExplicitPDDOFracElement::recalcElasticConstants (:37-52) at E_i
        const double E = nodeE[i];
        if (E <= 0.0) { c11 = c12 = c66 = 0.0; }
        else if (dim >= 3 || !plane_stress) {
            const double d = (1.0 + nu) * (1.0 - 2.0 * nu);
            c11 = E * (1.0 - nu) / d; c12 = E * nu / d; c66 = E / (2.0 * (1.0 + nu));
        } else {
            const double d = 1.0 - nu * nu;
            c11 = E / d; c12 = E * nu / d; c66 = E / (2.0 * (1.0 + nu));
        }
    }
    double h11 = 0, h12 = 0, h66 = 0;
    if (Phat) {
        if (dim >= 3 || !plane_stress) {
            const double d = (1.0 + nu) * (1.0 - 2.0 * nu);
            h11 = 1.0 * (1.0 - nu) / d; h12 = 1.0 * nu / d; h66 = 1.0 / (2.0 * (1.0 + nu));
        } else {
            const double d = 1.0 - nu * nu;
            h11 = 1.0 / d; h12 = 1.0 * nu / d; h66 = 1.0 / (2.0 * (1.0 + nu));
        }
    }
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) {
        const double *op = ops + (size_t)t * nop;
        double p[6], q[6];
        int nc;
        if (dim >= 3) {
            const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
            const double aa = c12 + c66;
            p[0] = c11 * Gxx + c66 * (Gyy + Gzz);
            p[1] = c11 * Gyy + c66 * (Gxx + Gzz);
            p[2] = c11 * Gzz + c66 * (Gxx + Gyy);
            p[3] = aa * Gxy; p[4] = aa * Gxz; p[5] = aa * Gyz;
            const double ha = h12 + h66;
            q[0] = h11 * Gxx + h66 * (Gyy + Gzz);
            q[1] = h11 * Gyy + h66 * (Gxx + Gzz);
            q[2] = h11 * Gzz + h66 * (Gxx + Gyy);
            q[3] = ha * Gxy; q[4] = ha * Gxz; q[5] = ha * Gyz;
            nc = 6;
        } else {
            const double Gxx = op[2], Gxy = op[3], Gyy = op[4];
            p[0] = c11 * Gxx + c66 * Gyy;
            p[1] = c66 * Gxx + c11 * Gyy;
            p[2] = (c12 + c66) * Gxy;
            q[0] = h11 * Gxx + h66 * Gyy;
            q[1] = h66 * Gxx + h11 * Gyy;
            q[2] = (h12 + h66) * Gxy;
            nc = 3;
        }
        const long long tp = t + (bsh ? bsh[i] : 0);   // operator layout position
        for (int c = 0; c < nc; c++) {
            if (P64) P64[(size_t)c * pst + tp] = p[c];
            if (P32) P32[(size_t)c * pst + tp] = (float)p[c];
            if (Phat) Phat[(size_t)c * ps + t] = q[c];
        }
    }
}

// PXG_FRAC_FP32_OPERATORS: the FP64 operators rounded to FP32 and back
__global__ void k_round_f32(size_t count, double *P)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < count) P[q] = (double)(float)P[q];
}

// Bond state (health, break steps) lives in the step kernel's stream
// layout: CSR bond t of node i at t + bsh[i] (bsh null: CSR order).  These
// kernels convert at the API boundary; they are not on the step path.
__global__ void k_health_init(int n, const int *off, const int *bsh, const unsigned char *seed,
                              unsigned char *health, int *bs)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    const int sh = bsh ? bsh[i] : 0;
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) {
        health[t + sh] = seed[t];
        bs[t + sh] = seed[t] ? INT_MAX : -1;
    }
}

__global__ void k_health_from_double(int n, const int *off, const int *bsh, const double *h,
                                     unsigned char *health)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    const int sh = bsh ? bsh[i] : 0;
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) health[t + sh] = (h[t] > 0.0) ? 1 : 0;
}

__global__ void k_health_to_double(int n, const int *off, const int *bsh, const unsigned char *health,
                                   double *h)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    const int sh = bsh ? bsh[i] : 0;
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) h[t] = health[t + sh] ? 1.0 : 0.0;
}

__global__ void k_bonds_to_csr(int n, const int *off, const int *bsh, const unsigned char *health,
                               const int *bs, unsigned char *hc, int *bc)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    const int sh = bsh[i];
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) {
        if (hc) hc[t] = health[t + sh];
        if (bc) bc[t] = bs[t + sh];
    }
}

// Shared FP32 operator runs.  The order pattern of a family: a hash of its
// length and of its neighbour offsets in family order, in units of the grid
// spacing (rounded), so nodes on a regular lattice whose families visit the
// same offsets in the same order share a hash.
__global__ void k_pattern_hash(int n, int dim, const int *off, const int *nbr, const double *xp, double inv_h,
                               unsigned long long *hash)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long h = 1469598103934665603ull ^ (unsigned long long)(off[i + 1] - off[i]);
    for (int t = off[i]; t < off[i + 1]; t++) {
        const int j = nbr[t] - 1;
        for (int c = 0; c < dim; c++) {
            const long long q = llrint((xp[(size_t)j * dim + c] - xp[(size_t)i * dim + c]) * inv_h);
            h = (h ^ (unsigned long long)(q & 0xffff)) * 1099511628211ull;
        }
    }
    hash[i] = h;
}

// A node with candidate representative r shares r's operator run when the
// families have the same length and every FP64 operator of i is within
// tol x (the largest |operator| of i) of r's, bond by bond in family order
// (NaN never passes).  pofs[i] then moves i's operator loads onto r's run.
__global__ void k_class_check(int n, int nc, const int *off, const int *bsh, const int *rep, long long ps,
                              const double *P, double tol, int *pofs)
{
    // 16 lanes per node; a group's lanes are consecutive in one wave, so the
    // shuffles below stay inside the group
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    const int r = (i < n) ? rep[i] : -1;
    const int len = (i < n) ? off[i + 1] - off[i] : 0;
    const bool cand = i < n && r >= 0 && r != i && off[r + 1] - off[r] == len;
    double mx = 0.0;
    if (cand)
        for (int c = 0; c < nc; c++)
            for (int k = l; k < len; k += kConvLanes) {
                const double a = fabs(P[(size_t)c * ps + off[i] + k]);
                mx = a > mx ? a : mx;
            }
    for (int m = kConvLanes / 2; m >= 1; m >>= 1) { const double o = __shfl_xor(mx, m); mx = o > mx ? o : mx; }
    const double lim = tol * mx;
    int ok = cand ? 1 : 0;
    if (cand)
        for (int c = 0; c < nc && ok; c++)
            for (int k = l; k < len; k += kConvLanes)
                if (!(fabs(P[(size_t)c * ps + off[i] + k] - P[(size_t)c * ps + off[r] + k]) <= lim)) { ok = 0; break; }
    for (int m = kConvLanes / 2; m >= 1; m >>= 1) ok &= __shfl_xor(ok, m);
    if (i < n && l == 0) pofs[i] = ok ? (off[r] + bsh[r]) - (off[i] + bsh[i]) : 0;
}

// FNV-1a hash of node i's FP32 operator run in the step layout: its length
// and the bit patterns of every component, bond by bond (one thread per
// node; a set-up pass)
__global__ void k_run_hash32(int n, int nc, const int *off, const int *bsh, long long pst, const float *P32,
                             unsigned long long *hash)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int len = off[i + 1] - off[i];
    const long long p0 = off[i] + bsh[i];
    unsigned long long h = 1469598103934665603ull ^ (unsigned long long)len;
    for (int c = 0; c < nc; c++)
        for (int k = 0; k < len; k++)
            h = (h ^ (unsigned long long)__float_as_uint(P32[(size_t)c * pst + p0 + k])) * 1099511628211ull;
    hash[i] = h;
}

// Exact form of k_class_check on the FP32 operators themselves: node i
// shares candidate r's run when the runs have the same length and every
// FP32 operator is the same bit pattern, bond by bond.  The step then reads
// the very values of i's own run, so sharing cannot change a result.
__global__ void k_class_check32(int n, int nc, const int *off, const int *bsh, const int *rep, long long pst,
                                const float *P32, int *pofs)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    const int r = (i < n) ? rep[i] : -1;
    const int len = (i < n) ? off[i + 1] - off[i] : 0;
    const bool cand = i < n && r >= 0 && r != i && off[r + 1] - off[r] == len;
    int ok = cand ? 1 : 0;
    if (cand) {
        const long long pi = off[i] + bsh[i], pr = off[r] + bsh[r];
        for (int c = 0; c < nc && ok; c++)
            for (int k = l; k < len; k += kConvLanes)
                if (__float_as_uint(P32[(size_t)c * pst + pi + k]) != __float_as_uint(P32[(size_t)c * pst + pr + k])) {
                    ok = 0;
                    break;
                }
    }
    for (int m = kConvLanes / 2; m >= 1; m >>= 1) ok &= __shfl_xor(ok, m);
    if (i < n && l == 0) pofs[i] = ok ? (off[r] + bsh[r]) - (off[i] + bsh[i]) : 0;
}

// Nodal output projections of the explicit element (ExplicitPDDOFracElement::
// computeNodalProjection, ExplicitPDDOFracElement.cpp:324-405): which = 0
// damage = 1 - (intact bonds)/|H_i|; otherwise the displacement gradient
// sum_j (u_j - u_i) (x) G(i,j) over the whole family in family order
// (first-derivative g-weights, health not consulted), the small strain, the
// Navier stress (sigma_zz = C12 tr(eps) in 2D plane strain, 0 in plane
// stress), von Mises (1), stress (2) or strain (3) as row-major 3x3.
//
// A one-wave block holds 4 nodes of 16 lanes.  Per chunk of 16 bonds each
// lane forms its bond's 9 products (u_j - u_i)_a G_b with coalesced loads of
// the family's contiguous ids and g-weights; lanes 0..8 then add component
// a*3+b over the chunk in family order, so every sum is the sequential one.
constexpr int kProjLanes = 16, kProjNodes = 64 / kProjLanes;

__device__ void frac_project_out(int dim, int which, int plane_stress, double C12, double C66,
                                 const double *gu, double *o)
{
    double eps[6] = {0, 0, 0, 0, 0, 0};
    if (dim >= 3) {
        eps[0] = gu[0]; eps[1] = gu[4]; eps[2] = gu[8];
        eps[3] = 0.5 * (gu[1] + gu[3]); eps[4] = 0.5 * (gu[2] + gu[6]); eps[5] = 0.5 * (gu[5] + gu[7]);
    } else {
        eps[0] = gu[0]; eps[1] = gu[4]; eps[3] = 0.5 * (gu[1] + gu[3]);
    }
    const double tr = eps[0] + eps[1] + eps[2];
    double sig[6];
    sig[0] = C12 * tr + 2.0 * C66 * eps[0]; sig[1] = C12 * tr + 2.0 * C66 * eps[1];
    sig[2] = (dim >= 3) ? (C12 * tr + 2.0 * C66 * eps[2]) : (!plane_stress ? C12 * tr : 0.0);
    sig[3] = 2.0 * C66 * eps[3]; sig[4] = 2.0 * C66 * eps[4]; sig[5] = 2.0 * C66 * eps[5];
    if (which == 1) {
        const double sm = (sig[0] + sig[1] + sig[2]) / 3.0;
        const double a = sig[0] - sm, c = sig[1] - sm, d = sig[2] - sm;
        o[0] = sqrt(1.5 * (a * a + c * c + d * d + 2.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5])));
        return;
    }
    const double *v = (which == 2) ? sig : eps;
    o[0] = v[0]; o[1] = v[3]; o[2] = v[4];
    o[3] = v[3]; o[4] = v[1]; o[5] = v[5];
    o[6] = v[4]; o[7] = v[5]; o[8] = v[2];
}

template <int DIM>
__global__ void __launch_bounds__(64) k_frac_project(int n, int nop, int which, int plane_stress, double C12,
                                                     double C66, const int *off, const int *nbr,
                                                     const double *ops, const unsigned char *health,
                                                     const int *bsh, const double *u, double *out)
{
    __shared__ double s_p[kProjNodes][kProjLanes][9];
    __shared__ double s_gu[kProjNodes][9];
    const int g = threadIdx.x / kProjLanes, l = threadIdx.x % kProjLanes;
    const int i = blockIdx.x * kProjNodes + g;
    const bool act = i < n;
    const int b = act ? off[i] : 0;
    const int nb = act ? off[i + 1] - b : 0;
    if (which == 0) {
        // intact-bond count: integer, so the lane split does not change it
        int live = 0;
        if (act) {
            const int sh = bsh ? bsh[i] : 0;
            for (int t = b + l; t < b + nb; t += kProjLanes) live += health[t + sh] ? 1 : 0;
        }
        for (int m = kProjLanes / 2; m >= 1; m >>= 1) live += __shfl_xor(live, m);
        if (act && l == 0) out[i] = nb <= 0 ? 0.0 : 1.0 - (double)live / (double)nb;
        return;
    }
    double ui[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) ui[d] = act ? u[(size_t)i * DIM + d] : 0.0;
    // block-uniform chunk count (the block is one wave)
    int mx = nb;
    for (int m = 32; m >= 1; m >>= 1) mx = max(mx, __shfl_xor(mx, m));
    double acc = 0.0;   // lane l < 9 owns component l = a*3 + b
    for (int c = 0; c < mx; c += kProjLanes) {
        double p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int t = b + c + l;
        if (c + l < nb) {
            const size_t rowJ = (size_t)(nbr[t] - 1) * DIM;
            const double *op = ops + (size_t)t * nop;
            double G[DIM], du[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) { G[d] = op[d]; du[d] = u[rowJ + d] - ui[d]; }
#pragma unroll
            for (int a = 0; a < DIM; a++)
#pragma unroll
                for (int q = 0; q < DIM; q++) p[a * 3 + q] = du[a] * G[q];
        }
#pragma unroll
        for (int k = 0; k < 9; k++) s_p[g][l][k] = p[k];
        __syncthreads();
        if (l < 9) {
            const int cnt = min(kProjLanes, nb - c);
            for (int q = 0; q < cnt; q++) acc += s_p[g][q][l];
        }
        __syncthreads();
    }
    if (l < 9) s_gu[g][l] = acc;
    __syncthreads();
    if (act && l == 0) {
        double gu[9];
#pragma unroll
        for (int k = 0; k < 9; k++) gu[k] = s_gu[g][k];
        frac_project_out(DIM, which, plane_stress, C12, C66, gu, out + (size_t)i * ((which >= 2) ? 9 : 1));
    }
}

__global__ void k_rev(int n, const int *off, const int *nbr, int *rev)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes) {
        const int j = nbr[t] - 1;
        int r = -1;
        for (int s = off[j]; s < off[j + 1]; s++) if (nbr[s] == i + 1) { r = s; break; }
        rev[t] = r;
    }
}

} // namespace pxg

using namespace pxg;

static unsigned gridn(long long n, int b) { return (unsigned)((n + b - 1) / b); }

namespace pxg {

int launch_frac_ops(pxg_ctx *ctx, const double *nodeE_dev)
{
    FracState &F = ctx->frac;
    hipLaunchKernelGGL(k_frac_ops, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, ctx->stream, ctx->n, ctx->dim,
                       ctx->nop, ctx->off.p, ctx->ops.p, nodeE_dev, F.C11, F.C12, F.C66, F.nu, F.plane_stress,
                       F.ps, F.bsh.p, F.pst, F.P64.p, (float *)nullptr, (double *)nullptr);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

int bond_state_csr(pxg_ctx *ctx, const unsigned char **health, const int **break_step)
{
    FracState &F = ctx->frac;
    if (!F.bsh.p) {
        if (health) *health = F.health.p;
        if (break_step) *break_step = F.break_step.p;
        return PXG_OK;
    }
    if (health) PXG_HIP(F.health_csr.alloc(ctx->nb));
    if (break_step) PXG_HIP(F.bs_csr.alloc(ctx->nb));
    hipLaunchKernelGGL(k_bonds_to_csr, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, ctx->stream, ctx->n, ctx->off.p,
                       F.bsh.p, F.health.p, F.break_step.p, health ? F.health_csr.p : nullptr,
                       break_step ? F.bs_csr.p : nullptr);
    PXG_HIP(hipGetLastError());
    if (health) *health = F.health_csr.p;
    if (break_step) *break_step = F.bs_csr.p;
    return PXG_OK;
}

// FP32C shared operator runs (F.pofs): a node reads a representative's FP32
// operator run instead of its own.
//  * 2D: only runs that are bit for bit the node's own FP32 operators, so
//    the step's results are exactly those of every node on its own run.  On
//    the 400x200 plate the interior nodes' FP64 operators differ by about
//    1e-14 relative, far inside one FP32 rounding, and the 81,200 runs
//    collapse to about 3,900 distinct ones, which stay in L2.
//  * 3D: nodes of one family order pattern whose FP64 operators agree to
//    kShareTol.  The 122-bond 3D runs almost never round to identical FP32
//    values (about 41,000 distinct runs for 75,200 nodes of an 80x40x20
//    block), so exact sharing would find little; a tolerance class applies
//    one FP32 rounding of the representative to every member.  On the
//    160x80x40 block its trajectory is the own-operator trajectory to 1e-12.
//    On the 2D plate such classes turned the members' scattered roundings
//    into a coherent drift that moved the crack's branching, hence the exact
//    rule there.
constexpr double kShareTol = 1e-12;

static int build_op_classes(pxg_ctx *ctx, const double *nodeE_dev)
{
    FracState &F = ctx->frac;
    F.pofs.release();
    F.nshared = 0;
    F.nclasses = 0;
    if (!F.bsh.p || F.precision != PXG_FP32C || (F.opts & PXG_FRAC_NO_SHARED_RUNS)) return PXG_OK;
    const int n = ctx->n, nc = (ctx->dim == 3) ? 6 : 3;
    const bool exact = ctx->dim != 3;
    hipStream_t st = ctx->stream;
    dbuf<double> P;                     // FP64 operators, CSR order (tolerance classes)
    dbuf<unsigned long long> hash;
    PXG_HIP(hash.alloc(n));
    if (exact) {
        hipLaunchKernelGGL(k_run_hash32, dim3(gridn(n, 128)), dim3(128), 0, st, n, nc, ctx->off.p, F.bsh.p, F.pst,
                           F.P32.p, hash.p);
    } else {
        PXG_HIP(P.alloc((size_t)F.ps * nc));
        hipLaunchKernelGGL(k_frac_ops, dim3(gridn(n, kConvNodes)), dim3(kConvBlock), 0, st, n, ctx->dim, ctx->nop,
                           ctx->off.p, ctx->ops.p, nodeE_dev, F.C11, F.C12, F.C66, F.nu, F.plane_stress, F.ps,
                           (const int *)nullptr, F.ps, P.p, (float *)nullptr, (double *)nullptr);
        PXG_HIP(hipGetLastError());
        const double h = std::max(ctx->dxyz[0], std::max(ctx->dxyz[1], ctx->dxyz[2]));
        PXG_CHECK(h > 0.0, PXG_ERR_STATE, "build_op_classes: zero grid spacing");
        hipLaunchKernelGGL(k_pattern_hash, dim3(gridn(n, 128)), dim3(128), 0, st, n, ctx->dim, ctx->off.p,
                           ctx->nbr.p, ctx->xp.p, 1.0 / h, hash.p);
    }
    PXG_HIP(hipGetLastError());
    std::vector<unsigned long long> hh((size_t)n);
    PXG_HIP(hipMemcpyAsync(hh.data(), hash.p, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    // candidate representative: the first node of each pattern
    std::unordered_map<unsigned long long, int> first;
    first.reserve((size_t)n / 8 + 16);
    std::vector<int> rep((size_t)n, -1);
    for (int i = 0; i < n; i++) {
        auto it = first.emplace(hh[(size_t)i], i).first;
        rep[(size_t)i] = it->second;
    }
    dbuf<int> drep;
    PXG_HIP(drep.alloc(n));
    PXG_HIP(F.pofs.alloc(n));
    PXG_HIP(hipMemcpyAsync(drep.p, rep.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    if (exact)
        hipLaunchKernelGGL(k_class_check32, dim3(gridn(n, kConvNodes)), dim3(kConvBlock), 0, st, n, nc, ctx->off.p,
                           F.bsh.p, drep.p, F.pst, F.P32.p, F.pofs.p);
    else
        hipLaunchKernelGGL(k_class_check, dim3(gridn(n, kConvNodes)), dim3(kConvBlock), 0, st, n, nc, ctx->off.p,
                           F.bsh.p, drep.p, F.ps, P.p, kShareTol, F.pofs.p);
    PXG_HIP(hipGetLastError());
    std::vector<int> pofs((size_t)n);
    PXG_HIP(hipMemcpyAsync(pofs.data(), F.pofs.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    std::vector<char> used((size_t)n, 0);
    for (int i = 0; i < n; i++)
        if (pofs[(size_t)i] != 0) {
            F.nshared++;
            if (!used[(size_t)rep[(size_t)i]]) { used[(size_t)rep[(size_t)i]] = 1; F.nclasses++; }
        }
    return PXG_OK;
}

int launch_preset(pxg_ctx *ctx, double *u, double t, int step)
{
    FracState &F = ctx->frac;
    const int cnt = F.bc.count;
    hipLaunchKernelGGL(k_bc_preset, dim3(cnt > 0 ? gridn(cnt, 256) : 1), dim3(256), 0, ctx->stream, cnt,
                       ctx->dim, F.bc.node.p, F.bc.src.p, F.bc.mode.p, F.bc.val0.p, F.bc.vel.p, t, u,
                       F.gmax.p, F.ghist.p, F.gstate.p, step);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

} // namespace pxg

extern "C" int pxg_frac_setup(pxg_ctx *ctx, const pxg_frac_params *p, int precision,
                              const double *node_E, double nu, int plane_stress)
{
    PXG_CHECK(ctx && p, PXG_ERR_ARG, "pxg_frac_setup: bad arguments");
    PXG_CHECK(ctx->nop > 0, PXG_ERR_STATE, "pxg_frac_setup: build the operators first");
    PXG_CHECK(p->dim == ctx->dim, PXG_ERR_ARG, "pxg_frac_setup: dim mismatch");
    PXG_CHECK(precision == PXG_FP64 || precision == PXG_FP32C, PXG_ERR_ARG, "pxg_frac_setup: bad precision");
    PXG_CHECK(p->rho > 0.0, PXG_ERR_ARG, "pxg_frac_setup: rho must be > 0");
    PXG_FAULT(ctx, "pxg_frac_setup");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    F.dim = p->dim; F.C11 = p->C11; F.C12 = p->C12; F.C66 = p->C66; F.rho = p->rho;
    F.b[0] = p->bx; F.b[1] = p->by; F.b[2] = p->bz; F.sc = p->sc;
    F.tension_only = p->tension_only; F.damage_on = p->damage_on;
    F.precision = precision; F.nu = nu; F.plane_stress = plane_stress;
    F.has_nodeE = node_E ? 1 : 0;
    const int nc = (ctx->dim == 3) ? 6 : 3;
    const long long nb = ctx->nb;
    F.ps = (nb + 63) / 64 * 64;
    if (node_E) {
        PXG_HIP(F.nodeE.alloc(ctx->n));
        PXG_HIP(hipMemcpyAsync(F.nodeE.p, node_E, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
        PXG_HIP(F.Phat.alloc((size_t)F.ps * nc));
        PXG_HIP(F.rev.alloc(nb));
        hipLaunchKernelGGL(k_rev, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->n, ctx->off.p, ctx->nbr.p, F.rev.p);
    } else {
        F.nodeE.release(); F.Phat.release(); F.rev.release();
    }
    {
        const int rc = build_step_gather(ctx);   // node tiles and the operator layout
        if (rc) return rc;
    }
    if (precision == PXG_FP64) { PXG_HIP(F.P64.alloc((size_t)F.pst * nc)); F.P32.release(); }
    else { PXG_HIP(F.P32.alloc((size_t)F.pst * nc)); if (!node_E) F.P64.release(); else PXG_HIP(F.P64.alloc((size_t)F.pst * nc)); }
    hipLaunchKernelGGL(k_frac_ops, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->n, ctx->dim, ctx->nop,
                       ctx->off.p, ctx->ops.p, node_E ? F.nodeE.p : nullptr, p->C11, p->C12, p->C66, nu,
                       plane_stress, F.ps, F.bsh.p, F.pst, F.P64.p, F.P32.p, node_E ? F.Phat.p : nullptr);
    PXG_HIP(hipGetLastError());
    if (precision == PXG_FP64 && (F.opts & PXG_FRAC_FP32_OPERATORS)) {
        const size_t cnt = (size_t)F.pst * nc;
        hipLaunchKernelGGL(k_round_f32, dim3(gridn((long long)cnt, 256)), dim3(256), 0, st, cnt, F.P64.p);
        PXG_HIP(hipGetLastError());
    }
    {
        const int rc = build_op_classes(ctx, node_E ? F.nodeE.p : nullptr);
        if (rc) return rc;
    }
    if (node_E) {
        const int rc = launch_transpose_ops(ctx);
        if (rc) return rc;
    } else {
        F.PhatT.release();
    }
    PXG_HIP(F.health.alloc((size_t)F.pst));
    PXG_HIP(F.break_step.alloc((size_t)F.pst));
    // layout padding: dead in every mode (health 0, a break step below any step)
    PXG_HIP(hipMemsetAsync(F.break_step.p, 0x80, sizeof(int) * (size_t)F.pst, st));
    F.health_csr.release(); F.bs_csr.release();
    hipLaunchKernelGGL(k_health_init, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->n, ctx->off.p, F.bsh.p,
                       ctx->seed.p, F.health.p, F.break_step.p);
    for (int k = 0; k < 2; k++) PXG_HIP(F.u[k].alloc((size_t)ctx->n * ctx->dim));
    PXG_HIP(F.accel.alloc((size_t)ctx->n * ctx->dim));
    PXG_HIP(F.preset_mask.alloc(ctx->n));
    PXG_HIP(hipMemsetAsync(F.preset_mask.p, 0, ctx->n, st));
    PXG_HIP(F.gmax.alloc(kGuardWords));
    PXG_HIP(F.ghist.alloc(50));
    PXG_HIP(F.ghist64.alloc(64));
    // empty BC program until pxg_frac_set_bcs
    PXG_HIP(F.bc.prog_of.alloc(ctx->n));
    PXG_HIP(hipMemsetAsync(F.bc.prog_of.p, 0xFF, sizeof(int) * ctx->n, st));
    PXG_HIP(F.bc.inv_off.alloc((size_t)ctx->n + 1));
    PXG_HIP(hipMemsetAsync(F.bc.inv_off.p, 0, sizeof(int) * ((size_t)ctx->n + 1), st));
    PXG_HIP(F.bc.inv_q.alloc(1));
    PXG_HIP(F.bc.node.alloc(1));               // the step reads entry 0 whatever the program
    PXG_HIP(hipMemsetAsync(F.bc.node.p, 0, sizeof(int), st));
    PXG_HIP(F.gstate.alloc(4));
    if (!F.gstate_host) PXG_HIP(hipHostMalloc((void **)&F.gstate_host, 4 * sizeof(int), hipHostMallocDefault));
    PXG_HIP(F.anybroke.alloc(1));
    PXG_HIP(hipMemsetAsync(F.gmax.p, 0, sizeof(unsigned long long) * kGuardWords, st));
    PXG_HIP(hipMemsetAsync(F.ghist.p, 0, sizeof(double) * 50, st));
    PXG_HIP(hipMemsetAsync(F.ghist64.p, 0, sizeof(double) * 64, st));
    PXG_HIP(hipMemsetAsync(F.gstate.p, 0, sizeof(int) * 4, st));
    PXG_HIP(hipMemsetAsync(F.anybroke.p, 0, sizeof(int), st));
    F.bc.count = 0;
    F.health_uploaded = 0;
    F.cur = 0; F.t = 0.0; F.step = 0; F.dt_prev = 0.0;
    {
        const int rc = build_step_records(ctx);
        if (rc) return rc;
    }
    PXG_HIP(hipStreamSynchronize(st));
    F.ready = 1;
    return PXG_OK;
}

extern "C" int pxg_frac_options(pxg_ctx *ctx, int flags)
{
    PXG_CHECK(ctx && (flags & ~(PXG_FRAC_NO_SHARED_RUNS | PXG_FRAC_FP32_OPERATORS)) == 0, PXG_ERR_ARG,
              "pxg_frac_options: unknown flags");
    ctx->frac.opts = flags;
    return PXG_OK;
}

extern "C" int pxg_frac_set_bcs(pxg_ctx *ctx, int nbc, const pxg_bc *bcs)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_set_bcs: call pxg_frac_setup first");
    PXG_CHECK(nbc >= 0 && (nbc == 0 || bcs), PXG_ERR_ARG, "pxg_frac_set_bcs: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    const int n = ctx->n, dim = ctx->dim;
    // resolve the BC list (later entries override earlier ones) into one
    // program entry per target node
    std::vector<int> slot(n, -1), tnode, tsrc, tmode;
    std::vector<double> tval, tvel;
    for (int e = 0; e < nbc; e++) {
        const pxg_bc &b = bcs[e];
        PXG_CHECK(b.ncomp >= 0 && b.ncomp <= dim && (b.nnodes == 0 || b.nodes), PXG_ERR_ARG,
                  "pxg_frac_set_bcs: bad BC entry");
        for (int q = 0; q < b.nnodes; q++) {
            const int id = b.nodes[q];
            PXG_CHECK(id >= 1 && id <= n, PXG_ERR_ARG, "pxg_frac_set_bcs: node outside the PD mesh");
            const int mb = ctx->hmirror[id - 1];
            if (b.type == 0) {
                PXG_CHECK(mb >= 1, PXG_ERR_ARG, "pxg_frac_set_bcs: neumann node is not a boundary ghost "
                                                "with a mirror bulk (NeumannBC.cpp:60-78)");
            } else if (b.has_box) {
                bool in = true;
                for (int ax = 0; ax < 3; ax++) {
                    const double lo = b.box[2 * ax], hi = b.box[2 * ax + 1];
                    if (lo > hi) continue;
                    const double x = ctx->hcoords[3 * (size_t)(id - 1) + ax];
                    if (x < lo || x > hi) { in = false; break; }
                }
                if (!in) continue;
            }
            int s = slot[id - 1];
            if (s < 0) {
                s = (int)tnode.size();
                slot[id - 1] = s;
                tnode.push_back(id - 1); tsrc.push_back(mb >= 1 ? mb - 1 : -1);
                for (int c = 0; c < 3; c++) { tmode.push_back(0); tval.push_back(0.0); tvel.push_back(0.0); }
            }
            for (int k = 0; k < b.ncomp; k++) {
                const int c = b.comp[k];
                PXG_CHECK(c >= 0 && c < dim, PXG_ERR_ARG, "pxg_frac_set_bcs: component out of range");
                int m;
                if (b.type == 0) m = 1;
                else m = (mb >= 1 && !b.direct) ? 2 : 3;
                tmode[3 * s + c] = m;
                tval[3 * s + c] = b.value[k];
                tvel[3 * s + c] = b.velocity[k];
            }
        }
    }
    // the merged program is exact when no target is read as a mirror source
    for (size_t s = 0; s < tnode.size(); s++) {
        bool reads = false;
        for (int c = 0; c < 3; c++) reads |= (tmode[3 * s + c] == 1 || tmode[3 * s + c] == 2);
        if (reads && tsrc[s] >= 0) {
            PXG_CHECK(slot[tsrc[s]] < 0, PXG_ERR_ARG,
                      "pxg_frac_set_bcs: a BC reads a node that another BC writes; chained presets "
                      "are not supported on the device");
        }
    }
    FracState &F = ctx->frac;
    const int cnt = (int)tnode.size();
    F.bc.count = cnt;
    std::vector<unsigned char> mask(n, 0);
    for (int s = 0; s < cnt; s++) mask[tnode[s]] = 1;
    const int m1 = cnt > 0 ? cnt : 1;
    PXG_HIP(F.bc.node.alloc(m1)); PXG_HIP(F.bc.src.alloc(m1)); PXG_HIP(F.bc.mode.alloc(3 * m1));
    PXG_HIP(F.bc.val0.alloc(3 * m1)); PXG_HIP(F.bc.vel.alloc(3 * m1));
    hipStream_t st = ctx->stream;
    if (cnt > 0) {
        PXG_HIP(hipMemcpyAsync(F.bc.node.p, tnode.data(), sizeof(int) * cnt, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(F.bc.src.p, tsrc.data(), sizeof(int) * cnt, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(F.bc.mode.p, tmode.data(), sizeof(int) * 3 * cnt, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(F.bc.val0.p, tval.data(), sizeof(double) * 3 * cnt, hipMemcpyHostToDevice, st));
        PXG_HIP(hipMemcpyAsync(F.bc.vel.p, tvel.data(), sizeof(double) * 3 * cnt, hipMemcpyHostToDevice, st));
    }
    PXG_HIP(hipMemcpyAsync(F.preset_mask.p, mask.data(), n, hipMemcpyHostToDevice, st));
    PXG_HIP(F.bc.prog_of.alloc(n));
    PXG_HIP(hipMemcpyAsync(F.bc.prog_of.p, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    // inverse of the source map (adjoint of the preset), entries in program order
    std::vector<int> ioff(n + 1, 0), iq;
    for (int s2 = 0; s2 < cnt; s2++) if (tsrc[s2] >= 0) ioff[tsrc[s2] + 1]++;
    for (int m = 0; m < n; m++) {
        PXG_CHECK(ioff[m + 1] <= 0xffff, PXG_ERR_ARG, "pxg_frac_set_bcs: a node mirrors more than 65535 targets");
        ioff[m + 1] += ioff[m];
    }
    iq.assign(ioff[n] > 0 ? ioff[n] : 1, 0);
    {
        std::vector<int> cur(ioff.begin(), ioff.end() - 1);
        for (int s2 = 0; s2 < cnt; s2++) if (tsrc[s2] >= 0) iq[cur[tsrc[s2]]++] = s2;
    }
    PXG_HIP(F.bc.inv_off.alloc(n + 1));
    PXG_HIP(F.bc.inv_q.alloc(iq.size()));
    PXG_HIP(hipMemcpyAsync(F.bc.inv_off.p, ioff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(F.bc.inv_q.p, iq.data(), sizeof(int) * iq.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipStreamSynchronize(st));
    return build_step_records(ctx);
}

extern "C" int pxg_frac_accel(pxg_ctx *ctx, const double *u, double *health_inout, double *accel_out,
                              int *any_broke)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_accel: call pxg_frac_setup first");
    PXG_CHECK(u && accel_out, PXG_ERR_ARG, "pxg_frac_accel: bad arguments");
    PXG_FAULT(ctx, "pxg_frac_accel");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    if (health_inout && !F.health_uploaded) {
        PXG_HIP(ctx->sk.alloc(ctx->nb));
        PXG_HIP(hipMemcpyAsync(ctx->sk.p, health_inout, sizeof(double) * ctx->nb, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_health_from_double, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->n, ctx->off.p,
                           F.bsh.p, ctx->sk.p, F.health.p);
        F.health_uploaded = 1;
    }
    PXG_HIP(hipMemcpyAsync(F.u[F.cur].p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemsetAsync(F.anybroke.p, 0, sizeof(int), st));
    int rc = launch_step(ctx, MODE_ACCEL, F.u[F.cur].p, F.u[F.cur].p, F.accel.p, 0.0, 0.0, F.step, false, INT_MIN);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(accel_out, F.accel.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    int broke = 0;
    PXG_HIP(hipMemcpyAsync(&broke, F.anybroke.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    if (broke && health_inout) {
        PXG_HIP(ctx->sk.alloc(ctx->nb));
        hipLaunchKernelGGL(k_health_to_double, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->n, ctx->off.p,
                           F.bsh.p, F.health.p, ctx->sk.p);
        PXG_HIP(hipMemcpyAsync(health_inout, ctx->sk.p, sizeof(double) * ctx->nb, hipMemcpyDeviceToHost, st));
        PXG_HIP(hipStreamSynchronize(st));
    }
    if (any_broke) *any_broke = broke ? 1 : 0;
    return PXG_OK;
}

extern "C" int pxg_frac_init(pxg_ctx *ctx, const double *u0, const double *v0, double dt)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_init: call pxg_frac_setup first");
    PXG_CHECK(u0 && dt > 0.0, PXG_ERR_ARG, "pxg_frac_init: bad arguments");
    PXG_FAULT(ctx, "pxg_frac_init");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    F.cur = 0;
    // a new run starts from the seeded bond health
    hipLaunchKernelGGL(k_health_init, dim3(gridn(ctx->n, kConvNodes)), dim3(kConvBlock), 0, st, ctx->n, ctx->off.p, F.bsh.p,
                       ctx->seed.p, F.health.p, F.break_step.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipMemcpyAsync(F.u[0].p, u0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    if (v0) PXG_HIP(hipMemcpyAsync(F.accel.p, v0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    else PXG_HIP(hipMemsetAsync(F.accel.p, 0, sizeof(double) * nd, st));
    PXG_HIP(hipMemsetAsync(F.gmax.p, 0, sizeof(unsigned long long) * kGuardWords, st));
    PXG_HIP(hipMemsetAsync(F.gstate.p, 0, sizeof(int) * 4, st));
    PXG_HIP(hipMemsetAsync(F.ghist64.p, 0, sizeof(double) * 64, st));
    // t = 0 rate evaluation (damage included), then u^{-1} = u0 - dt v0 + dt^2/2 a0
    const double half = 0.5 * dt * dt;
    int rc = launch_step(ctx, MODE_INIT, F.u[0].p, F.accel.p, F.u[1].p, dt, half, 0, false, INT_MIN);
    if (rc) return rc;
    PXG_HIP(hipStreamSynchronize(st));
    F.t = 0.0; F.dt_prev = dt; F.step = 0; F.cur0 = 0;
    return PXG_OK;
}

// closes the open step-timing pair after the last launch of a run
static hipError_t close_step_events(pxg_ctx *ctx)
{
    FracState &F = ctx->frac;
    if (!F.ev_open) return hipSuccess;
    const hipError_t e = hipEventRecord(F.ev[2 * F.nev + 1], ctx->stream);
    if (e != hipSuccess) return e;
    F.nev++; F.nlaunch += F.ev_pending;
    F.ev_open = 0; F.ev_pending = 0;
    return hipSuccess;
}

extern "C" int pxg_frac_run(pxg_ctx *ctx, double dt, double total, int nsteps_max, int *steps_done)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_run: call pxg_frac_setup first");
    PXG_CHECK(dt > 0.0 && total > 0.0, PXG_ERR_ARG, "pxg_frac_run: dt and total_time must be positive");
    PXG_FAULT(ctx, "pxg_frac_run");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    const double tEps = 1.0e-12 * total;
    int done = 0;
    int rc = PXG_OK;
    while (F.t < total - tEps && (nsteps_max <= 0 || done < nsteps_max)) {
        const int step = F.step + 1;
        const double dtA = fmin(dt, total - F.t);
        const double tNew = F.t + dtA;
        double *ucur = F.u[F.cur].p, *uold = F.u[F.cur ^ 1].p;
        if (F.prof && !F.ev_open) {
            // one event pair spans each run of back-to-back launches between
            // guard polls, so the markers do not sit between the kernels
            if ((size_t)(2 * F.nev + 2) > F.ev.size()) {
                hipEvent_t a, b;
                PXG_HIP(hipEventCreate(&a));
                PXG_HIP(hipEventCreate(&b));
                F.ev.push_back(a); F.ev.push_back(b);
            }
            PXG_HIP(hipEventRecord(F.ev[2 * F.nev], ctx->stream));
            F.ev_open = 1; F.ev_pending = 0;
        }
        // one launch per step: damage, forces, update, BC presets at tNew and
        // the guard decision on the previous step
        if (dtA == F.dt_prev) {
            rc = launch_step_fused(ctx, MODE_CD, ucur, uold, uold, dtA * dtA, 0.0, step, tNew);
        } else {
            const double rr = dtA / F.dt_prev;
            const double w = 0.5 * dtA * (dtA + F.dt_prev);
            rc = launch_step_fused(ctx, MODE_VAR, ucur, uold, uold, rr, w, step, tNew);
        }
        if (rc) return rc;
        if (F.prof) F.ev_pending++;
        F.cur ^= 1;
        F.t = tNew; F.dt_prev = dtA; F.step = step;
        done++;
        if ((done & 255) == 0) {
            if (F.prof) PXG_HIP(close_step_events(ctx));
            int gs[3];
            PXG_HIP(hipMemcpyAsync(gs, F.gstate.p, sizeof(gs), hipMemcpyDeviceToHost, ctx->stream));
            PXG_HIP(hipStreamSynchronize(ctx->stream));
            if (gs[0]) break;
        }
    }
    if (F.prof) PXG_HIP(close_step_events(ctx));
    int gs[3];
    if (done > 0) {
        // the last step's guard decision, delivered to the pinned mirror
        rc = launch_guard_final(ctx, F.step);
        if (rc) return rc;
        PXG_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < 3; c++) gs[c] = __atomic_load_n(&F.gstate_host[c], __ATOMIC_ACQUIRE);
    } else {
        PXG_HIP(hipMemcpyAsync(gs, F.gstate.p, sizeof(gs), hipMemcpyDeviceToHost, ctx->stream));
        PXG_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (gs[0]) {
        // state at the aborted step: the buffer written by that step
        const int abort_step = gs[1];
        F.cur = F.cur0 ^ (abort_step & 1);
        F.step = abort_step;
        if (steps_done) *steps_done = abort_step - 1;
        ::pxg::set_error("pxg_frac_run: solution blew up at step " + std::to_string(abort_step) +
                         " (non-finite or 1e3x growth over 50 steps; reduce dt below the CFL limit)");
        return PXG_ERR_BLOWUP;
    }
    if (steps_done) *steps_done = done;
    return PXG_OK;
}

extern "C" int pxg_frac_get_state(pxg_ctx *ctx, double *u, double *uold, unsigned char *health, int *break_step)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_get_state: no explicit state");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    if (u) PXG_HIP(hipMemcpyAsync(u, F.u[F.cur].p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    if (uold) PXG_HIP(hipMemcpyAsync(uold, F.u[F.cur ^ 1].p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    const unsigned char *hc = nullptr;
    const int *bc = nullptr;
    if (health || break_step) {
        const int rc = bond_state_csr(ctx, health ? &hc : nullptr, break_step ? &bc : nullptr);
        if (rc) return rc;
    }
    if (health) PXG_HIP(hipMemcpyAsync(health, hc, (size_t)ctx->nb, hipMemcpyDeviceToHost, st));
    if (break_step) PXG_HIP(hipMemcpyAsync(break_step, bc, sizeof(int) * (size_t)ctx->nb, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_frac_project_dev(pxg_ctx *ctx, int which, double *out_dev)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_project: no explicit state");
    PXG_CHECK(which >= 0 && which <= 3 && out_dev, PXG_ERR_ARG, "pxg_frac_project: bad arguments");
    PXG_FAULT(ctx, "pxg_frac_project");
    PXG_HIP(hipSetDevice(ctx->device));
    const FracState &F = ctx->frac;
    hipLaunchKernelGGL(ctx->dim == 3 ? k_frac_project<3> : k_frac_project<2>, dim3(gridn(ctx->n, kProjNodes)),
                       dim3(64), 0, ctx->stream, ctx->n, ctx->nop, which, F.plane_stress, F.C12, F.C66,
                       ctx->off.p, ctx->nbr.p, ctx->ops.p, F.health.p, F.bsh.p, F.u[F.cur].p, out_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_frac_project(pxg_ctx *ctx, int which, double *out)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_project: no explicit state");
    PXG_CHECK(which >= 0 && which <= 3 && out, PXG_ERR_ARG, "pxg_frac_project: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    const size_t cnt = (size_t)ctx->n * (which >= 2 ? 9 : 1);
    PXG_HIP(ctx->sf.alloc(cnt));
    int rc = pxg_frac_project_dev(ctx, which, ctx->sf.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(out, ctx->sf.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_frac_layout(pxg_ctx *ctx, long long *info)
{
    PXG_CHECK(ctx && ctx->frac.ready && info, PXG_ERR_ARG, "pxg_frac_layout: bad arguments");
    const FracState &F = ctx->frac;
    info[0] = F.pst;
    info[1] = F.bsh.p ? 1 : 0;
    info[2] = F.nclasses;
    info[3] = F.nshared;
    return PXG_OK;
}

extern "C" int pxg_frac_state_ptrs(pxg_ctx *ctx, double **u_dev, double **uold_dev)
{
    PXG_CHECK(ctx && ctx->frac.ready, PXG_ERR_STATE, "pxg_frac_state_ptrs: no explicit state");
    if (u_dev) *u_dev = ctx->frac.u[ctx->frac.cur].p;
    if (uold_dev) *uold_dev = ctx->frac.u[ctx->frac.cur ^ 1].p;
    return PXG_OK;
}

extern "C" int pxg_frac_profile(pxg_ctx *ctx, int enable)
{
    PXG_CHECK(ctx, PXG_ERR_ARG, "pxg_frac_profile: ctx is NULL");
    ctx->frac.prof = enable ? 1 : 0;
    ctx->frac.nev = 0;
    ctx->frac.nlaunch = 0;
    ctx->frac.ev_open = 0;
    ctx->frac.ev_pending = 0;
    return PXG_OK;
}

extern "C" int pxg_frac_kernel_time(pxg_ctx *ctx, double *ms_total, int *launches)
{
    PXG_CHECK(ctx && ms_total && launches, PXG_ERR_ARG, "pxg_frac_kernel_time: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    FracState &F = ctx->frac;
    double tot = 0.0;
    for (int k = 0; k < F.nev; k++) {
        float ms = 0.0f;
        PXG_HIP(hipEventElapsedTime(&ms, F.ev[2 * k], F.ev[2 * k + 1]));
        tot += ms;
    }
    *ms_total = tot;
    *launches = (int)F.nlaunch;
    return PXG_OK;
}

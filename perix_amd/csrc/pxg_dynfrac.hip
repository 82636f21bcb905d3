// pxg_dynfrac.hip -- the implicit backward-Euler dynamic fracture element
// (PDDODynamicFracElement) on the device: the per-step commit (velocity
// history and irreversible bond failure on the converged u_old,
// PDDODynamicFracElement.cpp:148-204) and its CSR assembly
// (PDDODynamicFracElement.cpp:207-297 through the assembler of
// FormResidualAndJacobian.cpp:25-132).
//
// The assembly runs one wavefront per node (lanes over its bonds); every sum
// of the node's diagonal block and residual is taken in the reference's order
// (nodal term first, then the bonds in family order), so K and the residual
// come out bit for bit.  The commit (one thread per node) reads u_old of the
// neighbours and writes only the node's own velocity, history and bonds.
#include <algorithm>
#include <vector>
#include "pxg_internal.h"

using namespace pxg;

namespace {

unsigned gridd(long long n, int b) { return (unsigned)((n + b - 1) / b); }

template <int DIM>
__global__ void k_dynfrac_commit(int n, const double *coords, const int *off, const int *nbr, double sc,
                                 int tension_only, int damage_on, double dt_prev, const double *uold,
                                 double *uprev, double *vel, unsigned char *health)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t ri = (size_t)i * DIM;
    for (int a = 0; a < DIM; a++) {
        const double uo = uold[ri + a];
        vel[ri + a] = (uo - uprev[ri + a]) / dt_prev;
        uprev[ri + a] = uo;
    }
    if (!(damage_on && sc > 0.0)) return;
    const double xi0 = coords[3 * (size_t)i], yi0 = coords[3 * (size_t)i + 1];
    const double zi0 = (DIM >= 3) ? coords[3 * (size_t)i + 2] : 0.0;
    for (int t = off[i]; t < off[i + 1]; t++) {
        if (!health[t]) continue;
        const int j = nbr[t] - 1;
        const size_t rj = (size_t)j * DIM;
        double xi[3], eta[3];
        xi[0] = coords[3 * (size_t)j] - xi0;
        xi[1] = coords[3 * (size_t)j + 1] - yi0;
        xi[2] = (DIM >= 3) ? coords[3 * (size_t)j + 2] - zi0 : 0.0;
        eta[0] = uold[rj] - uold[ri];
        eta[1] = uold[rj + 1] - uold[ri + 1];
        eta[2] = (DIM >= 3) ? uold[rj + DIM - 1] - uold[ri + DIM - 1] : 0.0;
        const double r = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
        if (r <= 0.0) continue;
        const double yx = xi[0] + eta[0], yy = xi[1] + eta[1], yz = xi[2] + eta[2];
        const double yr = sqrt(yx * yx + yy * yy + yz * yz);
        const double s = (yr - r) / r;
        if (tension_only ? (s > sc) : (fabs(s) > sc)) health[t] = 0;
    }
}

// Assembly, one wavefront per node with lanes over its bonds: lane l
// computes bond off[i] + 64c + l, writes its K_IJ block into the node's CSR
// rows (the 64 lanes write into the same DIM rows, so the stores coalesce)
// and stages its contributions to the diagonal block and the residual in
// LDS; one lane per diagonal-block entry / residual component then adds them
// in family order onto the nodal term, the reference's order.
constexpr int kDynWaves = 4;   // nodes per block

template <int DIM>
__global__ void __launch_bounds__(64 * kDynWaves)
k_dynfrac_assemble(int n, int nop, const int *off, const int *nbr, const double *ops, const int *rowptr,
                   const int *slot, const int *diag, const int *mslot, double C11, double C12, double C66,
                   double rho, double b0, double b1, double b2, double dt, const double *vel,
                   const unsigned char *health, const double *u, const double *uold, double *kvals, double *rhs)
{
    constexpr int NK = DIM * DIM, NS = NK + DIM;
    __shared__ double s_v[kDynWaves][NS][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * kDynWaves;
    const int i = i0 + w;
    const bool active = i < n;
    const int ic = active ? i : 0;   // reads of an idle wave stay in bounds
    int nch = 0;
#pragma unroll
    for (int q = 0; q < kDynWaves; q++)
        if (i0 + q < n) nch = max(nch, (off[i0 + q + 1] - off[i0 + q] + 63) >> 6);
    const size_t ri = (size_t)ic * DIM;
    const int rb = rowptr[ri];
    int rowoff[DIM];                          // row a starts at rowptr[ri + a]
#pragma unroll
    for (int a = 0; a < DIM; a++) rowoff[a] = rowptr[ri + a] - rb;
    if (active && lane < NK && mslot[i] >= 0) kvals[mslot[i] + rowoff[lane / DIM] + lane % DIM] = 0.0;
    double ui[DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++) ui[a] = u[ri + a];
    // nodal: R = b - c0 (u - uold - v dt), K = -c0 I (PDDODynamicFracElement.cpp:276-297)
    const double c0 = rho / (dt * dt);
    double acc = 0.0;
    if (lane < NK) {
        const int a = lane / DIM, c = lane % DIM;
        acc = 0.0 + (-((a == c) ? -c0 : 0.0));
    } else if (lane < NS) {
        const int a = lane - NK;
        const double bf = (a == 0) ? b0 : (a == 1) ? b1 : b2;
        acc = 0.0 + (bf - c0 * (ui[a] - uold[ri + a] - vel[ri + a] * dt));
    }
    const int t0 = active ? off[i] : 0, t1 = active ? off[i + 1] : 0;
    for (int ch = 0; ch < nch; ch++) {
        const int t = t0 + ch * 64 + lane;
        if (t < t1) {
            const size_t rj = (size_t)(nbr[t] - 1) * DIM;
            const double *op = ops + (size_t)t * nop;
            double R[DIM], P[DIM][DIM];
#pragma unroll
            for (int a = 0; a < DIM; a++) {
                R[a] = 0.0;
#pragma unroll
                for (int c = 0; c < DIM; c++) P[a][c] = 0.0;
            }
            const bool intact = health[t] != 0;
            if (intact) {
                const double mu = 1.0;
                if constexpr (DIM == 3) {
                    const double Gxx = op[3], Gxy = op[4], Gxz = op[5], Gyy = op[6], Gyz = op[7], Gzz = op[8];
                    const double aa = C12 + C66;
                    const double p11 = mu * (C11 * Gxx + C66 * (Gyy + Gzz));
                    const double p22 = mu * (C11 * Gyy + C66 * (Gxx + Gzz));
                    const double p33 = mu * (C11 * Gzz + C66 * (Gxx + Gyy));
                    const double p12 = mu * aa * Gxy, p13 = mu * aa * Gxz, p23 = mu * aa * Gyz;
                    const double dux = u[rj] - ui[0], duy = u[rj + 1] - ui[1], duz = u[rj + 2] - ui[2];
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
                    const double dux = u[rj] - ui[0], duy = u[rj + 1] - ui[1];
                    R[0] = p11 * dux + pxy * duy;
                    R[1] = pxy * dux + p22 * duy;
                    P[0][0] = p11; P[0][1] = pxy; P[1][0] = pxy; P[1][1] = p22;
                }
            }
            // scatter (FormResidualAndJacobian.cpp:119-128): K_II = -P, K_IJ = P;
            // a broken bond's local blocks are zero
            const int s0 = slot[t];
#pragma unroll
            for (int a = 0; a < DIM; a++) {
#pragma unroll
                for (int c = 0; c < DIM; c++) {
                    kvals[s0 + rowoff[a] + c] = 0.0 + (-(intact ? P[a][c] : 0.0));
                    s_v[w][a * DIM + c][lane] = intact ? P[a][c] : 0.0;
                }
                s_v[w][NK + a][lane] = R[a];
            }
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
    if (lane < NK) kvals[diag[i] + rowoff[lane / DIM] + lane % DIM] = acc;
    else if (lane < NS) rhs[ri + lane - NK] = acc;
}

__global__ void k_u8_from_seed(long long nb, const unsigned char *seed, unsigned char *h)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb) h[t] = seed[t];
}

__global__ void k_vel_init(int n, int dim, double v0, double v1, double v2, double *vel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v[3] = {v0, v1, v2};
    for (int a = 0; a < dim; a++) vel[(size_t)i * dim + a] = v[a];
}

} // namespace

extern "C" int pxg_dynfrac_setup(pxg_ctx *ctx, const pxg_frac_params *p, const double *v0,
                                 const double *uold0)
{
    PXG_CHECK(ctx && p && uold0, PXG_ERR_ARG, "pxg_dynfrac_setup: bad arguments");
    PXG_CHECK(ctx->nop > 0 && ctx->order >= 2, PXG_ERR_STATE,
              "pxg_dynfrac_setup: needs order-2 operators (PDDODynamicFracElement.cpp:105-109)");
    PXG_CHECK(ctx->nnz > 0 && ctx->csr_ndof == ctx->dim && p->dim == ctx->dim, PXG_ERR_STATE,
              "pxg_dynfrac_setup: needs a CSR pattern with dim dofs per node");
    PXG_CHECK(p->rho > 0.0, PXG_ERR_ARG, "pxg_dynfrac_setup: rho must be > 0");
    PXG_CHECK(!ctx->fault, PXG_ERR_FAULT, "pxg_dynfrac_setup: injected fault");
    PXG_HIP(hipSetDevice(ctx->device));
    DynFracState &D = ctx->dyn;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    D.C11 = p->C11; D.C12 = p->C12; D.C66 = p->C66; D.rho = p->rho;
    D.b[0] = p->bx; D.b[1] = p->by; D.b[2] = p->bz;
    D.sc = p->sc; D.tension_only = p->tension_only; D.damage_on = p->damage_on;
    PXG_HIP(D.health.alloc((size_t)ctx->nb));
    PXG_HIP(D.vel.alloc(nd));
    PXG_HIP(D.uprev.alloc(nd));
    hipLaunchKernelGGL(k_u8_from_seed, dim3(gridd(ctx->nb, 256)), dim3(256), 0, st, ctx->nb, ctx->seed.p,
                       D.health.p);
    hipLaunchKernelGGL(k_vel_init, dim3(gridd(ctx->n, 256)), dim3(256), 0, st, ctx->n, ctx->dim,
                       v0 ? v0[0] : 0.0, v0 ? v0[1] : 0.0, (v0 && ctx->dim == 3) ? v0[2] : 0.0, D.vel.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipMemcpyAsync(D.uprev.p, uold0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipStreamSynchronize(st));
    D.ready = 1;
    return PXG_OK;
}

extern "C" int pxg_dynfrac_commit_dev(pxg_ctx *ctx, const double *uold_dev, double dt_prev)
{
    PXG_CHECK(ctx && ctx->dyn.ready, PXG_ERR_STATE, "pxg_dynfrac_commit: call pxg_dynfrac_setup first");
    PXG_CHECK(uold_dev && dt_prev > 0.0, PXG_ERR_ARG, "pxg_dynfrac_commit: bad arguments");
    PXG_CHECK(!ctx->fault, PXG_ERR_FAULT, "pxg_dynfrac_commit: injected fault");
    PXG_HIP(hipSetDevice(ctx->device));
    DynFracState &D = ctx->dyn;
    const unsigned g = gridd(ctx->n, 128);
    if (ctx->dim == 3)
        hipLaunchKernelGGL(k_dynfrac_commit<3>, dim3(g), dim3(128), 0, ctx->stream, ctx->n, ctx->coords.p,
                           ctx->off.p, ctx->nbr.p, D.sc, D.tension_only, D.damage_on, dt_prev, uold_dev,
                           D.uprev.p, D.vel.p, D.health.p);
    else
        hipLaunchKernelGGL(k_dynfrac_commit<2>, dim3(g), dim3(128), 0, ctx->stream, ctx->n, ctx->coords.p,
                           ctx->off.p, ctx->nbr.p, D.sc, D.tension_only, D.damage_on, dt_prev, uold_dev,
                           D.uprev.p, D.vel.p, D.health.p);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_dynfrac_commit(pxg_ctx *ctx, const double *uold, double dt_prev)
{
    PXG_CHECK(ctx && ctx->dyn.ready, PXG_ERR_STATE, "pxg_dynfrac_commit: call pxg_dynfrac_setup first");
    PXG_CHECK(uold, PXG_ERR_ARG, "pxg_dynfrac_commit: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    const size_t nd = (size_t)ctx->n * ctx->dim;
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, ctx->stream));
    int rc = pxg_dynfrac_commit_dev(ctx, ctx->suo.p, dt_prev);
    if (rc) return rc;
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

extern "C" int pxg_assemble_dynfrac_dev(pxg_ctx *ctx, double dt, const double *u_dev, const double *uold_dev,
                                        double *kvals_dev, double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->dyn.ready, PXG_ERR_STATE, "pxg_assemble_dynfrac: call pxg_dynfrac_setup first");
    PXG_CHECK(ctx->csr_ndof == ctx->dim, PXG_ERR_STATE, "pxg_assemble_dynfrac: CSR pattern changed");
    PXG_CHECK(u_dev && uold_dev && kvals_dev && rhs_dev && dt > 0.0, PXG_ERR_ARG,
              "pxg_assemble_dynfrac: bad arguments (transient element, dt > 0)");
    PXG_CHECK(!ctx->fault, PXG_ERR_FAULT, "pxg_assemble_dynfrac: injected fault");
    PXG_HIP(hipSetDevice(ctx->device));
    const DynFracState &D = ctx->dyn;
    const unsigned g = gridd(ctx->n, kDynWaves);
    if (ctx->dim == 3)
        hipLaunchKernelGGL(k_dynfrac_assemble<3>, dim3(g), dim3(64 * kDynWaves), 0, ctx->stream, ctx->n, ctx->nop,
                           ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->rowptr.p, ctx->slot.p, ctx->diag.p,
                           ctx->mslot.p, D.C11, D.C12, D.C66, D.rho, D.b[0], D.b[1], D.b[2], dt, D.vel.p,
                           D.health.p, u_dev, uold_dev, kvals_dev, rhs_dev);
    else
        hipLaunchKernelGGL(k_dynfrac_assemble<2>, dim3(g), dim3(64 * kDynWaves), 0, ctx->stream, ctx->n, ctx->nop,
                           ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->rowptr.p, ctx->slot.p, ctx->diag.p,
                           ctx->mslot.p, D.C11, D.C12, D.C66, D.rho, D.b[0], D.b[1], D.b[2], dt, D.vel.p,
                           D.health.p, u_dev, uold_dev, kvals_dev, rhs_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_assemble_dynfrac(pxg_ctx *ctx, double dt, const double *u, const double *uold, double *kvals,
                                    double *rhs)
{
    PXG_CHECK(ctx && ctx->dyn.ready, PXG_ERR_STATE, "pxg_assemble_dynfrac: call pxg_dynfrac_setup first");
    PXG_CHECK(u && uold && kvals && rhs, PXG_ERR_ARG, "pxg_assemble_dynfrac: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim, nnz = (size_t)ctx->nnz;
    PXG_HIP(ctx->su.alloc(nd));
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(ctx->sk.alloc(nnz));
    PXG_HIP(ctx->sr.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    int rc = pxg_assemble_dynfrac_dev(ctx, dt, ctx->su.p, ctx->suo.p, ctx->sk.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_dynfrac_get_state(pxg_ctx *ctx, unsigned char *health, double *vel)
{
    PXG_CHECK(ctx && ctx->dyn.ready, PXG_ERR_STATE, "pxg_dynfrac_get_state: call pxg_dynfrac_setup first");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (health) PXG_HIP(hipMemcpyAsync(health, ctx->dyn.health.p, (size_t)ctx->nb, hipMemcpyDeviceToHost, st));
    if (vel)
        PXG_HIP(hipMemcpyAsync(vel, ctx->dyn.vel.p, sizeof(double) * (size_t)ctx->n * ctx->dim,
                               hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

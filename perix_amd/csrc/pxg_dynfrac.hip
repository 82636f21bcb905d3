// pxg_dynfrac.hip -- the implicit backward-Euler dynamic fracture element
// (PDDODynamicFracElement) on the device: the per-step commit (velocity
// history and irreversible bond failure on the converged u_old,
// PDDODynamicFracElement.cpp:148-204) and its CSR assembly
// (PDDODynamicFracElement.cpp:207-297 through the assembler of
// FormResidualAndJacobian.cpp:25-132).
//
// The assembly runs on the block-element kernels of pxg_stressch.hip
// (launch_dyn_rows: lanes over a node's bonds, every sum of the node's
// diagonal block and residual taken in the reference's order, so K and the
// residual come out bit for bit).  The commit (one thread per node) reads
// u_old of the neighbours and writes only the node's own velocity, history
// and bonds.
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
    PXG_FAULT(ctx, "pxg_dynfrac_setup");
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
    PXG_FAULT(ctx, "pxg_dynfrac_commit");
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
    PXG_FAULT(ctx, "pxg_assemble_dynfrac");
    PXG_HIP(hipSetDevice(ctx->device));
    return launch_dyn_rows(ctx, dt, u_dev, uold_dev, kvals_dev, rhs_dev);
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

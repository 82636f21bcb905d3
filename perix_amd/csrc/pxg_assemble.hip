// pxg_assemble.hip -- CSR residual/Jacobian assembly for the Poisson and
// diffusion elements, and the Cahn-Hilliard element's bond geometry and
// entry points (its assembly runs on the block-element kernels of
// pxg_stressch.hip, launch_ch_rows).
//
// Reference: PDSystem::formResidualAndJacobian (FormResidualAndJacobian.cpp:
// 25-132) with PoissonElement (PoissonElement.cpp:26-58) and
// CahnHilliardElement (CahnHilliardElement.cpp:66-292); device twins
// assembleKernel / cahnHilliardKernel (CudaAssembleKernels.cu:291-398,
// 572-649) are the behaviour spec.
//
// Poisson / diffusion rows are node-owned, so no atomics: a 256-thread block owns T nodes,
// lanes stream the tile's bonds (contiguous CSR order), write the
// off-diagonal blocks straight to their precomputed CSR slots (built once
// with the pattern, replacing the per-entry binary search of
// SparseMatrix::addValue) and park the diagonal / residual contributions in
// LDS; one lane per node then adds them in family order, reproducing the
// reference's accumulation exactly (K starts at 0 and receives the nodal
// block first, then each bond).
#include "pxg_internal.h"
#include <cmath>
#include <mutex>

namespace pxg {

constexpr int AT = 32;        // nodes per tile
constexpr int AK = 4;         // bonds per lane per chunk
constexpr int ACH = 256 * AK;

__device__ __forceinline__ double pd_lap(const double *op, int dim)
{
    // ElementBase::pdLaplacian (ElementBase.h:317-321)
    if (dim >= 3) return op[3] + op[6] + op[8];
    return op[2] + op[4];
}

__global__ void k_lap(long long nb, int dim, int nop, const double *ops, double *lap)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    lap[t] = pd_lap(ops + (size_t)t * nop, dim);
}

// Poisson (PoissonElement.cpp:26-58: coeff = sigma lap, nodal R = f, K = 0)
// and, with uold, diffusion (DiffusionElement.cpp:28-62: coeff = -D lap,
// nodal R = (u_i - uold_i)/dt - f, K = 1/dt) through the assembler's
// accumulation order: nodal term first, then the bonds in family order.
__global__ void __launch_bounds__(256) k_poisson(int n, const int *off, const int *nbr, const double *lap,
                                                 const int *slot, const int *diag, const int *mslot,
                                                 double sigma, const double *fnode, double fconst,
                                                 double dt, const double *uold,
                                                 const double *u, double *kvals, double *rhs)
{
    __shared__ int s_off[AT + 1];
    __shared__ double s_u[AT];
    __shared__ double s_c[ACH], s_r[ACH];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * AT;
    const int nt = min(AT, n - i0);
    if (tid <= nt) s_off[tid] = off[i0 + tid];
    if (tid < nt) s_u[tid] = u[i0 + tid];
    __syncthreads();
    const int bt0 = s_off[0], bt1 = s_off[nt];
    double kd = 0.0, r = 0.0;
    if (tid < nt) {
        const int i = i0 + tid;
        const double fv = fnode ? fnode[i] : fconst;
        if (uold) {
            r = r + ((s_u[tid] - uold[i]) / dt - fv);   // nodal R = (u - uold)/dt - f
            kd = kd + (-(1.0 / dt));                   // nodal K = 1/dt
        } else {
            r = r + fv;                                // nodal R = f
            kd = kd + (-0.0);                          // nodal K = 0
        }
        if (mslot[i] >= 0) kvals[mslot[i]] = 0.0;
    }
    for (int cs = bt0; cs < bt1; cs += ACH) {
        const int ce = min(bt1, cs + ACH);
        for (int k = 0; k < AK; k++) {
            const int lt = tid + k * 256, t = cs + lt;
            if (t >= ce) continue;
            int lo = 0, hi = nt - 1;
            while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_off[mid] <= t) lo = mid; else hi = mid - 1; }
            const int j = nbr[t] - 1;
            const double coeff = sigma * lap[t];
            s_c[lt] = coeff;
            s_r[lt] = coeff * (u[j] - s_u[lo]);
            kvals[slot[t]] = 0.0 + (-coeff);
        }
        __syncthreads();
        if (tid < nt) {
            const int tb = max(s_off[tid], cs), te = min(s_off[tid + 1], ce);
            for (int t = tb; t < te; t++) { r += s_r[t - cs]; kd += -(-s_c[t - cs]); }
        }
        __syncthreads();
    }
    if (tid < nt) {
        const int i = i0 + tid;
        rhs[i] = r;
        kvals[diag[i]] = kd;
    }
}

// ---- Cahn-Hilliard -------------------------------------------------------
__global__ void k_ch_geometry(int n, int dim, int nop, const double *coords, const int *off,
                              const int *nbr, const double *ops, const int *mirror, double *lap,
                              int *rev, unsigned char *isghost)
{
    // CahnHilliardElement.cpp:100-152
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi0 = coords[3 * (size_t)i], yi0 = coords[3 * (size_t)i + 1];
    const double zi0 = (dim >= 3) ? coords[3 * (size_t)i + 2] : 0.0;
    const double qi = xi0 * xi0 + yi0 * yi0 + zi0 * zi0;
    double lapq = 0.0;
    for (int t = off[i]; t < off[i + 1]; t++) {
        const int j = nbr[t] - 1;
        const double w = pd_lap(ops + (size_t)t * nop, dim);
        lap[t] = w;
        if (!(mirror[j] > 0)) {
            const double xj = coords[3 * (size_t)j], yj = coords[3 * (size_t)j + 1];
            const double zj = (dim >= 3) ? coords[3 * (size_t)j + 2] : 0.0;
            lapq += w * (xj * xj + yj * yj + zj * zj - qi);
        }
    }
    const double lam = (lapq > 1.0e-12) ? (2.0 * (double)dim / lapq) : 1.0;
    for (int t = off[i]; t < off[i + 1]; t++)
        if (!(mirror[nbr[t] - 1] > 0)) lap[t] *= lam;
    for (int t = off[i]; t < off[i + 1]; t++) {
        const int j = nbr[t] - 1;
        int r = -1;
        for (int s = off[j]; s < off[j + 1]; s++) if (nbr[s] == i + 1) { r = s; break; }
        rev[t] = r;
    }
    isghost[i] = (mirror[i] > 0) ? 1 : 0;
}

} // namespace pxg

using namespace pxg;

static unsigned gridb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// ctx->lap is shared: the plain pdLaplacian of the Poisson and diffusion
// assemblies, or the lambda-corrected one of the Cahn-Hilliard geometry.  It
// is rebuilt unless it holds the plain Laplacian of the current operators.
static int ensure_lap_poisson(pxg_ctx *ctx)
{
    if (ctx->lap_kind == LAP_POISSON && ctx->lap_gen == ctx->ops_gen && ctx->lap.n == (size_t)ctx->nb && ctx->lap.p)
        return PXG_OK;
    PXG_HIP(ctx->lap.alloc(ctx->nb));
    hipLaunchKernelGGL(k_lap, dim3(gridb(ctx->nb, 256)), dim3(256), 0, ctx->stream, ctx->nb, ctx->dim, ctx->nop,
                       ctx->ops.p, ctx->lap.p);
    PXG_HIP(hipGetLastError());
    ctx->lap_kind = LAP_POISSON;
    ctx->lap_gen = ctx->ops_gen;
    ctx->ch_ready = 0;          // the Cahn-Hilliard geometry was overwritten
    return PXG_OK;
}

extern "C" int pxg_assemble_poisson_dev(pxg_ctx *ctx, double sigma, const double *f_node_dev, double f_const,
                                        const double *u_dev, double *kvals_dev, double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->nop > 0 && ctx->nnz > 0 && ctx->csr_ndof == 1, PXG_ERR_STATE,
              "pxg_assemble_poisson: needs operators and a 1-dof CSR pattern");
    PXG_CHECK(u_dev && kvals_dev && rhs_dev, PXG_ERR_ARG, "pxg_assemble_poisson: bad arguments");
    PXG_FAULT(ctx, "pxg_assemble_poisson");
    PXG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_lap_poisson(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_poisson, dim3(gridb(ctx->n, AT)), dim3(256), 0, ctx->stream, ctx->n, ctx->off.p,
                       ctx->nbr.p, ctx->lap.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, sigma, f_node_dev,
                       f_const, 0.0, (const double *)nullptr, u_dev, kvals_dev, rhs_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_assemble_poisson(pxg_ctx *ctx, double sigma, const double *f_node, double f_const,
                                    const double *u, double *kvals, double *rhs)
{
    PXG_CHECK(ctx && ctx->nnz > 0 && u && kvals && rhs, PXG_ERR_ARG, "pxg_assemble_poisson: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PXG_HIP(ctx->su.alloc(ctx->n));
    PXG_HIP(ctx->sk.alloc(ctx->nnz));
    PXG_HIP(ctx->sr.alloc(ctx->n));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
    if (f_node) {
        PXG_HIP(ctx->sf.alloc(ctx->n));
        PXG_HIP(hipMemcpyAsync(ctx->sf.p, f_node, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
    }
    int rc = pxg_assemble_poisson_dev(ctx, sigma, f_node ? ctx->sf.p : nullptr, f_const, ctx->su.p, ctx->sk.p,
                                      ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

// ---- allocation-free Poisson assembly -------------------------------------
// pxg_assemble_poisson_noalloc keeps its one scratch buffer, the per-bond
// Laplacian, in this file-scope device array instead of the context's lap
// buffer (which the first Poisson call allocates).  Every GPU the library
// runs on holds its own copy of the array for the life of the process; host
// code reaches it only through the per-device address taken below, never
// through the symbol's name (the host-side name is a stand-in, not a device
// address).
constexpr size_t kScratchDoubles = (size_t)2 << 20;     // 16 MiB: 2,097,152 bonds
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static double *g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

// Called by pxg_create with `device` current: takes this device's address
// of the scratch array once per device.
int pxg::scratch_init(int device)
{
    PXG_CHECK(device >= 0 && device < kScratchDevices, PXG_ERR_ARG, "scratch_init: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return PXG_OK;
    void *p = nullptr;
    PXG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PXG_CHECK(p, PXG_ERR_HIP, "scratch_init: no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double *>(p);
    return PXG_OK;
}

extern "C" int pxg_assemble_poisson_noalloc(pxg_ctx *ctx, double sigma, const double *f_node_dev,
                                            double f_const, const double *u_dev, double *kvals_dev,
                                            double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->nop > 0 && ctx->nnz > 0 && ctx->csr_ndof == 1, PXG_ERR_STATE,
              "pxg_assemble_poisson_noalloc: needs operators and a 1-dof CSR pattern");
    PXG_CHECK(u_dev && kvals_dev && rhs_dev, PXG_ERR_ARG, "pxg_assemble_poisson_noalloc: bad arguments");
    PXG_CHECK((size_t)ctx->nb <= kScratchDoubles, PXG_ERR_ARG,
              "pxg_assemble_poisson_noalloc: " + std::to_string(ctx->nb) +
                  " directed bonds exceed the scratch limit of 2097152 (use pxg_assemble_poisson_dev)");
    PXG_FAULT(ctx, "pxg_assemble_poisson_noalloc");
    PXG_HIP(hipSetDevice(ctx->device));
    double *lap = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        lap = (ctx->device >= 0 && ctx->device < kScratchDevices) ? g_scratch_ptr[ctx->device] : nullptr;
    }
    PXG_CHECK(lap, PXG_ERR_STATE, "pxg_assemble_poisson_noalloc: scratch not set up on this device");
    // k_lap writes every one of the nb entries k_poisson reads: nothing to clear
    hipLaunchKernelGGL(k_lap, dim3(gridb(ctx->nb, 256)), dim3(256), 0, ctx->stream, ctx->nb, ctx->dim, ctx->nop,
                       ctx->ops.p, lap);
    hipLaunchKernelGGL(k_poisson, dim3(gridb(ctx->n, AT)), dim3(256), 0, ctx->stream, ctx->n, ctx->off.p,
                       ctx->nbr.p, lap, ctx->slot.p, ctx->diag.p, ctx->mslot.p, sigma, f_node_dev, f_const, 0.0,
                       (const double *)nullptr, u_dev, kvals_dev, rhs_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_assemble_diffusion_dev(pxg_ctx *ctx, double D, const double *f_node_dev, double f_const,
                                          double dt, const double *u_dev, const double *uold_dev,
                                          double *kvals_dev, double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->nop > 0 && ctx->nnz > 0 && ctx->csr_ndof == 1, PXG_ERR_STATE,
              "pxg_assemble_diffusion: needs operators and a 1-dof CSR pattern");
    PXG_CHECK(u_dev && uold_dev && kvals_dev && rhs_dev && dt > 0.0, PXG_ERR_ARG,
              "pxg_assemble_diffusion: bad arguments (the element is transient-only, dt > 0)");
    PXG_FAULT(ctx, "pxg_assemble_diffusion");
    PXG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_lap_poisson(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_poisson, dim3(gridb(ctx->n, AT)), dim3(256), 0, ctx->stream, ctx->n, ctx->off.p,
                       ctx->nbr.p, ctx->lap.p, ctx->slot.p, ctx->diag.p, ctx->mslot.p, -D, f_node_dev,
                       f_const, dt, uold_dev, u_dev, kvals_dev, rhs_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_assemble_diffusion(pxg_ctx *ctx, double D, const double *f_node, double f_const, double dt,
                                      const double *u, const double *uold, double *kvals, double *rhs)
{
    PXG_CHECK(ctx && ctx->nnz > 0 && u && uold && kvals && rhs, PXG_ERR_ARG,
              "pxg_assemble_diffusion: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PXG_HIP(ctx->su.alloc(ctx->n));
    PXG_HIP(ctx->suo.alloc(ctx->n));
    PXG_HIP(ctx->sk.alloc(ctx->nnz));
    PXG_HIP(ctx->sr.alloc(ctx->n));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
    if (f_node) {
        PXG_HIP(ctx->sf.alloc(ctx->n));
        PXG_HIP(hipMemcpyAsync(ctx->sf.p, f_node, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
    }
    int rc = pxg_assemble_diffusion_dev(ctx, D, f_node ? ctx->sf.p : nullptr, f_const, dt, ctx->su.p, ctx->suo.p,
                                        ctx->sk.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

int pxg::ch_geometry(pxg_ctx *ctx)
{
    PXG_FAULT(ctx, "pxg_ch_setup");
    PXG_HIP(hipSetDevice(ctx->device));
    PXG_HIP(ctx->lap.alloc(ctx->nb));
    PXG_HIP(ctx->rev.alloc(ctx->nb));
    PXG_HIP(ctx->isghost.alloc(ctx->n));
    hipLaunchKernelGGL(k_ch_geometry, dim3(gridb(ctx->n, 128)), dim3(128), 0, ctx->stream, ctx->n, ctx->dim,
                       ctx->nop, ctx->coords.p, ctx->off.p, ctx->nbr.p, ctx->ops.p, ctx->mirror.p, ctx->lap.p,
                       ctx->rev.p, ctx->isghost.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->lap_kind = LAP_CH;
    ctx->lap_gen = ctx->ops_gen;
    ctx->ch_ready = 1;
    return PXG_OK;
}

extern "C" int pxg_ch_setup(pxg_ctx *ctx)
{
    PXG_CHECK(ctx && ctx->nop > 0 && ctx->nnz > 0 && ctx->csr_ndof == 2, PXG_ERR_STATE,
              "pxg_ch_setup: needs operators and a 2-dof CSR pattern");
    return ch_geometry(ctx);
}

extern "C" int pxg_ch_set_geometry(pxg_ctx *ctx, const double *lapbond, const int *rev,
                                   const unsigned char *isghost)
{
    PXG_CHECK(ctx && ctx->nb > 0, PXG_ERR_STATE, "pxg_ch_set_geometry: no families");
    PXG_CHECK(lapbond && rev, PXG_ERR_ARG, "pxg_ch_set_geometry: bad arguments");
    for (long long t = 0; t < ctx->nb; t++)
        PXG_CHECK(rev[t] >= -1 && rev[t] < ctx->nb, PXG_ERR_ARG, "pxg_ch_set_geometry: reverse bond out of range");
    PXG_FAULT(ctx, "pxg_ch_set_geometry");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PXG_HIP(ctx->lap.alloc(ctx->nb));
    PXG_HIP(ctx->rev.alloc(ctx->nb));
    PXG_HIP(ctx->isghost.alloc(ctx->n));
    PXG_HIP(hipMemcpyAsync(ctx->lap.p, lapbond, sizeof(double) * ctx->nb, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->rev.p, rev, sizeof(int) * ctx->nb, hipMemcpyHostToDevice, st));
    if (isghost) PXG_HIP(hipMemcpyAsync(ctx->isghost.p, isghost, ctx->n, hipMemcpyHostToDevice, st));
    else PXG_HIP(hipMemsetAsync(ctx->isghost.p, 0, ctx->n, st));
    PXG_HIP(hipStreamSynchronize(st));
    ctx->lap_kind = LAP_CH;
    ctx->lap_gen = ctx->ops_gen;
    ctx->ch_ready = 1;
    return PXG_OK;
}

extern "C" int pxg_ch_set_mobility(pxg_ctx *ctx, const double *mnode)
{
    PXG_CHECK(ctx && ctx->n > 0, PXG_ERR_STATE, "pxg_ch_set_mobility: no points");
    if (!mnode) { ctx->mob_set = 0; return PXG_OK; }
    PXG_FAULT(ctx, "pxg_ch_set_mobility");
    PXG_HIP(hipSetDevice(ctx->device));
    PXG_HIP(ctx->mob.alloc(ctx->n));
    PXG_HIP(hipMemcpyAsync(ctx->mob.p, mnode, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->mob_set = 1;
    return PXG_OK;
}

extern "C" int pxg_ch_get_geometry(pxg_ctx *ctx, double *lapbond, int *rev, unsigned char *isghost)
{
    PXG_CHECK(ctx && ctx->ch_ready, PXG_ERR_STATE, "pxg_ch_get_geometry: call pxg_ch_setup first");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (lapbond) PXG_HIP(hipMemcpyAsync(lapbond, ctx->lap.p, sizeof(double) * ctx->nb, hipMemcpyDeviceToHost, st));
    if (rev) PXG_HIP(hipMemcpyAsync(rev, ctx->rev.p, sizeof(int) * ctx->nb, hipMemcpyDeviceToHost, st));
    if (isghost) PXG_HIP(hipMemcpyAsync(isghost, ctx->isghost.p, ctx->n, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_assemble_ch_dev(pxg_ctx *ctx, double kappa, double chi, double M0, double dt,
                                   const double *u_dev, const double *uold_dev, double *kvals_dev,
                                   double *rhs_dev)
{
    PXG_CHECK(ctx && ctx->ch_ready, PXG_ERR_STATE, "pxg_assemble_ch: call pxg_ch_setup first");
    PXG_CHECK(ctx->csr_ndof == 2, PXG_ERR_STATE, "pxg_assemble_ch: needs the 2-dof CSR pattern");
    PXG_CHECK(dt > 0.0, PXG_ERR_ARG, "pxg_assemble_ch: Cahn-Hilliard is transient-only (dt > 0)");
    PXG_CHECK(u_dev && uold_dev && kvals_dev && rhs_dev, PXG_ERR_ARG, "pxg_assemble_ch: bad arguments");
    PXG_FAULT(ctx, "pxg_assemble_ch");
    PXG_HIP(hipSetDevice(ctx->device));
    return launch_ch_rows(ctx, kappa, chi, M0, dt, u_dev, uold_dev, kvals_dev, rhs_dev);
}

extern "C" int pxg_assemble_ch(pxg_ctx *ctx, double kappa, double chi, double M0, double dt, const double *u,
                               const double *uold, double *kvals, double *rhs)
{
    PXG_CHECK(ctx && ctx->ch_ready && u && uold && kvals && rhs, PXG_ERR_ARG, "pxg_assemble_ch: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * 2;
    PXG_HIP(ctx->su.alloc(nd));
    PXG_HIP(ctx->suo.alloc(nd));
    PXG_HIP(ctx->sk.alloc(ctx->nnz));
    PXG_HIP(ctx->sr.alloc(nd));
    PXG_HIP(hipMemcpyAsync(ctx->su.p, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->suo.p, uold, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    int rc = pxg_assemble_ch_dev(ctx, kappa, chi, M0, dt, ctx->su.p, ctx->suo.p, ctx->sk.p, ctx->sr.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(kvals, ctx->sk.p, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(rhs, ctx->sr.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

// pxg_adjoint.hip -- reverse-mode derivative of the explicit fracture loop
// with respect to a per-node Young's modulus.  The reference has no adjoint;
// the forward map differentiated here is the explicit step of
// ExplicitPDDOFracElement.cpp:237-299 + TimeStepping.cpp:479-547 with bond
// health frozen from the recorded forward break steps:
//
//   u^{-1} = u^0 - dt v0 + dt^2/2 a(u^0; h^0)
//   w^k    = 2u^{k-1} - u^{k-2} + dt^2 a(u^{k-1}; h^k),   u^k = S w^k + g(t_k)
//   a_i    = (E_i (Khat_k u)_i + b) / rho,
//   (Khat_k u)_i = sum_j h^k_ij Phat_ij (u_j - u_i),  Phat from C(E = 1)
//   J      = sum_s (u^K_{s,comp})^2
//
// (C is linear in E, so dC/dE_i = C(E=1).)  Reverse sweep, k = K..1:
//   wbar   = S^T lam_k
//   Ebar_i += c (Khat_k u^{k-1})_i . wbar_i,            c = dt^2/rho
//   lam_{k-1} = lamp + 2 wbar + c Khat_k^T (E .* wbar),   lamp <- -wbar
// and finally Ebar_i += c/2 (Khat_0 u^0)_i . lam_{-1}.  Khat^T is applied in
// gather form through the reverse-bond index (families are symmetric), so
// every node writes only its own entries: no atomics.  Displacements are
// checkpointed every `ckpt` steps and each segment is recomputed forward
// before its reverse sweep.
#include "pxg_internal.h"
#include <climits>
#include <cmath>

namespace pxg {

constexpr int JT = 32;      // nodes per tile (2D); 3D uses 8
constexpr int JK = 4;

// wbar = S^T lam for the merged BC program
__global__ void k_preset_T(int n, int dim, const double *lam, const unsigned char *mask,
                           const int *bnode, const int *bmode, const int *inv_off, const int *inv_q,
                           const int *tslot, double *wbar)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    for (int c = 0; c < dim; c++) {
        double v = lam[(size_t)m * dim + c];
        if (mask[m]) {
            const int q = tslot[m];
            if (bmode[q * 3 + c] != 0) v = 0.0;
        }
        for (int p = inv_off[m]; p < inv_off[m + 1]; p++) {
            const int q = inv_q[p];
            const int md = bmode[q * 3 + c];
            const int g = bnode[q];
            if (md == 1) v += lam[(size_t)g * dim + c];
            else if (md == 2) v += -lam[(size_t)g * dim + c];
        }
        wbar[(size_t)m * dim + c] = v;
    }
}

// P(u) product with a symmetric per-bond operator (3 or 6 components)
template <int DIM>
__device__ __forceinline__ void pmul(const double *P, long long ps, long long t, const double *v, double *o)
{
    if (DIM == 2) {
        const double p11 = P[t], p22 = P[ps + t], p12 = P[2 * ps + t];
        o[0] = p11 * v[0] + p12 * v[1];
        o[1] = p12 * v[0] + p22 * v[1];
    } else {
        const double p11 = P[t], p22 = P[ps + t], p33 = P[2 * ps + t];
        const double p12 = P[3 * ps + t], p13 = P[4 * ps + t], p23 = P[5 * ps + t];
        o[0] = p11 * v[0] + p12 * v[1] + p13 * v[2];
        o[1] = p12 * v[0] + p22 * v[1] + p23 * v[2];
        o[2] = p13 * v[0] + p23 * v[1] + p33 * v[2];
    }
}

// One reverse step (LAST = the u^{-1} term only).
template <int DIM, bool LAST>
__global__ void __launch_bounds__(256) k_adj_step(int n, long long ps, const int *off, const int *nbr,
                                                  const int *rev, const int *bs, int k, const double *Phat,
                                                  const double *E, const double *u, const double *wbar,
                                                  double *lam, double *lamp, double *Ebar, double cf)
{
    constexpr int T = (DIM == 2) ? JT : 8;
    constexpr int CH = 256 * JK;
    __shared__ int s_off[T + 1];
    __shared__ double s_u[T][DIM], s_w[T][DIM];
    __shared__ double s_ku[CH][DIM], s_z[CH][DIM];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * T;
    const int nt = min(T, n - i0);
    if (tid <= nt) s_off[tid] = off[i0 + tid];
    if (tid < nt)
        for (int c = 0; c < DIM; c++) {
            s_u[tid][c] = u[(size_t)(i0 + tid) * DIM + c];
            s_w[tid][c] = wbar[(size_t)(i0 + tid) * DIM + c];
        }
    __syncthreads();
    const int bt0 = s_off[0], bt1 = s_off[nt];
    double ku[DIM], z[DIM];
    for (int c = 0; c < DIM; c++) { ku[c] = 0.0; z[c] = 0.0; }
    for (int cs = bt0; cs < bt1; cs += CH) {
        const int ce = min(bt1, cs + CH);
        for (int kk = 0; kk < JK; kk++) {
            const int lt = tid + kk * 256, t = cs + lt;
            if (t >= ce) continue;
            int lo = 0, hi = nt - 1;
            while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_off[mid] <= t) lo = mid; else hi = mid - 1; }
            const int i = i0 + lo;
            const int j = nbr[t] - 1;
            double du[DIM], a[DIM], zz[DIM];
            for (int c = 0; c < DIM; c++) { a[c] = 0.0; zz[c] = 0.0; }
            if (bs[t] > k) {
                for (int c = 0; c < DIM; c++) du[c] = u[(size_t)j * DIM + c] - s_u[lo][c];
                pmul<DIM>(Phat, ps, t, du, a);
                if (!LAST) {
                    double ew[DIM], o[DIM];
                    for (int c = 0; c < DIM; c++) ew[c] = E[i] * s_w[lo][c];
                    pmul<DIM>(Phat, ps, t, ew, o);
                    for (int c = 0; c < DIM; c++) zz[c] -= o[c];
                }
            }
            if (!LAST) {
                const int r = rev[t];
                if (r >= 0 && bs[r] > k) {
                    double ew[DIM], o[DIM];
                    for (int c = 0; c < DIM; c++) ew[c] = E[j] * wbar[(size_t)j * DIM + c];
                    pmul<DIM>(Phat, ps, r, ew, o);
                    for (int c = 0; c < DIM; c++) zz[c] += o[c];
                }
            }
            for (int c = 0; c < DIM; c++) { s_ku[lt][c] = a[c]; s_z[lt][c] = zz[c]; }
        }
        __syncthreads();
        if (tid < nt) {
            const int tb = max(s_off[tid], cs), te = min(s_off[tid + 1], ce);
            for (int t = tb; t < te; t++)
                for (int c = 0; c < DIM; c++) { ku[c] += s_ku[t - cs][c]; z[c] += s_z[t - cs][c]; }
        }
        __syncthreads();
    }
    if (tid < nt) {
        const int i = i0 + tid;
        double s = 0.0;
        for (int c = 0; c < DIM; c++) s += ku[c] * s_w[tid][c];
        Ebar[i] += cf * s;
        if (!LAST) {
            for (int c = 0; c < DIM; c++) {
                const size_t q = (size_t)i * DIM + c;
                const double wb = s_w[tid][c];
                const double np = lamp[q] + 2.0 * wb + cf * z[c];
                lamp[q] = -wb;
                lam[q] = np;
            }
        }
    }
}

__global__ void k_seed_lambda(int nsens, const int *sens, int dim, int comp, const double *uK, double *lam)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int q = 0; q < nsens; q++) {
        const size_t idx = (size_t)sens[q] * dim + comp;
        lam[idx] += 2.0 * uK[idx];
    }
}

__global__ void k_loss(int nsens, const int *sens, int dim, int comp, const double *uK, double *J)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int q = 0; q < nsens; q++) { const double v = uK[(size_t)sens[q] * dim + comp]; s += v * v; }
    *J = s;
}

// frozen-health forward over steps [k0+1, k1] starting from (um = u^{k0-1},
// uc = u^{k0}); store[k-k0] receives u^{k-1} for k in (k0, k1+1] when given.
static int frozen_segment(pxg_ctx *ctx, double *um, double *uc, int k0, int k1, double dt, double t0,
                          double *store, size_t nd)
{
    hipStream_t st = ctx->stream;
    double t = t0;
    if (store) PXG_HIP(hipMemcpyAsync(store, uc, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
    for (int k = k0 + 1; k <= k1; k++) {
        const double tNew = t + dt;
        int rc = launch_step(ctx, 0 /*MODE_CD*/, uc, um, um, dt * dt, 0.0, k, true, k);
        if (rc) return rc;
        rc = launch_preset(ctx, um, tNew, k);
        if (rc) return rc;
        double *tmp = um; um = uc; uc = tmp;
        if (store && k < k1) PXG_HIP(hipMemcpyAsync(store + (size_t)(k - k0) * nd, uc, sizeof(double) * nd,
                                                   hipMemcpyDeviceToDevice, st));
        t = tNew;
    }
    // caller's buffers: make (um, uc) hold (u^{k1-1}, u^{k1}) in the original slots
    return PXG_OK;
}

} // namespace pxg

using namespace pxg;

static unsigned gridn2(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// forward with frozen health from u0: fills checkpoints (pairs u^{k-1}, u^k at
// k = 0, ckpt, 2ckpt, ...) and returns u^K in F.u[...]
static int adj_forward(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int K, int ckpt,
                       dbuf<double> *cks, double **uK_out, double *tk)
{
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    PXG_HIP(hipMemcpyAsync(F.u[0].p, u0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    if (v0) PXG_HIP(hipMemcpyAsync(F.accel.p, v0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    else PXG_HIP(hipMemsetAsync(F.accel.p, 0, sizeof(double) * nd, st));
    PXG_HIP(hipMemsetAsync(F.gmax.p, 0, sizeof(unsigned long long) * 4, st));
    PXG_HIP(hipMemsetAsync(F.gstate.p, 0, sizeof(int) * 4, st));
    int rc = launch_step(ctx, 3 /*MODE_INIT*/, F.u[0].p, F.accel.p, F.u[1].p, dt, 0.5 * dt * dt, 0, false, 0);
    if (rc) return rc;
    double *um = F.u[1].p, *uc = F.u[0].p;
    double t = 0.0;
    for (int k = 1; k <= K; k++) {
        if (cks && (k - 1) % ckpt == 0) {
            const int c = (k - 1) / ckpt;
            PXG_HIP(hipMemcpyAsync(cks->p + (size_t)(2 * c) * nd, um, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
            PXG_HIP(hipMemcpyAsync(cks->p + (size_t)(2 * c + 1) * nd, uc, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
            tk[c] = t;
        }
        const double tNew = t + dt;
        rc = launch_step(ctx, 0, uc, um, um, dt * dt, 0.0, k, true, k);
        if (rc) return rc;
        rc = launch_preset(ctx, um, tNew, k);
        if (rc) return rc;
        double *tmp = um; um = uc; uc = tmp;
        t = tNew;
    }
    *uK_out = uc;
    return PXG_OK;
}

extern "C" int pxg_frac_loss(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int nsteps,
                             const double *node_E, int nsens, const int *sens, int comp, double *J_out)
{
    PXG_CHECK(ctx && ctx->frac.ready && ctx->frac.has_nodeE, PXG_ERR_STATE,
              "pxg_frac_loss: needs pxg_frac_setup with a per-node E");
    PXG_CHECK(u0 && dt > 0.0 && nsteps >= 1 && nsens >= 0 && (nsens == 0 || sens) && comp >= 0 &&
              comp < ctx->dim && J_out, PXG_ERR_ARG, "pxg_frac_loss: bad arguments");
    PXG_CHECK(!ctx->fault, PXG_ERR_FAULT, "pxg_frac_loss: injected fault");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    dbuf<double> eo;
    dbuf<int> ds;
    PXG_HIP(ds.alloc(nsens > 0 ? nsens : 1));
    if (nsens) PXG_HIP(hipMemcpyAsync(ds.p, sens, sizeof(int) * nsens, hipMemcpyHostToDevice, st));
    if (node_E) {
        PXG_HIP(eo.alloc(ctx->n));
        PXG_HIP(hipMemcpyAsync(eo.p, node_E, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
        int rc = launch_frac_ops(ctx, eo.p);
        if (rc) return rc;
    }
    double *uK = nullptr;
    int rc = adj_forward(ctx, u0, v0, dt, nsteps, 1, nullptr, &uK, nullptr);
    if (rc) return rc;
    dbuf<double> dJ;
    PXG_HIP(dJ.alloc(1));
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(1), 0, st, nsens, ds.p, ctx->dim, comp, uK, dJ.p);
    PXG_HIP(hipMemcpyAsync(J_out, dJ.p, sizeof(double), hipMemcpyDeviceToHost, st));
    if (node_E) { rc = launch_frac_ops(ctx, F.nodeE.p); if (rc) return rc; }
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_frac_adjoint(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int nsteps,
                                int ckpt, int nsens, const int *sens, int comp, double *J_out, double *grad)
{
    PXG_CHECK(ctx && ctx->frac.ready && ctx->frac.has_nodeE, PXG_ERR_STATE,
              "pxg_frac_adjoint: needs pxg_frac_setup with a per-node E (and a recorded forward run)");
    PXG_CHECK(u0 && dt > 0.0 && nsteps >= 1 && ckpt >= 1 && nsens >= 0 && (nsens == 0 || sens) &&
              comp >= 0 && comp < ctx->dim && grad, PXG_ERR_ARG, "pxg_frac_adjoint: bad arguments");
    PXG_CHECK(!ctx->fault, PXG_ERR_FAULT, "pxg_frac_adjoint: injected fault");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    hipStream_t st = ctx->stream;
    const int n = ctx->n, dim = ctx->dim, K = nsteps;
    const size_t nd = (size_t)n * dim;
    const int ncks = (K - 1) / ckpt + 1;
    dbuf<double> cks, seg, lam, lamp, wbar, Ebar, dJ;
    dbuf<int> ds, tslot;
    PXG_HIP(cks.alloc((size_t)2 * ncks * nd));
    PXG_HIP(seg.alloc((size_t)(ckpt + 1) * nd));
    PXG_HIP(lam.alloc(nd)); PXG_HIP(lamp.alloc(nd)); PXG_HIP(wbar.alloc(nd));
    PXG_HIP(Ebar.alloc(n)); PXG_HIP(dJ.alloc(1));
    PXG_HIP(ds.alloc(nsens > 0 ? nsens : 1));
    if (nsens) PXG_HIP(hipMemcpyAsync(ds.p, sens, sizeof(int) * nsens, hipMemcpyHostToDevice, st));
    // node -> BC program slot
    {
        std::vector<int> node(F.bc.count > 0 ? F.bc.count : 1), slot(n, 0);
        if (F.bc.count) PXG_HIP(hipMemcpy(node.data(), F.bc.node.p, sizeof(int) * F.bc.count, hipMemcpyDeviceToHost));
        for (int q = 0; q < F.bc.count; q++) slot[node[q]] = q;
        PXG_HIP(tslot.alloc(n));
        // synchronous: `slot` is released at the end of this block
        PXG_HIP(hipMemcpy(tslot.p, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    std::vector<double> tk(ncks, 0.0);
    double *uK = nullptr;
    int rc = adj_forward(ctx, u0, v0, dt, K, ckpt, &cks, &uK, tk.data());
    if (rc) return rc;
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(1), 0, st, nsens, ds.p, dim, comp, uK, dJ.p);
    PXG_HIP(hipMemsetAsync(lam.p, 0, sizeof(double) * nd, st));
    PXG_HIP(hipMemsetAsync(lamp.p, 0, sizeof(double) * nd, st));
    PXG_HIP(hipMemsetAsync(Ebar.p, 0, sizeof(double) * n, st));
    hipLaunchKernelGGL(k_seed_lambda, dim3(1), dim3(1), 0, st, nsens, ds.p, dim, comp, uK, lam.p);
    const double c = dt * dt / F.rho;
    const int T = (dim == 2) ? JT : 8;
    dbuf<double> wa, wb;
    PXG_HIP(wa.alloc(nd)); PXG_HIP(wb.alloc(nd));
    for (int cidx = ncks - 1; cidx >= 0; cidx--) {
        const int k0 = cidx * ckpt;                 // segment covers steps k0+1 .. k1
        const int k1 = (k0 + ckpt < K) ? k0 + ckpt : K;
        PXG_HIP(hipMemcpyAsync(wa.p, cks.p + (size_t)(2 * cidx) * nd, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
        PXG_HIP(hipMemcpyAsync(wb.p, cks.p + (size_t)(2 * cidx + 1) * nd, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
        // seg[k-1-k0] = u^{k-1} for k = k0+1 .. k1
        rc = frozen_segment(ctx, wa.p, wb.p, k0, k1, dt, tk[cidx], seg.p, nd);
        if (rc) return rc;
        for (int k = k1; k >= k0 + 1; k--) {
            const double *ukm1 = seg.p + (size_t)(k - 1 - k0) * nd;
            hipLaunchKernelGGL(k_preset_T, dim3(gridn2(n, 256)), dim3(256), 0, st, n, dim, lam.p,
                               F.preset_mask.p, F.bc.node.p, F.bc.mode.p, F.bc.inv_off.p, F.bc.inv_q.p,
                               tslot.p, wbar.p);
            if (dim == 2)
                hipLaunchKernelGGL((k_adj_step<2, false>), dim3(gridn2(n, T)), dim3(256), 0, st, n, F.ps,
                                   ctx->off.p, ctx->nbr.p, F.rev.p, F.break_step.p, k, F.Phat.p, F.nodeE.p,
                                   ukm1, wbar.p, lam.p, lamp.p, Ebar.p, c);
            else
                hipLaunchKernelGGL((k_adj_step<3, false>), dim3(gridn2(n, T)), dim3(256), 0, st, n, F.ps,
                                   ctx->off.p, ctx->nbr.p, F.rev.p, F.break_step.p, k, F.Phat.p, F.nodeE.p,
                                   ukm1, wbar.p, lam.p, lamp.p, Ebar.p, c);
            PXG_HIP(hipGetLastError());
        }
    }
    // u^{-1} term with h^0 at u^0 (checkpoint 0 holds u^{-1}, u^0)
    if (dim == 2)
        hipLaunchKernelGGL((k_adj_step<2, true>), dim3(gridn2(n, T)), dim3(256), 0, st, n, F.ps, ctx->off.p,
                           ctx->nbr.p, F.rev.p, F.break_step.p, 0, F.Phat.p, F.nodeE.p, cks.p + nd, lamp.p,
                           lam.p, lamp.p, Ebar.p, 0.5 * c);
    else
        hipLaunchKernelGGL((k_adj_step<3, true>), dim3(gridn2(n, T)), dim3(256), 0, st, n, F.ps, ctx->off.p,
                           ctx->nbr.p, F.rev.p, F.break_step.p, 0, F.Phat.p, F.nodeE.p, cks.p + nd, lamp.p,
                           lam.p, lamp.p, Ebar.p, 0.5 * c);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipMemcpyAsync(grad, Ebar.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (J_out) PXG_HIP(hipMemcpyAsync(J_out, dJ.p, sizeof(double), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

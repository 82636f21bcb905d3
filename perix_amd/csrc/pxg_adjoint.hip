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
// and finally Ebar_i += c/2 (Khat_0 u^0)_i . lam_{-1}.
//
// Khat_k^T y at node i = sum_j h_ji Phat_ji y_j - (sum_j h_ij Phat_ij) y_i:
// families are symmetric and stretch is symmetric, so h_ji = h_ij and the
// transposed operator Phat_ji is pre-gathered into bond order (PhatT[t] =
// Phat[rev(t)]) once; every node then writes only its own entries (no
// atomics) and every bond array is streamed, not gathered.  Displacements
// are checkpointed every `ckpt` steps and each segment is recomputed forward
// (one fused, frozen-health launch per step) before its reverse sweep.  With
// the E-gradient a reverse step is one launch of k_adj_step (pxg_step.hip,
// on the forward step's tiling: S^T lam_k is formed while the neighbour rows
// {u_j, E_j wbar_j} are staged in LDS); the body-force gradient alone needs
// no states and runs k_preset_T and k_adj_rev below.
#include "pxg_internal.h"
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

namespace pxg {

// wbar = S^T lam for the merged BC program, and ew = E .* wbar
// per-block sums of w over the block's nodes, per component, added to the
// block's own slot part[blockIdx.x * 3 + c] (a fixed summation order, no
// atomics: the body-force gradient is reproducible).  256 threads.
__device__ __forceinline__ void block_sum3(const double (&w)[3], int dim, double *part)
{
    __shared__ double red[4][3];
    double v[3] = {w[0], w[1], w[2]};
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] += __shfl_xor(v[c], o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int c = 0; c < 3; c++) red[wv][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < dim; c++)
            part[blockIdx.x * 3 + c] += ((red[0][c] + red[1][c]) + (red[2][c] + red[3][c]));
}

__global__ void __launch_bounds__(256) k_preset_T(int n, int dim, const double *lam, const int *prog_of,
                                                  const int *bnode, const int *bmode, const int *inv_off,
                                                  const int *inv_q, const double *E, double *wbar, double *ew,
                                                  double *bpart)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double wsum[3] = {0.0, 0.0, 0.0};
    if (m < n) {
        const int q = prog_of[m];
        for (int c = 0; c < dim; c++) {
            double v = lam[(size_t)m * dim + c];
            if (q >= 0 && bmode[q * 3 + c] != 0) v = 0.0;
            for (int p = inv_off[m]; p < inv_off[m + 1]; p++) {
                const int qq = inv_q[p];
                const int md = bmode[qq * 3 + c];
                const int g = bnode[qq];
                if (md == 1) v += lam[(size_t)g * dim + c];
                else if (md == 2) v += -lam[(size_t)g * dim + c];
            }
            wbar[(size_t)m * dim + c] = v;
            ew[(size_t)m * dim + c] = E[m] * v;
            wsum[c] = v;
        }
    }
    // the body force enters every central-difference value: dJ/db += c S^T lam
    if (bpart) block_sum3(wsum, dim, bpart);
}

// per-block sums of ubar^{-1} (the u^{-1} term of dJ/db)
__global__ void __launch_bounds__(256) k_body_last(int n, int dim, const double *lamp, double *bpart)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double w[3] = {0.0, 0.0, 0.0};
    if (m < n)
        for (int c = 0; c < dim; c++) w[c] = lamp[(size_t)m * dim + c];
    block_sum3(w, dim, bpart);
}

// dJ/db = c sum_k S^T lam_k + (c/2) ubar^{-1}, block partials in block order
__global__ void k_body_final(int nblk, int nblk_last, int dim, const double *part, const double *part_last,
                             double c, double *gb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int a = 0; a < dim; a++) {
        double s = 0.0, l = 0.0;
        for (int b = 0; b < nblk; b++) s += part[b * 3 + a];
        for (int b = 0; b < nblk_last; b++) l += part_last[b * 3 + a];
        gb[a] = c * s + 0.5 * c * l;
    }
}

// PhatT[c*ps + t] = Phat[c*ps + rev[t]] (0 for an unpaired bond)
__global__ void k_transpose_ops(long long nb, long long ps, int nc, const int *rev, const double *Phat,
                                double *PhatT)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int r = rev[t];
    for (int c = 0; c < nc; c++) PhatT[c * ps + t] = (r >= 0) ? Phat[c * ps + r] : 0.0;
}

// o = P v for a symmetric per-bond operator given by its 3 or 6 components
template <int DIM>
__device__ __forceinline__ void pmul(const double (&p)[DIM == 3 ? 6 : 3], const double (&v)[DIM], double (&o)[DIM])
{
    if constexpr (DIM == 2) {
        o[0] = p[0] * v[0] + p[2] * v[1];
        o[1] = p[2] * v[0] + p[1] * v[1];
    } else {
        o[0] = p[0] * v[0] + p[3] * v[1] + p[4] * v[2];
        o[1] = p[3] * v[0] + p[1] * v[1] + p[5] * v[2];
        o[2] = p[4] * v[0] + p[5] * v[1] + p[2] * v[2];
    }
}

struct AdjArgs {
    int n, k;                   // nodes, step (bond t is intact iff bs[t] > k)
    long long ps;
    const int *off, *nbr, *bs;
    const double *Phat, *PhatT;
    const double *u;            // u^{k-1}
    const double *wbar, *ew;    // S^T lam_k and E .* wbar
    double *lam, *lamp, *Ebar;
    double cf;
    // without the E-gradient: the family sums sp_k = sum_{j intact at k} Phat_ij
    // kept per node (spg) and grown as k decreases by the bonds that break at
    // step k + 1 (brk_t[cur[i] ..], sorted by break step, descending)
    double *spg;
    int *cur;
    const int *brk_off, *brk_t;
};

template <int DIM> struct AdjCfg;
template <> struct AdjCfg<2> { static constexpr int S = 8, G = 2; };
template <> struct AdjCfg<3> { static constexpr int S = 16, G = 2; };

// Two forms: LAST (= EGRAD) adds the u^{-1} term of the E-gradient; the
// other is one reverse step of the body-force gradient alone (no states, the
// intact family sums kept in spg).  The E-gradient's reverse steps run
// k_adj_step (pxg_step.hip).  S lanes per node walk the family in aligned
// groups of G bonds with vector loads of ids, break steps and operators; the
// per-lane FP64 partial sums are combined by a lane butterfly (a fixed
// order, so results are reproducible).
template <int DIM, bool LAST, bool EGRAD = LAST>
__global__ void __launch_bounds__(256) k_adj_rev(AdjArgs a)
{
    static_assert(EGRAD == LAST, "the u^{-1} term of the E-gradient, or a body-force step");
    constexpr int S = AdjCfg<DIM>::S, G = AdjCfg<DIM>::G;
    constexpr int NPW = 64 / S, W = S * G;
    constexpr int NC = (DIM == 3) ? 6 : 3;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ln = lane / S, sl = lane % S;
    const int i = (blockIdx.x * 4 + wv) * NPW + ln;
    const bool nvalid = i < a.n;
    int b0 = 0, b1 = 0;
    if (nvalid) { b0 = a.off[i]; b1 = a.off[i + 1]; }
    double ui[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) ui[c] = (EGRAD && nvalid) ? a.u[(size_t)i * DIM + c] : 0.0;
    double ku[DIM], zo[DIM], sp[NC];
#pragma unroll
    for (int c = 0; c < DIM; c++) { ku[c] = 0.0; zo[c] = 0.0; }
#pragma unroll
    for (int c = 0; c < NC; c++) sp[c] = 0.0;
    for (int t0 = (b0 & ~(G - 1)) + G * sl; t0 < b1; t0 += W) {
        int jv[G], bv[G];
        double p[NC][G], pt[LAST ? 1 : NC][G];
        const unsigned u0 = (unsigned)t0;
        if constexpr (G == 2) {
            const int2 j2 = *reinterpret_cast<const int2 *>(a.nbr + u0);
            const int2 b2 = *reinterpret_cast<const int2 *>(a.bs + u0);
            jv[0] = j2.x; jv[1] = j2.y; bv[0] = b2.x; bv[1] = b2.y;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                if constexpr (EGRAD) {
                    const double2 q = *reinterpret_cast<const double2 *>(a.Phat + c * a.ps + u0);
                    p[c][0] = q.x; p[c][1] = q.y;
                } else {
                    p[c][0] = 0.0; p[c][1] = 0.0;
                }
                if constexpr (!LAST) {
                    const double2 r = *reinterpret_cast<const double2 *>(a.PhatT + c * a.ps + u0);
                    pt[c][0] = r.x; pt[c][1] = r.y;
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int t = t0 + g;
            if (t < b0 || t >= b1 || bv[g] <= a.k) continue;
            const size_t jo = (size_t)(jv[g] - 1) * DIM;
            double pc[NC], o[DIM];
#pragma unroll
            for (int c = 0; c < NC; c++) pc[c] = p[c][g];
            if constexpr (EGRAD) {
                double du[DIM];
#pragma unroll
                for (int c = 0; c < DIM; c++) du[c] = a.u[jo + c] - ui[c];
                pmul<DIM>(pc, du, o);
#pragma unroll
                for (int c = 0; c < DIM; c++) ku[c] += o[c];
            }
            if constexpr (!LAST) {
                double ej[DIM], ptc[NC];
#pragma unroll
                for (int c = 0; c < DIM; c++) ej[c] = a.ew[jo + c];
#pragma unroll
                for (int c = 0; c < NC; c++) ptc[c] = pt[c][g];
                pmul<DIM>(ptc, ej, o);
#pragma unroll
                for (int c = 0; c < DIM; c++) zo[c] += o[c];
            }
        }
    }
#pragma unroll
    for (int m = 1; m < S; m <<= 1) {
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            if constexpr (EGRAD) ku[c] += __shfl_xor(ku[c], m);
            if constexpr (!LAST) zo[c] += __shfl_xor(zo[c], m);
        }
    }
    if (!nvalid || sl >= DIM) return;
    if constexpr (!EGRAD) {
        // sp_k = sp_{k+1} + Phat of the node's bonds that break at step k+1
        // (every component lane reads the same entries; lane 0 writes back)
        const size_t base = (size_t)i * NC;
#pragma unroll
        for (int c = 0; c < NC; c++) sp[c] = a.spg[base + c];
        int q = a.cur[i];
        const int qe = a.brk_off[i + 1];
        while (q < qe) {
            const int t = a.brk_t[q];
            if (a.bs[t] != a.k + 1) break;
#pragma unroll
            for (int c = 0; c < NC; c++) sp[c] += a.Phat[c * a.ps + t];
            q++;
        }
        if (sl == 0) {
#pragma unroll
            for (int c = 0; c < NC; c++) a.spg[base + c] = sp[c];
            a.cur[i] = q;
        }
    }
    double wi[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) wi[c] = a.wbar[(size_t)i * DIM + c];
    if (EGRAD && sl == 0) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < DIM; c++) s += ku[c] * wi[c];
        a.Ebar[i] += a.cf * s;
    }
    if constexpr (!LAST) {
        double ei[DIM], d[DIM];
#pragma unroll
        for (int c = 0; c < DIM; c++) ei[c] = a.ew[(size_t)i * DIM + c];
        pmul<DIM>(sp, ei, d);
        const int c = sl;
        const double z = (c == 0 ? zo[0] - d[0] : (c == 1 ? zo[1] - d[1] : zo[DIM - 1] - d[DIM - 1]));
        const double wb = (c == 0 ? wi[0] : (c == 1 ? wi[1] : wi[DIM - 1]));
        const size_t q = (size_t)i * DIM + c;
        const double np = a.lamp[q] + 2.0 * wb + a.cf * z;
        a.lamp[q] = -wb;
        a.lam[q] = np;
    }
}

// broken-bond lists for the family sums without the E-gradient: per node the
// bonds with a break step in [1, K], sorted by break step (descending)
__global__ void k_brk_count(int n, const int *off, const int *bs, int K, int *cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    for (int t = off[i]; t < off[i + 1]; t++) c += (bs[t] >= 1 && bs[t] <= K) ? 1 : 0;
    cnt[i] = c;
}

__global__ void k_brk_fill(int n, const int *off, const int *bs, int K, const int *brk_off, int *brk_t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int q = brk_off[i];
    const int q0 = q;
    for (int t = off[i]; t < off[i + 1]; t++) {
        if (!(bs[t] >= 1 && bs[t] <= K)) continue;
        int p = q++;                    // insertion: descending break step, then bond order
        while (p > q0 && bs[brk_t[p - 1]] < bs[t]) { brk_t[p] = brk_t[p - 1]; p--; }
        brk_t[p] = t;
    }
}

// sp_K[i] = sum over the bonds intact at step K of Phat (family order)
template <int NC>
__global__ void k_sp_init(int n, const int *off, const int *bs, int K, const double *Phat, long long ps,
                          const int *brk_off, double *spg, int *cur)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) s[c] = 0.0;
    for (int t = off[i]; t < off[i + 1]; t++) {
        if (bs[t] <= K) continue;
#pragma unroll
        for (int c = 0; c < NC; c++) s[c] += Phat[c * ps + t];
    }
#pragma unroll
    for (int c = 0; c < NC; c++) spg[(size_t)i * NC + c] = s[c];
    cur[i] = brk_off[i];
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

static unsigned gridn2(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int launch_transpose_ops(pxg_ctx *ctx)
{
    FracState &F = ctx->frac;
    const int nc = (ctx->dim == 3) ? 6 : 3;
    PXG_HIP(F.PhatT.alloc((size_t)F.ps * nc));
    hipLaunchKernelGGL(k_transpose_ops, dim3(gridn2(ctx->nb, 256)), dim3(256), 0, ctx->stream, ctx->nb, F.ps, nc,
                       F.rev.p, F.Phat.p, F.PhatT.p);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

template <int DIM, bool LAST, bool EGRAD = LAST>
static int launch_adj_rev(pxg_ctx *ctx, int k, const double *u, const double *wbar, const double *ew,
                          double *lam, double *lamp, double *Ebar, double cf)
{
    FracState &F = ctx->frac;
    AdjArgs a;
    // break steps in CSR order (bond_state_csr refreshed the copy at the start)
    a.n = ctx->n; a.k = k; a.ps = F.ps; a.off = ctx->off.p; a.nbr = ctx->nbr.p;
    a.bs = F.bsh.p ? F.bs_csr.p : F.break_step.p;
    a.Phat = F.Phat.p; a.PhatT = F.PhatT.p; a.u = u; a.wbar = wbar; a.ew = ew;
    a.lam = lam; a.lamp = lamp; a.Ebar = Ebar; a.cf = cf;
    a.spg = F.adj.spg.p; a.cur = F.adj.cur.p; a.brk_off = F.adj.brk_off.p; a.brk_t = F.adj.brk_t.p;
    constexpr int npb = 4 * (64 / AdjCfg<DIM>::S);
    hipLaunchKernelGGL((k_adj_rev<DIM, LAST, EGRAD>), dim3(gridn2(ctx->n, npb)), dim3(256), 0, ctx->stream, a);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

// frozen-health forward over steps [k0+1, k1]: seg[0] = u^{k0-1} and
// seg[1] = u^{k0} on entry; step k writes u^k into seg[k-k0+1]
static int frozen_segment(pxg_ctx *ctx, double *seg, size_t nd, int k0, int k1, double dt, double t0)
{
    double t = t0;
    for (int k = k0 + 1; k <= k1; k++) {
        const double tNew = t + dt;
        double *um = seg + (size_t)(k - k0 - 1) * nd, *uc = seg + (size_t)(k - k0) * nd;
        const int rc = launch_step_fused(ctx, MODE_CD, uc, um, uc + nd, dt * dt, 0.0, k, tNew, false, k);
        if (rc) return rc;
        t = tNew;
    }
    return PXG_OK;
}

// forward with frozen health from u0: fills checkpoints (pairs u^{k-1}, u^k at
// k = 0, ckpt, 2ckpt, ...) when asked and returns u^K in the workspace's own
// ping-pong buffers, so the state of the last pxg_frac_run (FracState::u,
// cur, step, t) stays as it was
static int adj_forward(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int K, int ckpt,
                       double *cks, double **uK_out, double *tk)
{
    FracState &F = ctx->frac;
    AdjWork &A = F.adj;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)ctx->n * ctx->dim;
    PXG_HIP(A.fu[0].alloc(nd));
    PXG_HIP(A.fu[1].alloc(nd));
    PXG_HIP(hipMemcpyAsync(A.fu[0].p, u0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    if (v0) PXG_HIP(hipMemcpyAsync(F.accel.p, v0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    else PXG_HIP(hipMemsetAsync(F.accel.p, 0, sizeof(double) * nd, st));
    int rc = launch_step(ctx, MODE_INIT, A.fu[0].p, F.accel.p, A.fu[1].p, dt, 0.5 * dt * dt, 0, false, 0);
    if (rc) return rc;
    double *um = A.fu[1].p, *uc = A.fu[0].p;
    double t = 0.0;
    for (int k = 1; k <= K; k++) {
        if (cks && (k - 1) % ckpt == 0) {
            const int c = (k - 1) / ckpt;
            PXG_HIP(hipMemcpyAsync(cks + (size_t)(2 * c) * nd, um, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
            PXG_HIP(hipMemcpyAsync(cks + (size_t)(2 * c + 1) * nd, uc, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
            tk[c] = t;
        }
        const double tNew = t + dt;
        rc = launch_step_fused(ctx, MODE_CD, uc, um, um, dt * dt, 0.0, k, tNew, false, k);
        if (rc) return rc;
        double *tmp = um; um = uc; uc = tmp;
        t = tNew;
    }
    *uK_out = uc;
    return PXG_OK;
}

} // namespace pxg

using namespace pxg;

extern "C" int pxg_frac_loss(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int nsteps,
                             const double *node_E, int nsens, const int *sens, int comp, double *J_out)
{
    PXG_CHECK(ctx && ctx->frac.ready && ctx->frac.has_nodeE, PXG_ERR_STATE,
              "pxg_frac_loss: needs pxg_frac_setup with a per-node E");
    PXG_CHECK(u0 && dt > 0.0 && nsteps >= 1 && nsens >= 0 && (nsens == 0 || sens) && comp >= 0 &&
              comp < ctx->dim && J_out, PXG_ERR_ARG, "pxg_frac_loss: bad arguments");
    PXG_FAULT(ctx, "pxg_frac_loss");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    AdjWork &A = F.adj;
    hipStream_t st = ctx->stream;
    PXG_HIP(A.ds.alloc(nsens > 0 ? nsens : 1));
    PXG_HIP(A.dJ.alloc(1));
    if (nsens) PXG_HIP(hipMemcpyAsync(A.ds.p, sens, sizeof(int) * nsens, hipMemcpyHostToDevice, st));
    if (node_E) {
        PXG_HIP(A.eo.alloc(ctx->n));
        PXG_HIP(hipMemcpyAsync(A.eo.p, node_E, sizeof(double) * ctx->n, hipMemcpyHostToDevice, st));
        int rc = launch_frac_ops(ctx, A.eo.p);
        if (rc) return rc;
    }
    double *uK = nullptr;
    int rc = adj_forward(ctx, u0, v0, dt, nsteps, 1, nullptr, &uK, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(1), 0, st, nsens, A.ds.p, ctx->dim, comp, uK, A.dJ.p);
    PXG_HIP(hipMemcpyAsync(J_out, A.dJ.p, sizeof(double), hipMemcpyDeviceToHost, st));
    if (node_E) { rc = launch_frac_ops(ctx, F.nodeE.p); if (rc) return rc; }
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_frac_set_body(pxg_ctx *ctx, const double *b)
{
    PXG_CHECK(ctx && ctx->frac.ready && b, PXG_ERR_ARG, "pxg_frac_set_body: bad arguments");
    for (int c = 0; c < 3; c++) ctx->frac.b[c] = (c < ctx->dim) ? b[c] : 0.0;
    return PXG_OK;
}

extern "C" int pxg_frac_adjoint(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int nsteps,
                                int ckpt, int nsens, const int *sens, int comp, double *J_out, double *grad)
{
    PXG_CHECK(grad, PXG_ERR_ARG, "pxg_frac_adjoint: bad arguments");
    return pxg_frac_adjoint_ex(ctx, u0, v0, dt, nsteps, ckpt, nsens, sens, comp, J_out, grad, nullptr);
}

extern "C" int pxg_frac_adjoint_ex(pxg_ctx *ctx, const double *u0, const double *v0, double dt, int nsteps,
                                   int ckpt, int nsens, const int *sens, int comp, double *J_out,
                                   double *grad, double *grad_b)
{
    PXG_CHECK(ctx && ctx->frac.ready && ctx->frac.has_nodeE, PXG_ERR_STATE,
              "pxg_frac_adjoint: needs pxg_frac_setup with a per-node E (and a recorded forward run)");
    PXG_CHECK(u0 && dt > 0.0 && nsteps >= 1 && ckpt >= 1 && nsens >= 0 && (nsens == 0 || sens) &&
              comp >= 0 && comp < ctx->dim, PXG_ERR_ARG, "pxg_frac_adjoint: bad arguments");
    PXG_FAULT(ctx, "pxg_frac_adjoint");
    PXG_HIP(hipSetDevice(ctx->device));
    FracState &F = ctx->frac;
    AdjWork &A = F.adj;
    hipStream_t st = ctx->stream;
    const int n = ctx->n, dim = ctx->dim, K = nsteps;
    const size_t nd = (size_t)n * dim;
    const int ncks = (K - 1) / ckpt + 1;
    // workspace kept by the context between calls (allocated on size changes);
    // without the E-gradient the reverse sweep reads no states: no checkpoints
    if (grad) {
        PXG_HIP(A.cks.alloc((size_t)2 * ncks * nd));
        PXG_HIP(A.seg.alloc((size_t)(ckpt + 2) * nd));
    }
    PXG_HIP(A.lam.alloc(nd)); PXG_HIP(A.lamp.alloc(nd));
    PXG_HIP(A.wbar.alloc(nd)); PXG_HIP(A.ew.alloc(nd));
    PXG_HIP(A.Ebar.alloc(n)); PXG_HIP(A.dJ.alloc(1));
    PXG_HIP(A.ds.alloc(nsens > 0 ? nsens : 1));
    // body-force partials: per block of the reverse step (node blocks of 256
    // without the E-gradient, the step tiling's blocks with it), then per
    // node block of 256 for the u^{-1} term
    const int nblk = (int)gridn2(n, 256);
    const int nblk_rev = grad ? adj_step_blocks(ctx) : nblk;
    PXG_HIP(A.bpart.alloc((size_t)3 * (nblk_rev + nblk)));
    PXG_HIP(A.gb.alloc(3));
    PXG_HIP(hipMemsetAsync(A.bpart.p, 0, sizeof(double) * 3 * (nblk_rev + nblk), st));
    double *bpart_last = A.bpart.p + (size_t)3 * nblk_rev;
    if (nsens) PXG_HIP(hipMemcpyAsync(A.ds.p, sens, sizeof(int) * nsens, hipMemcpyHostToDevice, st));
    std::vector<double> tk(ncks, 0.0);
    double *uK = nullptr;
    const int *bs_csr = nullptr;
    int rc = bond_state_csr(ctx, nullptr, &bs_csr);   // the recorded break steps, CSR order
    if (rc) return rc;
    rc = adj_forward(ctx, u0, v0, dt, K, ckpt, grad ? A.cks.p : nullptr, &uK, tk.data());
    if (rc) return rc;
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(1), 0, st, nsens, A.ds.p, dim, comp, uK, A.dJ.p);
    PXG_HIP(hipMemsetAsync(A.lam.p, 0, sizeof(double) * nd, st));
    PXG_HIP(hipMemsetAsync(A.lamp.p, 0, sizeof(double) * nd, st));
    PXG_HIP(hipMemsetAsync(A.Ebar.p, 0, sizeof(double) * n, st));
    hipLaunchKernelGGL(k_seed_lambda, dim3(1), dim3(1), 0, st, nsens, A.ds.p, dim, comp, uK, A.lam.p);
    const double c = dt * dt / F.rho;
    if (!grad) {
        // body-force gradient alone: the lambda recursion needs the family sums
        // of the intact operators, kept per node and grown by the bonds that
        // break at step k+1 as k decreases
        const int nc = (dim == 3) ? 6 : 3;
        PXG_HIP(A.cur.alloc(n));
        PXG_HIP(A.brk_off.alloc((size_t)n + 1));
        PXG_HIP(A.spg.alloc((size_t)n * nc));
        hipLaunchKernelGGL(k_brk_count, dim3(gridn2(n, 256)), dim3(256), 0, st, n, ctx->off.p, bs_csr, K, A.cur.p);
        PXG_HIP(hipGetLastError());
        std::vector<int> cnt((size_t)n), boff((size_t)n + 1, 0);
        PXG_HIP(hipMemcpyAsync(cnt.data(), A.cur.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        PXG_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < n; i++) boff[(size_t)i + 1] = boff[(size_t)i] + cnt[(size_t)i];
        PXG_HIP(A.brk_t.alloc((size_t)std::max(boff[(size_t)n], 1)));
        PXG_HIP(hipMemcpyAsync(A.brk_off.p, boff.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_brk_fill, dim3(gridn2(n, 256)), dim3(256), 0, st, n, ctx->off.p, bs_csr, K,
                           A.brk_off.p, A.brk_t.p);
        if (dim == 3)
            hipLaunchKernelGGL(k_sp_init<6>, dim3(gridn2(n, 256)), dim3(256), 0, st, n, ctx->off.p, bs_csr, K,
                               F.Phat.p, F.ps, A.brk_off.p, A.spg.p, A.cur.p);
        else
            hipLaunchKernelGGL(k_sp_init<3>, dim3(gridn2(n, 256)), dim3(256), 0, st, n, ctx->off.p, bs_csr, K,
                               F.Phat.p, F.ps, A.brk_off.p, A.spg.p, A.cur.p);
        PXG_HIP(hipGetLastError());
        for (int k = K; k >= 1; k--) {
            hipLaunchKernelGGL(k_preset_T, dim3(gridn2(n, 256)), dim3(256), 0, st, n, dim, A.lam.p,
                               F.bc.prog_of.p, F.bc.node.p, F.bc.mode.p, F.bc.inv_off.p, F.bc.inv_q.p,
                               F.nodeE.p, A.wbar.p, A.ew.p, grad_b ? A.bpart.p : nullptr);
            PXG_HIP(hipGetLastError());
            rc = (dim == 2)
                     ? launch_adj_rev<2, false, false>(ctx, k, nullptr, A.wbar.p, A.ew.p, A.lam.p, A.lamp.p, nullptr, c)
                     : launch_adj_rev<3, false, false>(ctx, k, nullptr, A.wbar.p, A.ew.p, A.lam.p, A.lamp.p, nullptr, c);
            if (rc) return rc;
        }
        if (grad_b) {
            hipLaunchKernelGGL(k_body_last, dim3(nblk), dim3(256), 0, st, n, dim, A.lamp.p, bpart_last);
            hipLaunchKernelGGL(k_body_final, dim3(1), dim3(64), 0, st, nblk_rev, nblk, dim, A.bpart.p, bpart_last,
                               c, A.gb.p);
            PXG_HIP(hipGetLastError());
            PXG_HIP(hipMemcpyAsync(grad_b, A.gb.p, sizeof(double) * dim, hipMemcpyDeviceToHost, st));
        }
        if (J_out) PXG_HIP(hipMemcpyAsync(J_out, A.dJ.p, sizeof(double), hipMemcpyDeviceToHost, st));
        PXG_HIP(hipStreamSynchronize(st));
        return PXG_OK;
    }
    // the reverse sweep: one fused launch per step (S^T lam_k staged with the
    // neighbour rows), lam ping-ponging between two buffers
    rc = adj_step_setup(ctx);
    if (rc) return rc;
    PXG_HIP(A.lam2.alloc(nd));
    double *lam_k = A.lam.p, *lam_km1 = A.lam2.p;
    for (int cidx = ncks - 1; cidx >= 0; cidx--) {
        const int k0 = cidx * ckpt;                 // segment covers steps k0+1 .. k1
        const int k1 = (k0 + ckpt < K) ? k0 + ckpt : K;
        PXG_HIP(hipMemcpyAsync(A.seg.p, A.cks.p + (size_t)(2 * cidx) * nd, sizeof(double) * 2 * nd,
                               hipMemcpyDeviceToDevice, st));
        rc = frozen_segment(ctx, A.seg.p, nd, k0, k1, dt, tk[cidx]);
        if (rc) return rc;
        for (int k = k1; k >= k0 + 1; k--) {
            const double *ukm1 = A.seg.p + (size_t)(k - k0) * nd;     // u^{k-1}
            rc = launch_adj_step(ctx, k, ukm1, lam_k, lam_km1, A.lamp.p, A.Ebar.p, grad_b ? A.bpart.p : nullptr, c);
            if (rc) return rc;
            double *tmp = lam_k; lam_k = lam_km1; lam_km1 = tmp;
        }
    }
    // u^{-1} term with h^0 at u^0 (checkpoint 0 holds u^{-1}, u^0)
    rc = (dim == 2) ? launch_adj_rev<2, true>(ctx, 0, A.cks.p + nd, A.lamp.p, nullptr, lam_k, A.lamp.p, A.Ebar.p, 0.5 * c)
                    : launch_adj_rev<3, true>(ctx, 0, A.cks.p + nd, A.lamp.p, nullptr, lam_k, A.lamp.p, A.Ebar.p, 0.5 * c);
    if (rc) return rc;
    if (grad_b) {
        hipLaunchKernelGGL(k_body_last, dim3(nblk), dim3(256), 0, st, n, dim, A.lamp.p, bpart_last);
        hipLaunchKernelGGL(k_body_final, dim3(1), dim3(64), 0, st, nblk_rev, nblk, dim, A.bpart.p, bpart_last, c,
                           A.gb.p);
        PXG_HIP(hipGetLastError());
        PXG_HIP(hipMemcpyAsync(grad_b, A.gb.p, sizeof(double) * dim, hipMemcpyDeviceToHost, st));
    }
    if (grad) PXG_HIP(hipMemcpyAsync(grad, A.Ebar.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (J_out) PXG_HIP(hipMemcpyAsync(J_out, A.dJ.p, sizeof(double), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

// pxg_ops.hip -- per-node PDDO moment matrix and per-bond g-weights (FP64).
//
// Reference: PDOperators::calcAMatrixCompute / calcOperatorsCompute
// (CalcPDOperators.cpp:108-354) with MatrixXd::solveMatrixChecked
// (MatrixXd.cpp:26-48), i.e. Eigen's FullPivLU applied to A X = diag(b).
//
// One 64-lane wavefront per node (block = one wave, so __syncthreads is a
// wave barrier):
//   1. lanes take 64 bonds of the family at a time, evaluate |xi|, the
//      Gaussian weight, the rim volume factor and the monomials eta^q into
//      LDS;
//   2. each lane owns one or two moment-matrix entries and accumulates them
//      over the staged bonds in family order -- the same additions, in the
//      same order, as the reference's nested loop, so A is bit-identical;
//   3. complete-pivot LU with lanes parallel over matrix entries (pivot =
//      first maximum in column-major order, as Eigen's maxCoeff visitor);
//   4. lanes j < nop solve column j of A X = P diag(b) with the panel
//      structure of Eigen's row-major-rhs triangular solve (panels of 4 rows,
//      one sequential dot product per panel over the solved rows, reciprocal
//      multiply on the diagonal);
//   5. consistency residual max|AX-B| <= 1e-8 * scale, then every bond's
//      op_k = ((sum_a X[a,k] eta^{q_a}) * w vc V_j) / delta^{|q_k|}.
// All arithmetic is compiled without FMA contraction so that it rounds like
// the reference's x86-64 build.  The build is latency-bound (dependent
// neighbour gathers, a short serial LU per node), so occupancy is what hides
// it: empty asm memory clobbers keep the LU factors and X in LDS rather than
// hoisted into registers, and the kernel fits 5 waves per SIMD.
#include "pxg_internal.h"
#include <cfloat>
#include <cmath>

namespace pxg {

__constant__ int c_qp[kMaxNop], c_qq[kMaxNop], c_qr[kMaxNop];
__constant__ double c_bdiag[kMaxNop];

template <int NOP>
__device__ __forceinline__ void bond_geometry(int i, int j1, const double *xp, int dim,
                                              const double *vols, double delta, double inv_delta,
                                              double dx_char, int volcorr, double &bw, double *mono)
{
    const int j = j1 - 1;
    const double xx = xp[(size_t)j * dim] - xp[(size_t)i * dim];
    const double xy = xp[(size_t)j * dim + 1] - xp[(size_t)i * dim + 1];
    const double xz = (dim == 3) ? xp[(size_t)j * dim + 2] - xp[(size_t)i * dim + 2] : 0.0;
    const double xn = sqrt(xx * xx + xy * xy + xz * xz);
    bw = 0.0;
    if (!(xn > 0.0)) return;
    // computeWeight / partialVolumeCorrection (CalcPDOperators.cpp:48-60)
    const double r = 2.0 * xn / delta;
    const double w = exp(-r * r);
    double vc = 1.0;
    if (volcorr && dx_char > 0.0 && !(xn <= delta - 0.5 * dx_char)) {
        vc = (delta - xn + 0.5 * dx_char) / dx_char;
        vc = vc < 0.0 ? 0.0 : (vc > 1.0 ? 1.0 : vc);
    }
    bw = w * vc * vols[j];
    if (!(bw > 0.0)) { bw = 0.0; return; }
    const double ex = xx * inv_delta, ey = xy * inv_delta, ez = xz * inv_delta;
#pragma unroll
    for (int k = 0; k < NOP; k++) {
        double m = 1.0;
        for (int t = 0; t < c_qp[k]; t++) m *= ex;
        for (int t = 0; t < c_qq[k]; t++) m *= ey;
        for (int t = 0; t < c_qr[k]; t++) m *= ez;
        mono[k] = m;
    }
}

template <int NOP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_operators(int n, int dim, const double *xp, const double *vols, const int *off, const int *nbr,
            double delta, double dx_char, int volcorr, double *ops, unsigned char *flags, int *ndeg)
{
    constexpr int NN = NOP * NOP;
    __shared__ double s_bw[64];
    __shared__ double s_mono[64][NOP];
    __shared__ double s_A[NN];
    __shared__ double s_lu[NN];
    __shared__ double s_X[NN];
    __shared__ int s_rowT[NOP], s_colT[NOP];
    __shared__ int s_nonzero;
    __shared__ double s_maxpiv;

    const int i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const int b0 = off[i], b1 = off[i + 1];
    const double inv_delta = 1.0 / delta;

    // ---- 1+2: moment matrix, bit-identical accumulation order ----------
    const int e0 = lane, e1 = lane + 64;
    const int a0 = e0 / NOP, c0 = e0 % NOP, a1 = e1 / NOP, c1 = e1 % NOP;
    double acc0 = 0.0, acc1 = 0.0;
    int usable = 0;
    for (int s = b0; s < b1; s += 64) {
        const int t = s + lane;
        double bw = 0.0, mono[NOP];
#pragma unroll
        for (int k = 0; k < NOP; k++) mono[k] = 0.0;
        if (t < b1) bond_geometry<NOP>(i, nbr[t], xp, dim, vols, delta, inv_delta, dx_char, volcorr, bw, mono);
        s_bw[lane] = bw;
#pragma unroll
        for (int k = 0; k < NOP; k++) s_mono[lane][k] = mono[k];
        usable += __popcll(__ballot(bw > 0.0));
        __syncthreads();
        // An unusable bond has bw = 0 and zero monomials, so its term is
        // +-0 and adding it leaves the sum bit-unchanged: no branch needed.
        const int cnt = min(64, b1 - s);
#pragma unroll 8
        for (int u = 0; u < cnt; u++) {
            const double bu = s_bw[u];
            acc0 += bu * s_mono[u][a0] * s_mono[u][c0];
            if (e1 < NN) acc1 += bu * s_mono[u][a1] * s_mono[u][c1];
        }
        __syncthreads();
    }
    if (e0 < NN) { s_A[e0] = acc0; s_lu[e0] = acc0; }
    if (e1 < NN) { s_A[e1] = acc1; s_lu[e1] = acc1; }
    if (lane == 0) { s_nonzero = NOP; s_maxpiv = 0.0; }
    __syncthreads();

    // ---- 3: complete-pivot LU (FullPivLU.h:360-439) ---------------------
    for (int k = 0; k < NOP; k++) {
        // biggest |a| of the corner, ties -> first in column-major order
        double best = -1.0; int bidx = 1 << 30;
        for (int e = lane; e < NN; e += 64) {
            const int r = e / NOP, c = e % NOP;
            if (r >= k && c >= k) {
                const double sc = fabs(s_lu[e]);
                const int ci = c * NOP + r;
                if (sc > best || (sc == best && ci < bidx)) { best = sc; bidx = ci; }
            }
        }
        for (int m = 32; m >= 1; m >>= 1) {
            const double ob = __shfl_xor(best, m);
            const int oi = __shfl_xor(bidx, m);
            if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
        }
        if (best == 0.0) {
            if (lane == 0) {
                s_nonzero = k;
                for (int q = k; q < NOP; q++) { s_rowT[q] = q; s_colT[q] = q; }
            }
            __syncthreads();
            break;
        }
        const int br = bidx % NOP, bc = bidx / NOP;
        if (lane == 0) {
            if (best > s_maxpiv) s_maxpiv = best;
            s_rowT[k] = br; s_colT[k] = bc;
        }
        if (br != k && lane < NOP) {
            const double t = s_lu[k * NOP + lane];
            s_lu[k * NOP + lane] = s_lu[br * NOP + lane];
            s_lu[br * NOP + lane] = t;
        }
        __syncthreads();
        if (bc != k && lane < NOP) {
            const double t = s_lu[lane * NOP + k];
            s_lu[lane * NOP + k] = s_lu[lane * NOP + bc];
            s_lu[lane * NOP + bc] = t;
        }
        __syncthreads();
        if (k < NOP - 1) {
            const double piv = s_lu[k * NOP + k];
            if (lane > k && lane < NOP) s_lu[lane * NOP + k] /= piv;
            __syncthreads();
            for (int e = lane; e < NN; e += 64) {
                const int r = e / NOP, c = e % NOP;
                if (r > k && c > k) s_lu[e] -= s_lu[r * NOP + k] * s_lu[k * NOP + c];
            }
            __syncthreads();
        }
    }
    __syncthreads();

    // ---- 4: column-parallel solve (FullPivLU.h:594-628) ----------------
    const int nonzero = s_nonzero;
    const double thr = fabs(s_maxpiv) * ((double)NOP * DBL_EPSILON);
    int rank = 0;
    for (int q = 0; q < nonzero; q++) rank += (fabs(s_lu[q * NOP + q]) > thr);
    if (lane < NOP) {
        const int j = lane;
        double c[NOP];
#pragma unroll
        for (int r = 0; r < NOP; r++) c[r] = (r == j) ? c_bdiag[j] : 0.0;
        if (rank > 0) {
            for (int k = 0; k < NOP; k++) {
                const int r = s_rowT[k];
                if (r != k) {
                    double tk = 0.0, tr = 0.0;
#pragma unroll
                    for (int q = 0; q < NOP; q++) { if (q == k) tk = c[q]; if (q == r) tr = c[q]; }
#pragma unroll
                    for (int q = 0; q < NOP; q++) { if (q == k) c[q] = tr; if (q == r) c[q] = tk; }
                }
            }
            // unit-lower, panels of 4 starting at row 0
#pragma unroll
            for (int j2 = 0; j2 < NOP; j2 += 4) {
                asm volatile("" ::: "memory");   // re-read L per panel, not all 81 in registers
                const int w = (NOP - j2 < 4) ? NOP - j2 : 4;
                if (j2 > 0) {
#pragma unroll
                    for (int r = j2; r < j2 + w; r++) {
                        double accd = 0.0;
#pragma unroll
                        for (int k = 0; k < j2; k++) accd += c[k] * s_lu[r * NOP + k];
                        c[r] = c[r] + (-1.0 * accd);
                    }
                }
#pragma unroll
                for (int r = j2; r < j2 + w; r++)
#pragma unroll
                    for (int k = j2; k < r; k++) c[r] -= c[k] * s_lu[r * NOP + k];
            }
            // upper on the rank corner, panels aligned to row 0, last first
            const int sz = rank;
            int j2 = sz - ((sz % 4) ? (sz % 4) : 4);
            for (; j2 >= 0; j2 -= 4) {
                asm volatile("" ::: "memory");   // same for U
                const int w = (sz - j2 < 4) ? sz - j2 : 4;
                const int po = j2 + w;
#pragma unroll
                for (int r = 0; r < NOP; r++) {
                    if (r < j2 || r >= po) continue;
                    if (po < sz) {
                        double accd = 0.0;
#pragma unroll
                        for (int k = 0; k < NOP; k++) if (k >= po && k < sz) accd += c[k] * s_lu[r * NOP + k];
                        c[r] = c[r] + (-1.0 * accd);
                    }
                }
#pragma unroll
                for (int r = NOP - 1; r >= 0; r--) {
                    if (r < j2 || r >= po) continue;
#pragma unroll
                    for (int k = 0; k < NOP; k++) if (k > r && k < po) c[r] -= c[k] * s_lu[r * NOP + k];
                    const double inv = 1.0 / s_lu[r * NOP + r];
                    c[r] *= inv;
                }
            }
        }
        // undo the column permutation
        int q[NOP];
#pragma unroll
        for (int r = 0; r < NOP; r++) q[r] = r;
        for (int k = 0; k < NOP; k++) {
            const int ck = s_colT[k];
            int qk = 0, qc = 0;
#pragma unroll
            for (int r = 0; r < NOP; r++) { if (r == k) qk = q[r]; if (r == ck) qc = q[r]; }
#pragma unroll
            for (int r = 0; r < NOP; r++) { if (r == k) q[r] = qc; if (r == ck) q[r] = qk; }
        }
#pragma unroll
        for (int r = 0; r < NOP; r++) s_X[q[r] * NOP + j] = (rank > 0 && r < rank) ? c[r] : 0.0;
    }
    __syncthreads();

    // ---- 5: consistency check (MatrixXd.cpp:37-47) ----------------------
    double resid = 0.0, bmax = 0.0, amax = 0.0, xmax = 0.0;
    for (int e = lane; e < NN; e += 64) {
        const int r = e / NOP, c = e % NOP;
        double sacc = 0.0;
#pragma unroll
        for (int k = 0; k < NOP; k++) sacc += s_A[r * NOP + k] * s_X[k * NOP + c];
        const double bv = (r == c) ? c_bdiag[r] : 0.0;
        resid = fmax(resid, fabs(sacc - bv));
        bmax = fmax(bmax, fabs(bv));
        amax = fmax(amax, fabs(s_A[e]));
        xmax = fmax(xmax, fabs(s_X[e]));
    }
    for (int m = 32; m >= 1; m >>= 1) {
        resid = fmax(resid, __shfl_xor(resid, m));
        bmax = fmax(bmax, __shfl_xor(bmax, m));
        amax = fmax(amax, __shfl_xor(amax, m));
        xmax = fmax(xmax, __shfl_xor(xmax, m));
    }
    const double scale = fmax(bmax, amax * xmax);
    const bool consistent = resid <= 1.0e-8 * fmax(scale, DBL_MIN);
    if (lane == 0) {
        unsigned char f = 0;
        if (usable < NOP) f |= 2;
        if (usable < NOP || !consistent) { f |= 1; atomicAdd(ndeg, 1); }
        flags[i] = f;
    }

    // ---- 6: per-bond operators (CalcPDOperators.cpp:307-354) ------------
    double idp[NOP];
#pragma unroll
    for (int k = 0; k < NOP; k++) {
        double v = 1.0;
        const int tot = c_qp[k] + c_qq[k] + c_qr[k];
        for (int t = 0; t < tot; t++) v *= inv_delta;
        idp[k] = v;
    }
    for (int s = b0; s < b1; s += 64) {
        const int t = s + lane;
        if (t >= b1) continue;
        double bw = 0.0, mono[NOP];
#pragma unroll
        for (int k = 0; k < NOP; k++) mono[k] = 0.0;
        bond_geometry<NOP>(i, nbr[t], xp, dim, vols, delta, inv_delta, dx_char, volcorr, bw, mono);
        double *op = ops + (size_t)t * NOP;
        // keep X in LDS (broadcast reads) instead of 81 hoisted registers
        asm volatile("" ::: "memory");
#pragma unroll
        for (int k = 0; k < NOP; k++) {
            asm volatile("" ::: "memory");
            double v = 0.0;
            if (bw > 0.0) {
                double sacc = 0.0;
#pragma unroll
                for (int a = 0; a < NOP; a++) sacc += s_X[a * NOP + k] * mono[a];
                v = sacc * bw * idp[k];
            }
            op[k] = v;
        }
    }
}

} // namespace pxg

using namespace pxg;

extern "C" int pxg_build_operators(pxg_ctx *ctx, int order, double delta, double dx_char, int *ndeg_out)
{
    PXG_CHECK(ctx && ctx->nb > 0, PXG_ERR_STATE, "pxg_build_operators: no families");
    PXG_CHECK(order == 2, PXG_ERR_ARG, "pxg_build_operators: Order 2 is the supported PDDO order "
                                       "(5 slots in 2D, 9 in 3D)");
    PXG_CHECK(delta > 0.0, PXG_ERR_ARG, "pxg_build_operators: delta must be > 0");
    PXG_FAULT(ctx, "pxg_build_operators");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int dim = ctx->dim;
    // multi-index table (PDOperators.cpp:51-75) and b_k = p!q!r!
    int qp[kMaxNop] = {0}, qq[kMaxNop] = {0}, qr[kMaxNop] = {0};
    double bd[kMaxNop] = {0};
    int k = 0;
    for (int o = 1; o <= order; o++) {
        if (dim == 3) {
            for (int p = o; p >= 0; --p)
                for (int q = o - p; q >= 0; --q) { qp[k] = p; qq[k] = q; qr[k] = o - p - q; k++; }
        } else {
            for (int p = o; p >= 0; --p) { qp[k] = p; qq[k] = o - p; qr[k] = 0; k++; }
        }
    }
    const int nop = k;
    auto fact = [](int m) { int r = 1; for (int i = 2; i <= m; i++) r *= i; return r; };
    for (int s = 0; s < nop; s++) bd[s] = (double)(fact(qp[s]) * fact(qq[s]) * fact(qr[s]));
    PXG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_qp), qp, sizeof(qp), 0, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_qq), qq, sizeof(qq), 0, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_qr), qr, sizeof(qr), 0, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_bdiag), bd, sizeof(bd), 0, hipMemcpyHostToDevice, st));
    // the operators are rewritten: every element built on the old ones is stale
    invalidate(ctx, INV_OPS);
    ctx->nop = 0;
    PXG_HIP(ctx->ops.alloc((size_t)ctx->nb * nop));
    PXG_HIP(ctx->flags.alloc(ctx->n));
    dbuf<int> nd;
    PXG_HIP(nd.alloc(1));
    PXG_HIP(hipMemsetAsync(nd.p, 0, sizeof(int), st));
    if (nop == 5)
        hipLaunchKernelGGL(k_operators<5>, dim3(ctx->n), dim3(64), 0, st, ctx->n, dim, ctx->xp.p,
                           ctx->vols.p, ctx->off.p, ctx->nbr.p, delta, dx_char, dx_char > 0.0 ? 1 : 0,
                           ctx->ops.p, ctx->flags.p, nd.p);
    else
        hipLaunchKernelGGL(k_operators<9>, dim3(ctx->n), dim3(64), 0, st, ctx->n, dim, ctx->xp.p,
                           ctx->vols.p, ctx->off.p, ctx->nbr.p, delta, dx_char, dx_char > 0.0 ? 1 : 0,
                           ctx->ops.p, ctx->flags.p, nd.p);
    PXG_HIP(hipGetLastError());
    int ndeg = 0;
    PXG_HIP(hipMemcpyAsync(&ndeg, nd.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    ctx->order = order; ctx->nop = nop; ctx->delta = delta; ctx->ndegenerate = ndeg;
    if (ndeg_out) *ndeg_out = ndeg;
    return PXG_OK;
}

extern "C" int pxg_get_operators(pxg_ctx *ctx, double *ops, unsigned char *flags)
{
    PXG_CHECK(ctx && ctx->nop > 0, PXG_ERR_STATE, "pxg_get_operators: no operators");
    PXG_HIP(hipSetDevice(ctx->device));
    if (ops) PXG_HIP(hipMemcpyAsync(ops, ctx->ops.p, sizeof(double) * (size_t)ctx->nb * ctx->nop, hipMemcpyDeviceToHost, ctx->stream));
    if (flags) PXG_HIP(hipMemcpyAsync(flags, ctx->flags.p, (size_t)ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

// pxg_ops.hip -- per-node PDDO moment matrix and per-bond g-weights (FP64).
//
// Reference: PDOperators::calcAMatrixCompute / calcOperatorsCompute
// (CalcPDOperators.cpp:108-354) with MatrixXd::solveMatrixChecked
// (MatrixXd.cpp:26-48), i.e. Eigen's FullPivLU applied to A X = diag(b).
//
// Four nodes per 64-lane wavefront, 16 lanes per node (block = one wave, so
// an LDS wait is the only barrier needed); a node's lanes run the whole build:
//   1. the 16 lanes take 16 bonds of the family at a time, evaluate |xi|,
//      the Gaussian weight, the rim volume factor and the monomials eta^q
//      into LDS;
//   2. lane c < nop owns column c of the moment matrix and accumulates it
//      over the staged bonds in family order -- the same products and
//      additions, in the same order, as the reference's nested loop, so A is
//      bit-identical;
//   3. complete-pivot LU with lanes parallel over matrix entries (pivot =
//      first maximum in column-major order, as Eigen's maxCoeff visitor);
//   4. lanes j < nop solve column j of A X = P diag(b) with the panel
//      structure of Eigen's row-major-rhs triangular solve (panels of 4 rows,
//      one sequential dot product per panel over the solved rows, reciprocal
//      multiply on the diagonal);
//   5. consistency residual max|AX-B| <= 1e-8 * scale, then every bond's
//      op_k = ((sum_a X[a,k] eta^{q_a}) * w vc V_j) / delta^{|q_k|}.
// All arithmetic is compiled without FMA contraction so that it rounds like
// the reference's x86-64 build.  The build is latency- and issue-bound (a
// short serial LU per node, the 9-lane column solve): four nodes per wave
// keep 36 of 64 lanes busy in the solve and run the pivot reductions of
// four nodes in one shuffle tree.  LDS is reused across phases (staged
// bonds, then A and X) so that the kernel keeps 5 waves per SIMD; empty asm
// memory clobbers keep the LU factors and X in LDS rather than hoisted into
// registers.
#include "pxg_internal.h"
#include <cfloat>
#include <cmath>
#include <utility>

namespace pxg {

// Multi-index table of the operator slots (PDOperators.cpp:51-75): for
// o = 1, 2 the exponents (p, q, r) with p + q + r = o, p descending, then q
// descending (r = 0 in 2D); b_k = p!q!r!.  The build supports order 2 only,
// so the table is a compile-time function of the slot count (5 in 2D, 9 in
// 3D) and every monomial is a fixed product sequence.
__host__ __device__ constexpr int slot_exp(int nop, int k, int axis)
{
    int s = 0;
    for (int o = 1; o <= 2; o++)
        for (int a = o; a >= 0; --a) {
            if (nop == 9) {
                for (int c = o - a; c >= 0; --c) {
                    if (s == k) return axis == 0 ? a : (axis == 1 ? c : o - a - c);
                    s++;
                }
            } else {
                if (s == k) return axis == 0 ? a : (axis == 1 ? o - a : 0);
                s++;
            }
        }
    return 0;
}
__host__ __device__ constexpr int fact(int m) { return m <= 1 ? 1 : m * fact(m - 1); }
__host__ __device__ constexpr double slot_b(int nop, int k)
{
    return (double)(fact(slot_exp(nop, k, 0)) * fact(slot_exp(nop, k, 1)) * fact(slot_exp(nop, k, 2)));
}
static_assert(slot_exp(9, 4, 0) == 1 && slot_exp(9, 4, 1) == 1 && slot_exp(9, 8, 2) == 2 && slot_exp(5, 2, 0) == 2,
              "order-2 slot table");

// x^P y^Q z^R as the reference forms it: 1.0 times x P times, then y, then z
template <int P, int Q, int R>
__device__ __forceinline__ double slot_monomial(const double (&e)[3])
{
    double m = 1.0;
#pragma unroll
    for (int t = 0; t < P; t++) m *= e[0];
#pragma unroll
    for (int t = 0; t < Q; t++) m *= e[1];
#pragma unroll
    for (int t = 0; t < R; t++) m *= e[2];
    return m;
}
template <int NOP, int... K>
__device__ __forceinline__ void monomials_seq(const double (&e)[3], double *mono, std::integer_sequence<int, K...>)
{
    ((mono[K] = slot_monomial<slot_exp(NOP, K, 0), slot_exp(NOP, K, 1), slot_exp(NOP, K, 2)>(e)), ...);
}
// delta^-|q_k| as 1.0 times 1/delta |q_k| times
template <int NOP, int... K>
__device__ __forceinline__ void inv_powers_seq(double inv_delta, double *idp, std::integer_sequence<int, K...>)
{
    const double e[3] = {inv_delta, inv_delta, inv_delta};
    ((idp[K] = slot_monomial<slot_exp(NOP, K, 0) + slot_exp(NOP, K, 1) + slot_exp(NOP, K, 2), 0, 0>(e)), ...);
}
// b_k for a lane-dependent k (a select chain over the constants)
template <int NOP, int... K>
__device__ __forceinline__ double slot_b_at(int k, std::integer_sequence<int, K...>)
{
    double v = 0.0;
    ((v = (k == K) ? slot_b(NOP, K) : v), ...);
    return v;
}

constexpr int kOpLanes = 16, kOpNodes = 64 / kOpLanes;

// neighbour j's coordinates and volume, loaded a chunk ahead of their use
struct NbrRow { double x[3]; double v; };
__device__ __forceinline__ void load_nbr(int j1, const double *xp, int dim, const double *vols, NbrRow &o)
{
    const int j = j1 - 1;
    o.x[0] = xp[(size_t)j * dim];
    o.x[1] = xp[(size_t)j * dim + 1];
    o.x[2] = (dim == 3) ? xp[(size_t)j * dim + 2] : 0.0;
    o.v = vols[j];
}

// weight bw = w vc V_j of bond (i, j) and its scaled offset e = xi / delta
// (bw = 0 for a bond that contributes nothing)
__device__ __forceinline__ void bond_geometry(const double (&xi)[3], const NbrRow &nj, int dim, double delta,
                                              double inv_delta, double dx_char, int volcorr, double &bw,
                                              double (&e)[3])
{
    const double xx = nj.x[0] - xi[0];
    const double xy = nj.x[1] - xi[1];
    const double xz = (dim == 3) ? nj.x[2] - xi[2] : 0.0;
    const double xn = sqrt(xx * xx + xy * xy + xz * xz);
    bw = 0.0;
    e[0] = e[1] = e[2] = 0.0;
    if (!(xn > 0.0)) return;
    // computeWeight / partialVolumeCorrection (CalcPDOperators.cpp:48-60)
    const double r = 2.0 * xn / delta;
    const double w = exp(-r * r);
    double vc = 1.0;
    if (volcorr && dx_char > 0.0 && !(xn <= delta - 0.5 * dx_char)) {
        vc = (delta - xn + 0.5 * dx_char) / dx_char;
        vc = vc < 0.0 ? 0.0 : (vc > 1.0 ? 1.0 : vc);
    }
    bw = w * vc * nj.v;
    if (!(bw > 0.0)) { bw = 0.0; return; }
    e[0] = xx * inv_delta; e[1] = xy * inv_delta; e[2] = xz * inv_delta;
}

// The block is one wavefront, whose LDS instructions execute in issue order:
// a barrier between its LDS writes and reads only has to stop the compiler
// from moving LDS accesses across it and wait for the wave's LDS traffic.
// Unlike __syncthreads it leaves the wave's global loads in flight, so the
// next chunk's neighbour rows load under the current chunk's work.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// monomials eta^{q_k} of a scaled offset (all zero for bw = 0)
template <int NOP>
__device__ __forceinline__ void monomials(double bw, const double (&e)[3], double *mono)
{
    monomials_seq<NOP>(e, mono, std::make_integer_sequence<int, NOP>{});
#pragma unroll
    for (int k = 0; k < NOP; k++) mono[k] = (bw > 0.0) ? mono[k] : 0.0;
}

// maximum over the 16 lanes of a node (xor partners stay inside the group)
__device__ __forceinline__ double group_max(double v)
{
#pragma unroll
    for (int m = kOpLanes / 2; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
    return v;
}

template <int NOP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_operators(int n, int dim, const double *xp, const double *vols, const int *off, const int *nbr,
            double delta, double dx_char, int volcorr, double *ops, unsigned char *flags, int *ndeg)
{
    using Seq = std::make_integer_sequence<int, NOP>;
    constexpr int NN = NOP * NOP;
    constexpr int L = kOpLanes;
    constexpr int NE = (NN + L - 1) / L;       // matrix entries per lane (LU, residual)
    constexpr int BR = 2 * NOP;                // staged doubles per bond: p_a = bw m_a, then m_a
    constexpr int NST = (64 * BR > 3 * kOpNodes * NN) ? 64 * BR : 3 * kOpNodes * NN;
    // staged bonds; after the moment sums the nodes' A, X and LU factors
    __shared__ __align__(16) double s_stage[NST];
    __shared__ int s_rowT[kOpNodes][NOP], s_colT[kOpNodes][NOP];

    const int lane = threadIdx.x, g = lane / L, l = lane % L;
    const int i = blockIdx.x * kOpNodes + g;
    const bool valid = i < n;
    const int b0 = valid ? off[i] : 0, b1 = valid ? off[i + 1] : 0;
    const double inv_delta = 1.0 / delta;
    double *A = s_stage + g * NN;
    double *X = s_stage + (kOpNodes + g) * NN;
    double *F = s_stage + (2 * kOpNodes + g) * NN;
    // LU and residual: the lane's matrix entries are e = l + m L (m < NE,
    // e < NN), row e / NOP and column e % NOP (recomputed where used: fewer
    // live registers)

    // ---- 1+2: moment matrix, bit-identical accumulation order ----------
    int nch = (b1 - b0 + L - 1) / L;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nch = max(nch, __shfl_xor(nch, m));
    // lane c < NOP of a node sums column c of A: A[r][c] += p_r m_c, with
    // p_r = bw m_r formed once per bond as the reference's left-to-right
    // product (bw m_r) m_c does, so the sums are bit-identical; the p_r are
    // broadcast reads shared by the node's lanes
    double acc[NOP];
#pragma unroll
    for (int r = 0; r < NOP; r++) acc[r] = 0.0;
    int usable = 0;
    double xi[3] = {0.0, 0.0, 0.0};
    if (valid) {
        xi[0] = xp[(size_t)i * dim];
        xi[1] = xp[(size_t)i * dim + 1];
        if (dim == 3) xi[2] = xp[(size_t)i * dim + 2];
    }
    // neighbour rows one chunk ahead, neighbour ids two chunks ahead
    NbrRow nx = {};
    int jnext = 1;
    if (b0 + l < b1) load_nbr(nbr[b0 + l], xp, dim, vols, nx);
    if (b0 + l + L < b1) jnext = nbr[b0 + l + L];
    for (int ch = 0; ch < nch; ch++) {
        const int s = b0 + ch * L, t = s + l;
        const NbrRow cur = nx;
        if (t + L < b1) load_nbr(jnext, xp, dim, vols, nx);
        if (t + 2 * L < b1) jnext = nbr[t + 2 * L];
        double bw = 0.0, mono[NOP];
#pragma unroll
        for (int k = 0; k < NOP; k++) mono[k] = 0.0;
        if (t < b1) {
            double e[3];
            bond_geometry(xi, cur, dim, delta, inv_delta, dx_char, volcorr, bw, e);
            monomials<NOP>(bw, e, mono);
        }
#pragma unroll
        for (int k = 0; k < NOP; k++) {
            s_stage[lane * BR + k] = bw * mono[k];
            s_stage[lane * BR + NOP + k] = mono[k];
        }
        usable += __popcll((__ballot(bw > 0.0) >> (g * L)) & 0xFFFFull);
        wave_lds_sync();
        // An unusable bond has bw = 0 and zero monomials, so its term is
        // +-0 and adding it leaves the sum bit-unchanged: no branch needed.
        const int cnt = min(L, b1 - s);
        if (l < NOP) {
#pragma unroll 1
            for (int u = 0; u < cnt; u++) {
                const double *pu = s_stage + (g * L + u) * BR;
                const double mc = pu[NOP + l];
#pragma unroll
                for (int r = 0; r < NOP; r++) acc[r] += pu[r] * mc;
            }
        }
        wave_lds_sync();
    }
    if (l < NOP) {
#pragma unroll
        for (int r = 0; r < NOP; r++) { A[r * NOP + l] = acc[r]; F[r * NOP + l] = acc[r]; }
    }
    wave_lds_sync();

    // ---- 3: complete-pivot LU (FullPivLU.h:360-439) ---------------------
    int nonzero = NOP;
    double maxpiv = 0.0;
    bool done = !valid;
    for (int k = 0; k < NOP; k++) {
        // biggest |a| of the corner, ties -> first in column-major order
        double best = -1.0; int bidx = 1 << 30;
        if (!done) {
#pragma unroll
            for (int m = 0; m < NE; m++) {
                const int e = l + m * L, r = e / NOP, c = e % NOP;
                if (e < NN && r >= k && c >= k) {
                    const double sc = fabs(F[l + m * L]);
                    const int ci = c * NOP + r;
                    if (sc > best || (sc == best && ci < bidx)) { best = sc; bidx = ci; }
                }
            }
        }
#pragma unroll
        for (int m = L / 2; m >= 1; m >>= 1) {
            const double ob = __shfl_xor(best, m);
            const int oi = __shfl_xor(bidx, m);
            if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
        }
        if (!done && best == 0.0) {
            done = true;
            nonzero = k;
            if (l == 0)
                for (int q = k; q < NOP; q++) { s_rowT[g][q] = q; s_colT[g][q] = q; }
        }
        const int br = bidx % NOP, bc = bidx / NOP;
        if (!done) {
            if (best > maxpiv) maxpiv = best;
            if (l == 0) { s_rowT[g][k] = br; s_colT[g][k] = bc; }
            if (br != k && l < NOP) {
                const double t = F[k * NOP + l];
                F[k * NOP + l] = F[br * NOP + l];
                F[br * NOP + l] = t;
            }
        }
        wave_lds_sync();
        if (!done && bc != k && l < NOP) {
            const double t = F[l * NOP + k];
            F[l * NOP + k] = F[l * NOP + bc];
            F[l * NOP + bc] = t;
        }
        wave_lds_sync();
        if (k < NOP - 1) {
            if (!done) {
                const double piv = F[k * NOP + k];
                if (l > k && l < NOP) F[l * NOP + k] /= piv;
            }
            wave_lds_sync();
            if (!done) {
#pragma unroll
                for (int m = 0; m < NE; m++) {
                    const int e = l + m * L, r = e / NOP, c = e % NOP;
                    if (e < NN && r > k && c > k) F[e] -= F[r * NOP + k] * F[k * NOP + c];
                }
            }
            wave_lds_sync();
        }
    }

    // ---- 4: column-parallel solve (FullPivLU.h:594-628) ----------------
    const double thr = fabs(maxpiv) * ((double)NOP * DBL_EPSILON);
    int rank = 0;
    for (int q = 0; q < nonzero; q++) rank += (fabs(F[q * NOP + q]) > thr);
    if (valid && l < NOP) {
        const int j = l;
        double c[NOP];
#pragma unroll
        for (int r = 0; r < NOP; r++) c[r] = (r == j) ? slot_b(NOP, r) : 0.0;
        if (rank > 0) {
            for (int k = 0; k < NOP; k++) {
                const int r = s_rowT[g][k];
                if (r != k) {
                    double tk = 0.0, tr = 0.0;
#pragma unroll
                    for (int qq = 0; qq < NOP; qq++) { if (qq == k) tk = c[qq]; if (qq == r) tr = c[qq]; }
#pragma unroll
                    for (int qq = 0; qq < NOP; qq++) { if (qq == k) c[qq] = tr; if (qq == r) c[qq] = tk; }
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
                        for (int k = 0; k < j2; k++) accd += c[k] * F[r * NOP + k];
                        c[r] = c[r] + (-1.0 * accd);
                    }
                }
#pragma unroll
                for (int r = j2; r < j2 + w; r++)
#pragma unroll
                    for (int k = j2; k < r; k++) c[r] -= c[k] * F[r * NOP + k];
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
                        for (int k = 0; k < NOP; k++) if (k >= po && k < sz) accd += c[k] * F[r * NOP + k];
                        c[r] = c[r] + (-1.0 * accd);
                    }
                }
#pragma unroll
                for (int r = NOP - 1; r >= 0; r--) {
                    if (r < j2 || r >= po) continue;
#pragma unroll
                    for (int k = 0; k < NOP; k++) if (k > r && k < po) c[r] -= c[k] * F[r * NOP + k];
                    const double inv = 1.0 / F[r * NOP + r];
                    c[r] *= inv;
                }
            }
        }
        // undo the column permutation
        int q[NOP];
#pragma unroll
        for (int r = 0; r < NOP; r++) q[r] = r;
        for (int k = 0; k < NOP; k++) {
            const int ck = s_colT[g][k];
            int qk = 0, qc = 0;
#pragma unroll
            for (int r = 0; r < NOP; r++) { if (r == k) qk = q[r]; if (r == ck) qc = q[r]; }
#pragma unroll
            for (int r = 0; r < NOP; r++) { if (r == k) q[r] = qc; if (r == ck) q[r] = qk; }
        }
#pragma unroll
        for (int r = 0; r < NOP; r++) X[q[r] * NOP + j] = (rank > 0 && r < rank) ? c[r] : 0.0;
    }
    wave_lds_sync();

    // ---- 5: consistency check (MatrixXd.cpp:37-47) ----------------------
    double resid = 0.0, bmax = 0.0, amax = 0.0, xmax = 0.0;
#pragma unroll 1   // unrolled, the entries' row and column loads are all hoisted
    for (int m = 0; m < NE; m++) {
        const int e = l + m * L, r = e / NOP, cc = e % NOP;
        if (e < NN) {
            double sacc = 0.0;
#pragma unroll
            for (int k = 0; k < NOP; k++) sacc += A[r * NOP + k] * X[k * NOP + cc];
            const double bv = (r == cc) ? slot_b_at<NOP>(r, Seq{}) : 0.0;
            resid = fmax(resid, fabs(sacc - bv));
            bmax = fmax(bmax, fabs(bv));
            amax = fmax(amax, fabs(A[l + m * L]));
            xmax = fmax(xmax, fabs(X[l + m * L]));
        }
    }
    resid = group_max(resid); bmax = group_max(bmax); amax = group_max(amax); xmax = group_max(xmax);
    const double scale = fmax(bmax, amax * xmax);
    const bool consistent = resid <= 1.0e-8 * fmax(scale, DBL_MIN);
    if (valid && l == 0) {
        unsigned char f = 0;
        if (usable < NOP) f |= 2;
        if (usable < NOP || !consistent) { f |= 1; atomicAdd(ndeg, 1); }
        flags[i] = f;
    }

    // ---- 6: per-bond operators (CalcPDOperators.cpp:307-354) ------------
    double idp[NOP];
    inv_powers_seq<NOP>(inv_delta, idp, Seq{});
    nx = NbrRow{};
    jnext = 1;
    if (b0 + l < b1) load_nbr(nbr[b0 + l], xp, dim, vols, nx);
    if (b0 + l + L < b1) jnext = nbr[b0 + l + L];
    for (int t = b0 + l; t < b1; t += L) {
        const NbrRow cur = nx;
        if (t + L < b1) load_nbr(jnext, xp, dim, vols, nx);
        if (t + 2 * L < b1) jnext = nbr[t + 2 * L];
        double bw, e[3], mono[NOP];
        bond_geometry(xi, cur, dim, delta, inv_delta, dx_char, volcorr, bw, e);
        monomials<NOP>(bw, e, mono);
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
                for (int a = 0; a < NOP; a++) sacc += X[a * NOP + k] * mono[a];
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
    const int nop = (dim == 3) ? 9 : 5;        // order-2 slots (the kernel's slot table)
    // the operators are rewritten: every element built on the old ones is stale
    invalidate(ctx, INV_OPS);
    ctx->nop = 0;
    PXG_HIP(ctx->ops.alloc((size_t)ctx->nb * nop));
    PXG_HIP(ctx->flags.alloc(ctx->n));
    dbuf<int> nd;
    PXG_HIP(nd.alloc(1));
    PXG_HIP(hipMemsetAsync(nd.p, 0, sizeof(int), st));
    const dim3 grid((unsigned)((ctx->n + kOpNodes - 1) / kOpNodes));
    if (nop == 5)
        hipLaunchKernelGGL(k_operators<5>, grid, dim3(64), 0, st, ctx->n, dim, ctx->xp.p,
                           ctx->vols.p, ctx->off.p, ctx->nbr.p, delta, dx_char, dx_char > 0.0 ? 1 : 0,
                           ctx->ops.p, ctx->flags.p, nd.p);
    else
        hipLaunchKernelGGL(k_operators<9>, grid, dim3(64), 0, st, ctx->n, dim, ctx->xp.p,
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

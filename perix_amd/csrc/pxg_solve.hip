// pxg_solve.hip -- direct solve of the implicit Newton system K dU = RHS on
// the device, and the Newton update that keeps the iterate there.
//
// The reference's default linear solver (src/LinearSolver/DefaultSolver.cpp)
// orders the equations by reverse Cuthill-McKee (computeOrdering, :52-150),
// takes for every equation j the envelope [first(j), j) of row j and column
// j together (buildProfile, :164-194) and factors A = L D U without pivoting,
// column by column inside that envelope (factorize, :287-370), then
// substitutes (solveFactored, :372-424).  Its arithmetic is a chain of dot
// products, one equation at a time.
//
// Here the same ordering and envelope are cut into SB x SB dense blocks and
// factored right-looking (A = L U, unit L; mathematically the same factors,
// D folded into U):
//   FB[I]   first block column of block row I's envelope = min over its rows
//           of first(j) / SB; block row I of L holds blocks (I, FB[I]..I-1),
//           block column I of U holds (FB[I]..I-1, I) -- fill-in stays inside
//           (an entry outside a row's envelope has only zero products);
//   step K  (1) factor the diagonal block D_K = L_KK U_KK in LDS and form
//           L_KK^-1 and U_KK^-1; (2) panels L_IK = A_IK U_KK^-1 and
//           U_KI = L_KK^-1 A_KI for the active set {I > K : FB[I] <= K};
//           (3) trailing update A_IJ -= L_IK U_KJ over active x active.
// Each step's work is SB x SB x SB products (the panel and update kernels are
// plain FP64 GEMM tiles, one workgroup per 16-row strip); the chain of n
// pivots is the serial part, as it is in the reference.  When the update has
// many blocks, the next step's diagonal factor runs inside it (k_update_diag).
// The per-step launch chains depend only on the pattern and are replayed as
// HIP graphs.  Substitution walks the blocks with
// the inverted diagonal factors: forward y_K = L_KK^-1 y_K, then y_I -= L_IK
// y_K for the active I; backward x_K = U_KK^-1 z_K, then z_I -= U_IK x_K for
// I in [FB[K], K).  The orderings (RCM, envelope) are computed on the host
// once per CSR pattern, exactly as the reference does (std::sort on degree,
// so equal degrees fall in the same order).
#include <algorithm>
#include <climits>
#include <cmath>
#include <queue>
#include <vector>
#include "pxg_internal.h"

using namespace pxg;

namespace {

constexpr int SB = kSolveBlock;
constexpr int SB2 = SB * SB;
constexpr int kThreads = 256;
constexpr int TS = SB + 2;            // LDS row stride of GEMM tiles (16-byte aligned rows)

__global__ __launch_bounds__(256) void k_fill(double *p, size_t count, double v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(64) void k_fill_int(int *p, int count, int v)
{
    const int i = (int)threadIdx.x;
    if (i < count) p[i] = v;
}

// CSR values into the block storage; the padded equations n..nb*SB-1 get a
// unit diagonal so that their blocks factor trivially
__global__ __launch_bounds__(256) void k_scatter(long long nnz, const double *__restrict__ vals,
                                                 const long long *__restrict__ loc, double *F, int n, int npad,
                                                 double *D)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) F[loc[i]] = vals[i];
    const long long e = n + i;
    if (e < npad) D[(e / SB) * SB2 + (e % SB) * (SB + 1)] = 1.0;
}

// (1) D_K = L U in place (unit L below the diagonal, U on and above), with
// L^-1 = X (unit lower) and U^-1 = Y (upper), blocked in 8 x 8 sub-blocks
// (rows and columns b0..b0+7 of round kb), all three matrices in LDS:
//   (a) wave 0 factors the diagonal sub-block, lane r holding its row r in
//       registers, the pivot row read across lanes (readlane); its unit-L
//       and U inverses (Lx, Uy) come out of the same walk;
//   (b) the panels A_i,kb = A_i,kb Uy and A_kb,j = Lx A_kb,j, and the
//       inverses' own block row / column X_kb,: = Lx X_kb,: and
//       Y_:,kb = Y_:,kb Uy, one thread per row or column;
//   (c) the trailing updates A_ij -= A_i,kb A_kb,j, X_i,: -= A_i,kb X_kb,:
//       and Y_:,j -= Y_:,kb A_kb,j, four consecutive entries per item.
// A zero pivot records the least failing equation (new order) in *zpiv and
// continues with 1, so the arithmetic stays finite; the host reports it.
// The pivots stay a chain of 64 per block: the walks take about 13 us of a
// block's 29 us, the panels and updates the rest.
constexpr int BD = 8;                 // sub-block edge of the blocked diagonal factor

__device__ __forceinline__ double read_lane(double v, int lane)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_readlane(hi, lane), __builtin_amdgcn_readlane(lo, lane));
}

__device__ __forceinline__ void diag_factor(int K, double *__restrict__ Dall, double *__restrict__ Liall,
                                                    double *__restrict__ Uiall, int *zpiv)
{
    __shared__ __align__(16) double A[SB * TS], X[SB * TS], Y[SB * TS];
    __shared__ double Lx[BD * BD], Uy[BD * BD];
    const int t = (int)threadIdx.x;
    double *D = Dall + (size_t)K * SB2;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = 2 * (m * kThreads + t), r = e / SB, c = e % SB;
        *reinterpret_cast<double2 *>(A + r * TS + c) = *reinterpret_cast<const double2 *>(D + e);
        *reinterpret_cast<double2 *>(X + r * TS + c) = make_double2(r == c ? 1.0 : 0.0, r == c + 1 ? 1.0 : 0.0);
        *reinterpret_cast<double2 *>(Y + r * TS + c) = make_double2(r == c ? 1.0 : 0.0, r == c + 1 ? 1.0 : 0.0);
    }
    __syncthreads();
    int zfirst = INT_MAX;
    for (int kb = 0; kb < SB / BD; kb++) {
        const int b0 = BD * kb, nr = SB - b0 - BD;          // rows (columns) after the sub-block
        // (a) the diagonal sub-block, wave 0 (lanes 0..7 carry its rows)
        if (t < 64) {
            const int r = t & (BD - 1);
            double a[BD], x[BD], y[BD];
#pragma unroll
            for (int j = 0; j < BD; j++) {
                a[j] = A[(b0 + r) * TS + b0 + j];
                x[j] = (r == j) ? 1.0 : 0.0;
                y[j] = x[j];
            }
#pragma unroll
            for (int k = 0; k < BD; k++) {
                double p = read_lane(a[k], k);
                if (p == 0.0) {
                    zfirst = min(zfirst, K * SB + b0 + k);
                    p = 1.0;
                }
                double rinv = __builtin_amdgcn_rcp(p);
                rinv = fma(fma(-p, rinv, 1.0), rinv, rinv);
                rinv = fma(fma(-p, rinv, 1.0), rinv, rinv);
                double uk[BD], xk[BD];
#pragma unroll
                for (int j = 0; j < BD; j++) {
                    uk[j] = j > k ? read_lane(a[j], k) : 0.0;
                    xk[j] = j <= k ? read_lane(x[j], k) : 0.0;
                }
                const double l = (r > k) ? a[k] * rinv : 0.0;
                const double yk = y[k] * rinv;
#pragma unroll
                for (int j = 0; j < BD; j++) {
                    if (j > k) {
                        a[j] = fma(-l, uk[j], a[j]);
                        y[j] = fma(-yk, uk[j], y[j]);
                    } else {
                        x[j] = fma(-l, xk[j], x[j]);
                    }
                }
                if (r > k) a[k] = l;
                y[k] = yk;
            }
            if (t < BD) {
#pragma unroll
                for (int j = 0; j < BD; j++) {
                    A[(b0 + r) * TS + b0 + j] = a[j];
                    Lx[r * BD + j] = x[j];
                    Uy[r * BD + j] = y[j];
                }
            }
        }
        __syncthreads();
        // (b) panels and the inverses' block row / column: 2 nr + 2 (b0 + 8) <= 128 items
        {
            double v[BD], o[BD];
            int kind = -1, idx = 0;
            if (t < nr) { kind = 0; idx = b0 + BD + t; }                               // row of A below
            else if (t < 2 * nr) { kind = 1; idx = b0 + BD + t - nr; }                 // column of A to the right
            else if (t < 2 * nr + b0 + BD) { kind = 2; idx = t - 2 * nr; }             // column of X (< b0+8)
            else if (t < 2 * nr + 2 * (b0 + BD)) { kind = 3; idx = t - 2 * nr - b0 - BD; }  // row of Y (< b0+8)
            if (kind == 0 || kind == 3) {
                double *M = (kind == 0) ? A : Y;
#pragma unroll
                for (int m = 0; m < BD; m++) v[m] = M[idx * TS + b0 + m];
#pragma unroll
                for (int c = 0; c < BD; c++) {
                    double acc = 0.0;
#pragma unroll
                    for (int m = 0; m < BD; m++) acc = fma(v[m], Uy[m * BD + c], acc);
                    o[c] = acc;
                }
            } else if (kind == 1 || kind == 2) {
                double *M = (kind == 1) ? A : X;
#pragma unroll
                for (int m = 0; m < BD; m++) v[m] = M[(b0 + m) * TS + idx];
#pragma unroll
                for (int c = 0; c < BD; c++) {
                    double acc = 0.0;
#pragma unroll
                    for (int m = 0; m < BD; m++) acc = fma(Lx[c * BD + m], v[m], acc);
                    o[c] = acc;
                }
            }
            __syncthreads();
            if (kind == 0 || kind == 3) {
                double *M = (kind == 0) ? A : Y;
#pragma unroll
                for (int c = 0; c < BD; c++) M[idx * TS + b0 + c] = o[c];
            } else if (kind == 1 || kind == 2) {
                double *M = (kind == 1) ? A : X;
#pragma unroll
                for (int c = 0; c < BD; c++) M[(b0 + c) * TS + idx] = o[c];
            }
        }
        __syncthreads();
        // (c) trailing updates, four consecutive entries per item:
        //   A rows/cols >= b0+8; X rows >= b0+8, cols < b0+8; Y rows < b0+8, cols >= b0+8
        if (nr > 0) {
            const int q4 = nr / 4, qx = (b0 + BD) / 4;
            const int nA = nr * q4, nX = nr * qx, nY = (b0 + BD) * q4;
            for (int w = t; w < nA + nX + nY; w += kThreads) {
                const double *Lf;   // left factor row (8 values in the kb column block)
                const double *Rf;   // right factor rows b0..b0+7, four columns
                double *Cp;
                if (w < nA) {
                    const int i = b0 + BD + w / q4, j = b0 + BD + 4 * (w % q4);
                    Lf = A + i * TS + b0; Rf = A + b0 * TS + j; Cp = A + i * TS + j;
                } else if (w < nA + nX) {
                    const int w2 = w - nA, i = b0 + BD + w2 / qx, j = 4 * (w2 % qx);
                    Lf = A + i * TS + b0; Rf = X + b0 * TS + j; Cp = X + i * TS + j;
                } else {
                    const int w3 = w - nA - nX, i = w3 / q4, j = b0 + BD + 4 * (w3 % q4);
                    Lf = Y + i * TS + b0; Rf = A + b0 * TS + j; Cp = Y + i * TS + j;
                }
                double2 c01 = *reinterpret_cast<const double2 *>(Cp), c23 = *reinterpret_cast<const double2 *>(Cp + 2);
#pragma unroll
                for (int m = 0; m < BD; m++) {
                    const double lv = Lf[m];
                    const double2 r01 = *reinterpret_cast<const double2 *>(Rf + m * TS);
                    const double2 r23 = *reinterpret_cast<const double2 *>(Rf + m * TS + 2);
                    c01.x = fma(-lv, r01.x, c01.x);
                    c01.y = fma(-lv, r01.y, c01.y);
                    c23.x = fma(-lv, r23.x, c23.x);
                    c23.y = fma(-lv, r23.y, c23.y);
                }
                *reinterpret_cast<double2 *>(Cp) = c01;
                *reinterpret_cast<double2 *>(Cp + 2) = c23;
            }
        }
        __syncthreads();
    }
    if (t == 0 && zfirst != INT_MAX) atomicMin(zpiv, zfirst);
    double *Li = Liall + (size_t)K * SB2, *Ui = Uiall + (size_t)K * SB2;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = 2 * (m * kThreads + t), r = e / SB, c = e % SB;
        *reinterpret_cast<double2 *>(D + e) = *reinterpret_cast<const double2 *>(A + r * TS + c);
        *reinterpret_cast<double2 *>(Li + e) = *reinterpret_cast<const double2 *>(X + r * TS + c);
        *reinterpret_cast<double2 *>(Ui + e) = *reinterpret_cast<const double2 *>(Y + r * TS + c);
    }
}

__global__ __launch_bounds__(256) void k_diag(int K, double *__restrict__ Dall, double *__restrict__ Liall,
                                              double *__restrict__ Uiall, int *zpiv)
{
    diag_factor(K, Dall, Liall, Uiall, zpiv);
}

// Strip products of SB x SB row-major blocks, one workgroup per 16-wide
// strip of the output (1024 entries, 4 per thread), every global load issued
// before the first use.  C may alias the operand whose strip is read: the
// workgroup reads only the strip it writes from that operand.
constexpr int STRIP = 16;
constexpr int NSTRIP = SB / STRIP;

// rows [s0, s0 + 16) of C (op)= A B: thread t owns row s0 + t / 16, columns
// 4 (t % 16)..; A's strip and all of B go to LDS
template <bool SUB>
__device__ __forceinline__ void row_strip_gemm(const double *A, const double *B, double *C, int s0, double *As,
                                               double *Bs)
{
    const int t = (int)threadIdx.x, r = t >> 4, c0 = 4 * (t & 15);
    double2 va[2], vb[8], vc[2];
#pragma unroll
    for (int m = 0; m < 2; m++) va[m] = *reinterpret_cast<const double2 *>(A + (size_t)s0 * SB + 2 * (m * kThreads + t));
#pragma unroll
    for (int m = 0; m < 8; m++) vb[m] = *reinterpret_cast<const double2 *>(B + 2 * (m * kThreads + t));
    if (SUB) {
        vc[0] = *reinterpret_cast<const double2 *>(C + (size_t)(s0 + r) * SB + c0);
        vc[1] = *reinterpret_cast<const double2 *>(C + (size_t)(s0 + r) * SB + c0 + 2);
    }
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const int e = 2 * (m * kThreads + t);
        *reinterpret_cast<double2 *>(As + (e / SB) * TS + e % SB) = va[m];
    }
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = 2 * (m * kThreads + t);
        *reinterpret_cast<double2 *>(Bs + (e / SB) * TS + e % SB) = vb[m];
    }
    __syncthreads();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int k = 0; k < SB; k++) {
        const double av = As[r * TS + k];
        const double2 b01 = *reinterpret_cast<const double2 *>(Bs + k * TS + c0);
        const double2 b23 = *reinterpret_cast<const double2 *>(Bs + k * TS + c0 + 2);
        acc[0] = fma(av, b01.x, acc[0]);
        acc[1] = fma(av, b01.y, acc[1]);
        acc[2] = fma(av, b23.x, acc[2]);
        acc[3] = fma(av, b23.y, acc[3]);
    }
    double *row = C + (size_t)(s0 + r) * SB + c0;
    if (SUB) {
        *reinterpret_cast<double2 *>(row) = make_double2(vc[0].x - acc[0], vc[0].y - acc[1]);
        *reinterpret_cast<double2 *>(row + 2) = make_double2(vc[1].x - acc[2], vc[1].y - acc[3]);
    } else {
        *reinterpret_cast<double2 *>(row) = make_double2(acc[0], acc[1]);
        *reinterpret_cast<double2 *>(row + 2) = make_double2(acc[2], acc[3]);
    }
}

// columns [s0, s0 + 16) of C = A B: thread t owns column s0 + t % 16, rows
// 4 (t / 16)..; all of A (transposed) and B's strip go to LDS
__device__ __forceinline__ void col_strip_gemm(const double *A, const double *B, double *C, int s0, double *At,
                                               double *Bs)
{
    const int t = (int)threadIdx.x, c = t & 15, r0 = 4 * (t >> 4);
    double2 va[8], vb[2];
#pragma unroll
    for (int m = 0; m < 8; m++) va[m] = *reinterpret_cast<const double2 *>(A + 2 * (m * kThreads + t));
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const int e = m * kThreads + t, k = e / (STRIP / 2), cp = e % (STRIP / 2);
        vb[m] = *reinterpret_cast<const double2 *>(B + (size_t)k * SB + s0 + 2 * cp);
    }
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = 2 * (m * kThreads + t), rr = e / SB, kk = e % SB;
        At[kk * TS + rr] = va[m].x;
        At[(kk + 1) * TS + rr] = va[m].y;
    }
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const int e = m * kThreads + t, k = e / (STRIP / 2), cp = e % (STRIP / 2);
        *reinterpret_cast<double2 *>(Bs + k * (STRIP + 2) + 2 * cp) = vb[m];
    }
    __syncthreads();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int k = 0; k < SB; k++) {
        const double2 a01 = *reinterpret_cast<const double2 *>(At + k * TS + r0);
        const double2 a23 = *reinterpret_cast<const double2 *>(At + k * TS + r0 + 2);
        const double bv = Bs[k * (STRIP + 2) + c];
        acc[0] = fma(a01.x, bv, acc[0]);
        acc[1] = fma(a01.y, bv, acc[1]);
        acc[2] = fma(a23.x, bv, acc[2]);
        acc[3] = fma(a23.y, bv, acc[3]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) C[(size_t)(r0 + i) * SB + s0 + c] = acc[i];
}

struct Layout {
    double *F;                        // [0, ns) L strips, [ns, 2ns) U strips, [2ns, 2ns+nb) diagonal
    long long ns;
    const int *fb;                    // FB[I]
    const long long *lp;              // first strip block of block row / column I
    __device__ double *L(int I, int K) const { return F + (size_t)(lp[I] + (K - fb[I])) * SB2; }
    __device__ double *U(int K, int J) const { return F + (size_t)(ns + lp[J] + (K - fb[J])) * SB2; }
    __device__ double *D(int I) const { return F + (size_t)(2 * ns + I) * SB2; }
};

// (2) workgroup w: block w / NSTRIP of the panel list, strip w % NSTRIP.
// First na blocks: L_IK = A_IK U_KK^-1 by row strips; next na: U_KJ =
// L_KK^-1 A_KJ by column strips (each in place, see the strip products)
__global__ __launch_bounds__(256) void k_panel(int K, const int *__restrict__ act, int na, Layout Y,
                                               const double *__restrict__ Li, const double *__restrict__ Ui)
{
    __shared__ __align__(16) double As[SB * TS];
    __shared__ __align__(16) double Bs[SB * TS];
    const int blk = (int)blockIdx.x / NSTRIP, s0 = STRIP * ((int)blockIdx.x % NSTRIP);
    if (blk < na) {
        double *c = Y.L(act[blk], K);
        row_strip_gemm<false>(c, Ui + (size_t)K * SB2, c, s0, As, Bs);
    } else {
        double *c = Y.U(K, act[blk - na]);
        col_strip_gemm(Li + (size_t)K * SB2, c, c, s0, As, Bs);
    }
}

// (3) A_IJ -= L_IK U_KJ for (I, J) in active x active, one row strip per workgroup
__global__ __launch_bounds__(256) void k_update(int K, const int *__restrict__ act, int na, Layout Y)
{
    __shared__ __align__(16) double As[STRIP * TS];
    __shared__ __align__(16) double Bs[SB * TS];
    const int blk = (int)blockIdx.x / NSTRIP, s0 = STRIP * ((int)blockIdx.x % NSTRIP);
    const int I = act[blk / na], J = act[blk % na];
    double *c = (I > J) ? Y.L(I, J) : (I < J) ? Y.U(I, J) : Y.D(I);
    row_strip_gemm<true>(Y.L(I, K), Y.U(K, J), c, s0, As, Bs);
}

// (3') the trailing update of step K when block K+1 is active (act[0] == K+1),
// fused with the next step's diagonal factor: workgroup 0 applies all four
// strips of A_{K+1,K+1} -= L_{K+1,K} U_{K,K+1} itself and then factors
// D_{K+1} (as k_diag), while the other workgroups update the remaining tiles;
// the next diagonal factor so runs beside this step's update instead of after it.
__global__ __launch_bounds__(256) void k_update_diag(int K, const int *__restrict__ act, int na, Layout Y,
                                                     double *__restrict__ Dall, double *__restrict__ Liall,
                                                     double *__restrict__ Uiall, int *zpiv)
{
    __shared__ __align__(16) double As[STRIP * TS];
    __shared__ __align__(16) double Bs[SB * TS];
    if (blockIdx.x == 0) {
        // D_{K+1} -= L_{K+1,K} U_{K,K+1}: U_{K,K+1} staged once, the four
        // 16-row strips of L and of D loaded up front, then strip by strip
        const int I = act[0], t = (int)threadIdx.x, r = t >> 4, c0 = 4 * (t & 15);
        const double *Lb = Y.L(I, K), *Ub = Y.U(K, I);
        double *Db = Y.D(I);
        double2 vb[8], va[8], vc[8];
#pragma unroll
        for (int m = 0; m < 8; m++) {
            vb[m] = *reinterpret_cast<const double2 *>(Ub + 2 * (m * kThreads + t));
            va[m] = *reinterpret_cast<const double2 *>(Lb + 2 * (m * kThreads + t));
        }
#pragma unroll
        for (int q = 0; q < NSTRIP; q++) {
            vc[2 * q] = *reinterpret_cast<const double2 *>(Db + (size_t)(STRIP * q + r) * SB + c0);
            vc[2 * q + 1] = *reinterpret_cast<const double2 *>(Db + (size_t)(STRIP * q + r) * SB + c0 + 2);
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int e = 2 * (m * kThreads + t);
            *reinterpret_cast<double2 *>(Bs + (e / SB) * TS + e % SB) = vb[m];
        }
#pragma unroll
        for (int q = 0; q < NSTRIP; q++) {
            // va[2q], va[2q+1] hold strip q of L (rows 16q..16q+15)
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const int e = 2 * (m * kThreads + t);
                *reinterpret_cast<double2 *>(As + (e / SB) * TS + e % SB) = va[2 * q + m];
            }
            __syncthreads();
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
            for (int k = 0; k < SB; k++) {
                const double av = As[r * TS + k];
                const double2 b01 = *reinterpret_cast<const double2 *>(Bs + k * TS + c0);
                const double2 b23 = *reinterpret_cast<const double2 *>(Bs + k * TS + c0 + 2);
                acc[0] = fma(av, b01.x, acc[0]);
                acc[1] = fma(av, b01.y, acc[1]);
                acc[2] = fma(av, b23.x, acc[2]);
                acc[3] = fma(av, b23.y, acc[3]);
            }
            double *row = Db + (size_t)(STRIP * q + r) * SB + c0;
            *reinterpret_cast<double2 *>(row) = make_double2(vc[2 * q].x - acc[0], vc[2 * q].y - acc[1]);
            *reinterpret_cast<double2 *>(row + 2) = make_double2(vc[2 * q + 1].x - acc[2], vc[2 * q + 1].y - acc[3]);
            __syncthreads();          // LDS reuse; the strips' stores precede the factor's loads
        }
        diag_factor(I, Dall, Liall, Uiall, zpiv);
        return;
    }
    const int w = (int)blockIdx.x - 1, blk = 1 + w / NSTRIP, s0 = STRIP * (w % NSTRIP);
    const int I = act[blk / na], J = act[blk % na];
    double *c = (I > J) ? Y.L(I, J) : (I < J) ? Y.U(I, J) : Y.D(I);
    row_strip_gemm<true>(Y.L(I, K), Y.U(K, J), c, s0, As, Bs);
}

// Two SB x SB blocks into LDS (row stride TS) with every load issued up
// front (a caller with one block passes it twice).  256 threads, 8 double2
// each per block.
__device__ __forceinline__ void stage_two(const double *M, const double *N, double *Ms, double *Ns)
{
    const int t = (int)threadIdx.x;
    double2 vm[8], vn[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        vm[m] = *reinterpret_cast<const double2 *>(M + 2 * (m * kThreads + t));
        vn[m] = *reinterpret_cast<const double2 *>(N + 2 * (m * kThreads + t));
    }
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = 2 * (m * kThreads + t), r = e / SB, c = e % SB;
        *reinterpret_cast<double2 *>(Ms + r * TS + c) = vm[m];
        *reinterpret_cast<double2 *>(Ns + r * TS + c) = vn[m];
    }
}

// row r = t / 4 of M times v (LDS, SB values): each of the row's 4 threads
// sums 16 columns, the quad reduces with lane exchanges
__device__ __forceinline__ double row_dot(const double *Ms, const double *v)
{
    const int t = (int)threadIdx.x, r = t >> 2, q = t & 3;
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int c = 2 * q + 8 * m;
        const double2 mv = *reinterpret_cast<const double2 *>(Ms + r * TS + c);
        s = fma(mv.x, v[c], s);
        s = fma(mv.y, v[c + 1], s);
    }
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    return s;
}

// forward step K: z_K = L_KK^-1 y_K; y_I -= L_IK z_K for the active I
__global__ __launch_bounds__(256) void k_fwd(int K, const int *__restrict__ act, int na, Layout Y,
                                             const double *__restrict__ Li, double *y, double *zout)
{
    __shared__ __align__(16) double Ms[SB * TS], Ns[SB * TS];
    __shared__ double ys[SB], v[SB];
    const int w = (int)blockIdx.x, t = (int)threadIdx.x;
    const bool two = w < na;
    const int I = two ? act[w] : 0;
    const double *Lk = Li + (size_t)K * SB2;
    stage_two(Lk, two ? Y.L(I, K) : Lk, Ms, Ns);
    if (t < SB) ys[t] = y[(size_t)K * SB + t];
    __syncthreads();
    const double s = row_dot(Ms, ys);
    if ((t & 3) == 0) v[t >> 2] = s;
    __syncthreads();
    if (w == 0 && t < SB) zout[(size_t)K * SB + t] = v[t];
    if (!two) return;
    const double u = row_dot(Ns, v);
    if ((t & 3) == 0) y[(size_t)I * SB + (t >> 2)] -= u;
}

// backward step K: x_K = U_KK^-1 z_K; z_I -= U_IK x_K for I in [FB[K], K)
__global__ __launch_bounds__(256) void k_bwd(int K, Layout Y, const double *__restrict__ Ui, double *z,
                                             double *xout)
{
    __shared__ __align__(16) double Ms[SB * TS], Ns[SB * TS];
    __shared__ double zs[SB], v[SB];
    const int w = (int)blockIdx.x, t = (int)threadIdx.x;
    const int I = Y.fb[K] + w;
    const bool two = I < K;
    const double *Uk = Ui + (size_t)K * SB2;
    stage_two(Uk, two ? Y.U(I, K) : Uk, Ms, Ns);
    if (t < SB) zs[t] = z[(size_t)K * SB + t];
    __syncthreads();
    const double s = row_dot(Ms, zs);
    if ((t & 3) == 0) v[t >> 2] = s;
    __syncthreads();
    if (w == 0 && t < SB) xout[(size_t)K * SB + t] = v[t];
    if (!two) return;
    const double u = row_dot(Ns, v);
    if ((t & 3) == 0) z[(size_t)I * SB + (t >> 2)] -= u;
}

__global__ __launch_bounds__(256) void k_perm_in(int n, int npad, const int *__restrict__ inv,
                                                 const double *__restrict__ b, double *y)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) y[inv[i]] = b[i];
    else if (i < npad) y[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_perm_out(int n, const int *__restrict__ inv, const double *__restrict__ w,
                                                  double *x)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) x[i] = w[inv[i]];
}

// ---- Newton update (NonlinearSolver.cpp:108-125) ---------------------------
constexpr int kNormBlocks = 256;

// partial sums of squares of a and b (deterministic: fixed grid, fixed tree)
__global__ __launch_bounds__(256) void k_sumsq2(long long n, const double *__restrict__ a,
                                                const double *__restrict__ b, double *part)
{
    __shared__ double sa[4], sb[4];
    double xa = 0.0, xb = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        xa = fma(a[i], a[i], xa);
        xb = fma(b[i], b[i], xb);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        xa += __shfl_xor(xa, o, 64);
        xb += __shfl_xor(xb, o, 64);
    }
    const int wv = (int)threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { sa[wv] = xa; sb[wv] = xb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sa[0] + sa[1]) + (sa[2] + sa[3]);
        part[2 * blockIdx.x + 1] = (sb[0] + sb[1]) + (sb[2] + sb[3]);
    }
}

// norms[0] = |a|, norms[1] = |b| from the partials (one wave)
__global__ __launch_bounds__(64) void k_norm_final(int nparts, const double *__restrict__ part, double *norms)
{
    double xa = 0.0, xb = 0.0;
    for (int i = (int)threadIdx.x; i < nparts; i += 64) {
        xa += part[2 * i];
        xb += part[2 * i + 1];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        xa += __shfl_xor(xa, o, 64);
        xb += __shfl_xor(xb, o, 64);
    }
    if (threadIdx.x == 0) {
        norms[0] = sqrt(xa);
        norms[1] = sqrt(xb);
    }
}

// U += dU when |dU| is finite (the reference applies no update otherwise)
__global__ __launch_bounds__(256) void k_newton_add(long long n, const double *__restrict__ du,
                                                    const double *__restrict__ norms, double *u)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && isfinite(norms[1])) u[i] += du[i];
}

// BCSystem::presetSolution (BCSystem.cpp:114-129) from the resolved BC row
// program: Dirichlet reflection U(g) = 2 g(t) - U(b), direct U(g) = g(t)
// (DirichletBC.cpp:107-147), Neumann U(g) = U(b) (NeumannBC.cpp:88-105).
// `gate` (may be NULL) skips the preset when the Newton update was refused.
__global__ __launch_bounds__(256) void k_csr_preset(int count, const int *__restrict__ prow,
                                                    const int *__restrict__ ppdof,
                                                    const signed char *__restrict__ pkind,
                                                    const double *__restrict__ pval0,
                                                    const double *__restrict__ pvel, double t,
                                                    const double *__restrict__ gate, double *U)
{
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= count) return;
    if (gate && !isfinite(gate[1])) return;
    const int g = prow[q], kind = pkind[q];
    if (kind == BCROW_MIRROR) {
        U[g] = U[ppdof[q]];
    } else {
        const double value = pval0[q] + pvel[q] * t;
        U[g] = (kind == BCROW_REFLECT) ? 2.0 * value - U[ppdof[q]] : value;
    }
}

unsigned grid_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

// DefaultSolver.cpp:52-150: reverse Cuthill-McKee over the CSR pattern; BFS
// from the first least-degree unvisited vertex, new neighbours marked on
// enqueue and queued by ascending degree (std::sort, as the reference).
void rcm_order(int n, const std::vector<int> &rp, const std::vector<int> &cols, bool use_rcm, std::vector<int> &perm)
{
    perm.resize(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    if (!use_rcm || n == 0) return;
    std::vector<int> deg(n);
    for (int v = 0; v < n; v++) {
        int d = 0;
        for (int p = rp[v]; p < rp[v + 1]; p++) d += cols[p] != v;
        deg[v] = d;
    }
    std::vector<char> seen(n, 0);
    std::vector<int> order, fresh;
    order.reserve(n);
    std::queue<int> q;
    int from = 0;
    for (;;) {
        int s = -1, sd = INT_MAX;
        for (int v = from; v < n && sd > 0; v++)
            if (!seen[v] && deg[v] < sd) { s = v; sd = deg[v]; }
        if (s < 0) break;
        from = s + 1;
        seen[s] = 1;
        q.push(s);
        while (!q.empty()) {
            const int v = q.front();
            q.pop();
            order.push_back(v);
            fresh.clear();
            for (int p = rp[v]; p < rp[v + 1]; p++) {
                const int w = cols[p];
                if (w != v && !seen[w]) { seen[w] = 1; fresh.push_back(w); }
            }
            std::sort(fresh.begin(), fresh.end(), [&deg](int a, int b) { return deg[a] < deg[b]; });
            for (int w : fresh) q.push(w);
        }
    }
    if ((int)order.size() == n)
        for (int k = 0; k < n; k++) perm[k] = order[n - 1 - k];
}

// which pattern the solver factors: its own (pxg_solve_set_pattern) or the
// context's implicit CSR, each with its own generation
long long pattern_key(const pxg_ctx *ctx)
{
    return ctx->solve.own ? -1 - ctx->solve.own_gen : ctx->csr_gen;
}

int symbolic(pxg_ctx *ctx)
{
    SolveState &S = ctx->solve;
    S.ready = 0;
    S.drop_graphs();
    hipStream_t st = ctx->stream;
    std::vector<int> rp, cols;
    if (S.own) {
        rp = S.own_rp;
        cols = S.own_cols;
    } else {
        rp.resize((size_t)ctx->n * ctx->csr_ndof + 1);
        cols.resize((size_t)ctx->nnz);
        PXG_HIP(hipMemcpyAsync(rp.data(), ctx->rowptr.p, sizeof(int) * rp.size(), hipMemcpyDeviceToHost, st));
        PXG_HIP(hipMemcpyAsync(cols.data(), ctx->cols.p, sizeof(int) * cols.size(), hipMemcpyDeviceToHost, st));
        PXG_HIP(hipStreamSynchronize(st));
    }
    const int n = (int)rp.size() - 1;
    const long long nnz = (long long)cols.size();
    std::vector<int> perm, inv(n);
    rcm_order(n, rp, cols, S.use_rcm != 0, perm);
    for (int k = 0; k < n; k++) inv[perm[k]] = k;
    // DefaultSolver.cpp:177-194: first(j) = least k < j with (j,k) or (k,j) stored
    std::vector<int> first(n);
    for (int j = 0; j < n; j++) first[j] = j;
    for (int io = 0; io < n; io++) {
        const int i = inv[io];
        for (int p = rp[io]; p < rp[io + 1]; p++) {
            const int j = inv[cols[p]];
            const int lo = std::min(i, j), hi = std::max(i, j);
            if (lo < first[hi]) first[hi] = lo;
        }
    }
    std::vector<long long> jp((size_t)n + 1, 0);
    for (int j = 0; j < n; j++) jp[j + 1] = jp[j] + (j - first[j]);
    // the reference's storage guard (DefaultSolver.cpp:196-238): the skyline
    // (two doubles per profile entry plus the diagonal) must fit max_profile_gb
    const double gib = ((double)jp[n] * 2.0 + (double)n) * 8.0 / (1024.0 * 1024.0 * 1024.0);
    PXG_CHECK(!(S.max_gib > 0.0 && gib > S.max_gib), PXG_ERR_ARG,
              "pxg_solve: the skyline profile needs " + std::to_string(gib) + " GiB (n=" + std::to_string(n) +
                  ", profile=" + std::to_string(jp[n]) + "), exceeding max_profile_gb=" +
                  std::to_string(S.max_gib) + " (DefaultSolver.cpp:196-238); raise it or set it <= 0");
    const int nb = (n + SB - 1) / SB;
    std::vector<int> fb(nb);
    for (int I = 0; I < nb; I++) {
        int f = I;
        for (int j = I * SB; j < std::min(n, (I + 1) * SB); j++) f = std::min(f, first[j] / SB);
        fb[I] = f;
    }
    std::vector<long long> lp((size_t)nb + 1, 0);
    for (int I = 0; I < nb; I++) lp[I + 1] = lp[I] + (I - fb[I]);
    const long long ns = lp[nb];
    // active set of step K: block rows I > K whose envelope reaches column K
    std::vector<int> ap((size_t)nb + 1, 0), al((size_t)std::max<long long>(ns, 1));
    for (int I = 0; I < nb; I++)
        for (int K = fb[I]; K < I; K++) ap[K + 1]++;
    for (int K = 0; K < nb; K++) ap[K + 1] += ap[K];
    {
        std::vector<int> fill(ap.begin(), ap.end() - 1);
        for (int I = 0; I < nb; I++)
            for (int K = fb[I]; K < I; K++) al[(size_t)fill[K]++] = I;
    }
    int amax = 0;
    for (int K = 0; K < nb; K++) amax = std::max(amax, ap[K + 1] - ap[K]);
    // value slot of every CSR entry in the block storage
    std::vector<long long> loc((size_t)nnz);
    for (int io = 0; io < n; io++) {
        const int i = inv[io], I = i / SB;
        for (int p = rp[io]; p < rp[io + 1]; p++) {
            const int j = inv[cols[p]], J = j / SB;
            long long blk;
            if (I == J) blk = 2 * ns + I;
            else if (I > J) blk = lp[I] + (J - fb[I]);
            else blk = ns + lp[J] + (I - fb[J]);
            loc[(size_t)p] = blk * SB2 + (long long)(i % SB) * SB + (j % SB);
        }
    }
    const size_t fblocks = (size_t)(2 * ns + nb);
    PXG_HIP(S.F.alloc(fblocks * SB2));
    PXG_HIP(S.Li.alloc((size_t)nb * SB2));
    PXG_HIP(S.Ui.alloc((size_t)nb * SB2));
    PXG_HIP(S.y.alloc((size_t)nb * SB));
    PXG_HIP(S.z.alloc((size_t)nb * SB));
    PXG_HIP(S.w.alloc((size_t)nb * SB));
    PXG_HIP(S.inv.alloc((size_t)n));
    PXG_HIP(S.loc.alloc((size_t)nnz));
    PXG_HIP(S.fb.alloc((size_t)nb));
    PXG_HIP(S.lp.alloc((size_t)nb + 1));
    PXG_HIP(S.al.alloc(al.size()));
    PXG_HIP(S.zpiv.alloc(1));
    PXG_HIP(S.part.alloc(2 * kNormBlocks));
    PXG_HIP(S.norms.alloc(2));
    PXG_HIP(hipMemcpyAsync(S.inv.p, inv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(S.loc.p, loc.data(), sizeof(long long) * loc.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(S.fb.p, fb.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(S.lp.p, lp.data(), sizeof(long long) * lp.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(S.al.p, al.data(), sizeof(int) * al.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipStreamSynchronize(st));
    S.hperm = perm;
    S.hjp = jp;
    S.hap = ap;
    S.hal = al;
    S.hfb = fb;
    S.n = n;
    S.nb = nb;
    S.ns = ns;
    S.amax = amax;
    S.nnz = nnz;
    S.gen = pattern_key(ctx);
    S.factored = 0;
    S.ready = 1;
    return PXG_OK;
}

int ensure_symbolic(pxg_ctx *ctx)
{
    PXG_CHECK(ctx && (ctx->solve.own || (ctx->nnz > 0 && ctx->csr_ndof > 0)), PXG_ERR_STATE,
              "pxg_solve: call pxg_build_csr, pxg_set_csr or pxg_solve_set_pattern first");
    const SolveState &S = ctx->solve;
    if (S.ready && S.gen == pattern_key(ctx)) return PXG_OK;
    return symbolic(ctx);
}

Layout layout(SolveState &S)
{
    return Layout{S.F.p, S.ns, S.fb.p, S.lp.p};
}

// the per-step launch chains depend only on the pattern and the solver's own
// buffers: enqueued directly, or captured once into a HIP graph and replayed
// the next diagonal factor joins the trailing update when the update has at
// least this many active blocks (below it the update is shorter than the
// fused kernel's extra block product)
constexpr int kFuseMinActive = 16;

int enqueue_factor_steps(SolveState &S, hipStream_t st)
{
    const Layout Y = layout(S);
    double *D = S.F.p + (size_t)(2 * S.ns) * SB2;
    if (S.nb > 0) hipLaunchKernelGGL(k_diag, dim3(1), dim3(256), 0, st, 0, D, S.Li.p, S.Ui.p, S.zpiv.p);
    for (int K = 0; K < S.nb; K++) {
        // D_K is factored (by k_diag or by the previous step's fused update)
        const int na = S.hap[K + 1] - S.hap[K];
        const bool next = K + 1 < S.nb;
        if (na == 0) {
            if (next) hipLaunchKernelGGL(k_diag, dim3(1), dim3(256), 0, st, K + 1, D, S.Li.p, S.Ui.p, S.zpiv.p);
            continue;
        }
        const int *act = S.al.p + S.hap[K];
        hipLaunchKernelGGL(k_panel, dim3(2 * na * NSTRIP), dim3(256), 0, st, K, act, na, Y, S.Li.p, S.Ui.p);
        if (S.hal[(size_t)S.hap[K]] == K + 1 && na >= kFuseMinActive) {
            hipLaunchKernelGGL(k_update_diag, dim3(1 + (na * na - 1) * NSTRIP), dim3(256), 0, st, K, act, na, Y, D,
                               S.Li.p, S.Ui.p, S.zpiv.p);
        } else {
            hipLaunchKernelGGL(k_update, dim3(na * na * NSTRIP), dim3(256), 0, st, K, act, na, Y);
            if (next) hipLaunchKernelGGL(k_diag, dim3(1), dim3(256), 0, st, K + 1, D, S.Li.p, S.Ui.p, S.zpiv.p);
        }
    }
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

int enqueue_sweeps(SolveState &S, hipStream_t st)
{
    const Layout Y = layout(S);
    for (int K = 0; K < S.nb; K++) {
        const int na = S.hap[K + 1] - S.hap[K];
        hipLaunchKernelGGL(k_fwd, dim3(std::max(1, na)), dim3(256), 0, st, K, S.al.p + S.hap[K], na, Y, S.Li.p,
                           S.y.p, S.z.p);
    }
    for (int K = S.nb - 1; K >= 0; K--) {
        const int nu = K - S.hfb[K];
        hipLaunchKernelGGL(k_bwd, dim3(std::max(1, nu)), dim3(256), 0, st, K, Y, S.Ui.p, S.z.p, S.w.p);
    }
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

int run_chain(SolveState &S, hipStream_t st, hipGraphExec_t &exec, int (*enqueue)(SolveState &, hipStream_t))
{
    if (!S.use_graphs) return enqueue(S, st);
    if (!exec) {
        hipGraph_t g = nullptr;
        PXG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(S, st);
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PXG_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) exec = nullptr;
        PXG_HIP(ei);
    }
    PXG_HIP(hipGraphLaunch(exec, st));
    return PXG_OK;
}

}  // namespace

// numeric factorisation of the device CSR values (DefaultSolver.cpp:248-370)
extern "C" int pxg_solve_factor_dev(pxg_ctx *ctx, const double *kvals_dev)
{
    PXG_CHECK(ctx && kvals_dev, PXG_ERR_ARG, "pxg_solve_factor_dev: bad arguments");
    PXG_FAULT(ctx, "pxg_solve_factor_dev");
    PXG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_symbolic(ctx);
    if (rc) return rc;
    SolveState &S = ctx->solve;
    S.factored = 0;
    hipStream_t st = ctx->stream;
    const size_t fcount = (size_t)(2 * S.ns + S.nb) * SB2;
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, st, S.F.p, fcount, 0.0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, S.zpiv.p, 1, INT_MAX);
    const long long npad = (long long)S.nb * SB;
    const long long work = std::max(S.nnz, npad - S.n);
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(work, 256)), dim3(256), 0, st, S.nnz, kvals_dev, S.loc.p, S.F.p,
                       S.n, (int)npad, S.F.p + (size_t)(2 * S.ns) * SB2);
    PXG_HIP(hipGetLastError());
    rc = run_chain(S, st, S.gfac, enqueue_factor_steps);
    if (rc) return rc;
    int zp = INT_MAX;
    PXG_HIP(hipMemcpyAsync(&zp, S.zpiv.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    PXG_CHECK(zp == INT_MAX || zp >= S.n, PXG_ERR_SINGULAR,
              "pxg_solve: zero diagonal pivot at equation " + std::to_string(zp + 1) +
                  " during LDU factorization (DefaultSolver.cpp:362-367)");
    S.factored = 1;
    return PXG_OK;
}

// substitution with the current factors (DefaultSolver.cpp:372-424)
extern "C" int pxg_solve_apply_dev(pxg_ctx *ctx, const double *rhs_dev, double *x_dev)
{
    PXG_CHECK(ctx && rhs_dev && x_dev, PXG_ERR_ARG, "pxg_solve_apply_dev: bad arguments");
    PXG_CHECK(ctx->solve.factored && ctx->solve.gen == pattern_key(ctx), PXG_ERR_STATE,
              "pxg_solve_apply_dev: call pxg_solve_factor_dev first");
    PXG_FAULT(ctx, "pxg_solve_apply_dev");
    PXG_HIP(hipSetDevice(ctx->device));
    SolveState &S = ctx->solve;
    hipStream_t st = ctx->stream;
    const int npad = S.nb * SB;
    hipLaunchKernelGGL(k_perm_in, dim3(grid_for(npad, 256)), dim3(256), 0, st, S.n, npad, S.inv.p, rhs_dev, S.y.p);
    int rc = run_chain(S, st, S.gsub, enqueue_sweeps);
    if (rc) return rc;
    hipLaunchKernelGGL(k_perm_out, dim3(grid_for(S.n, 256)), dim3(256), 0, st, S.n, S.inv.p, S.w.p, x_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_solve_dev(pxg_ctx *ctx, const double *kvals_dev, const double *rhs_dev, double *x_dev)
{
    int rc = pxg_solve_factor_dev(ctx, kvals_dev);
    if (rc) return rc;
    return pxg_solve_apply_dev(ctx, rhs_dev, x_dev);
}

extern "C" int pxg_solve(pxg_ctx *ctx, const double *kvals, const double *rhs, double *x)
{
    PXG_CHECK(ctx && kvals && rhs && x, PXG_ERR_ARG, "pxg_solve: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_symbolic(ctx);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const size_t rows = (size_t)ctx->solve.n, nnz = (size_t)ctx->solve.nnz;
    PXG_HIP(ctx->sk.alloc(nnz));
    PXG_HIP(ctx->sr.alloc(rows));
    PXG_HIP(ctx->su.alloc(rows));
    PXG_HIP(hipMemcpyAsync(ctx->sk.p, kvals, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->sr.p, rhs, sizeof(double) * rows, hipMemcpyHostToDevice, st));
    rc = pxg_solve_dev(ctx, ctx->sk.p, ctx->sr.p, ctx->su.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(x, ctx->su.p, sizeof(double) * rows, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_solve_set_pattern(pxg_ctx *ctx, int n, const int *rowptr, const int *cols)
{
    PXG_CHECK(ctx && n > 0 && rowptr && cols, PXG_ERR_ARG, "pxg_solve_set_pattern: bad arguments");
    PXG_CHECK(rowptr[0] == 0 && rowptr[n] > 0, PXG_ERR_ARG, "pxg_solve_set_pattern: bad row offsets");
    for (int i = 0; i < n; i++) {
        PXG_CHECK(rowptr[i + 1] >= rowptr[i], PXG_ERR_ARG, "pxg_solve_set_pattern: row offsets must not decrease");
        bool diag = false;
        for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
            PXG_CHECK(cols[p] >= 0 && cols[p] < n, PXG_ERR_ARG, "pxg_solve_set_pattern: column out of range");
            PXG_CHECK(p == rowptr[i] || cols[p] > cols[p - 1], PXG_ERR_ARG,
                      "pxg_solve_set_pattern: columns must be sorted and distinct within a row");
            diag = diag || cols[p] == i;
        }
        PXG_CHECK(diag, PXG_ERR_ARG, "pxg_solve_set_pattern: a row without its diagonal");
    }
    SolveState &S = ctx->solve;
    S.own_rp.assign(rowptr, rowptr + n + 1);
    S.own_cols.assign(cols, cols + rowptr[n]);
    S.own = 1;
    S.own_gen++;
    S.ready = 0;
    S.factored = 0;
    return PXG_OK;
}

extern "C" int pxg_solve_options(pxg_ctx *ctx, int use_rcm, double max_profile_gib, int flags)
{
    PXG_CHECK(ctx && (flags & ~PXG_SOLVE_NO_GRAPHS) == 0, PXG_ERR_ARG, "pxg_solve_options: bad arguments");
    SolveState &S = ctx->solve;
    S.use_rcm = use_rcm != 0;
    S.max_gib = max_profile_gib;
    S.use_graphs = !(flags & PXG_SOLVE_NO_GRAPHS);
    S.ready = 0;                      // the next solve re-runs the symbolic phase
    S.factored = 0;
    return PXG_OK;
}

extern "C" int pxg_solve_info(pxg_ctx *ctx, long long *out)
{
    PXG_CHECK(ctx && out, PXG_ERR_ARG, "pxg_solve_info: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_symbolic(ctx);
    if (rc) return rc;
    const SolveState &S = ctx->solve;
    out[0] = S.n;
    out[1] = S.nb;
    out[2] = SB;
    out[3] = S.hjp[(size_t)S.n];
    out[4] = S.ns;
    out[5] = S.amax;
    out[6] = (long long)((2 * S.ns + 3 * (long long)S.nb) * SB2 * (long long)sizeof(double));
    long long tiles = 0;
    for (int K = 0; K < S.nb; K++) {
        const long long na = S.hap[K + 1] - S.hap[K];
        tiles += na * na;
    }
    out[7] = tiles;
    return PXG_OK;
}

extern "C" int pxg_solve_ordering(pxg_ctx *ctx, int *perm, long long *jp)
{
    PXG_CHECK(ctx, PXG_ERR_ARG, "pxg_solve_ordering: bad arguments");
    PXG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_symbolic(ctx);
    if (rc) return rc;
    const SolveState &S = ctx->solve;
    if (perm) std::copy(S.hperm.begin(), S.hperm.end(), perm);
    if (jp) std::copy(S.hjp.begin(), S.hjp.end(), jp);
    return PXG_OK;
}

extern "C" int pxg_csr_preset_dev(pxg_ctx *ctx, double t, double *u_dev)
{
    PXG_CHECK(ctx && ctx->csrbc.set, PXG_ERR_STATE, "pxg_csr_preset_dev: call pxg_csr_set_bcs first");
    PXG_CHECK(u_dev, PXG_ERR_ARG, "pxg_csr_preset_dev: bad arguments");
    PXG_CHECK(!ctx->csrbc.chained, PXG_ERR_STATE,
              "pxg_csr_preset_dev: a preset reads a constrained dof; the list needs its sequential order");
    PXG_FAULT(ctx, "pxg_csr_preset_dev");
    PXG_HIP(hipSetDevice(ctx->device));
    const CsrBcProgram &P = ctx->csrbc;
    if (P.count == 0) return PXG_OK;
    hipLaunchKernelGGL(k_csr_preset, dim3(grid_for(P.count, 256)), dim3(256), 0, ctx->stream, P.count, P.row.p,
                       P.pdof.p, P.kind.p, P.val0.p, P.vel.p, t, (const double *)nullptr, u_dev);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

extern "C" int pxg_newton_update_dev(pxg_ctx *ctx, double t, double *u_dev, const double *du_dev,
                                     const double *rhs_dev, double *norms)
{
    PXG_CHECK(ctx && ctx->csrbc.set, PXG_ERR_STATE, "pxg_newton_update_dev: call pxg_csr_set_bcs first");
    PXG_CHECK(u_dev && du_dev && rhs_dev && norms, PXG_ERR_ARG, "pxg_newton_update_dev: bad arguments");
    PXG_CHECK(!ctx->csrbc.chained, PXG_ERR_STATE,
              "pxg_newton_update_dev: a preset reads a constrained dof; the list needs its sequential order");
    PXG_FAULT(ctx, "pxg_newton_update_dev");
    PXG_HIP(hipSetDevice(ctx->device));
    SolveState &S = ctx->solve;
    PXG_HIP(S.part.alloc(2 * kNormBlocks));
    PXG_HIP(S.norms.alloc(2));
    hipStream_t st = ctx->stream;
    const long long rows = (long long)ctx->n * ctx->csr_ndof;
    hipLaunchKernelGGL(k_sumsq2, dim3(kNormBlocks), dim3(256), 0, st, rows, rhs_dev, du_dev, S.part.p);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(64), 0, st, kNormBlocks, S.part.p, S.norms.p);
    hipLaunchKernelGGL(k_newton_add, dim3(grid_for(rows, 256)), dim3(256), 0, st, rows, du_dev, S.norms.p, u_dev);
    const CsrBcProgram &P = ctx->csrbc;
    if (P.count > 0)
        hipLaunchKernelGGL(k_csr_preset, dim3(grid_for(P.count, 256)), dim3(256), 0, st, P.count, P.row.p, P.pdof.p,
                           P.kind.p, P.val0.p, P.vel.p, t, (const double *)S.norms.p, u_dev);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipMemcpyAsync(norms, S.norms.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

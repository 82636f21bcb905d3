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
// plain FP64 GEMM tiles, one workgroup per block); the chain of n pivots is the
// serial part, as it is in the reference.  Substitution walks the blocks with
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
constexpr int LS = SB + 1;            // LDS row stride of the diagonal factor (bank spread)
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

// (1) D_K = L U in place (unit L below the diagonal, U on and above), then
// L^-1 (unit lower) and U^-1 (upper).  A zero pivot records the least failing
// equation (new order) in *zpiv and continues with 1 so that the arithmetic
// stays finite; the host reports the failure after the factorisation.
__global__ __launch_bounds__(256) void k_diag(int K, double *__restrict__ Dall, double *__restrict__ Liall,
                                              double *__restrict__ Uiall, int *zpiv)
{
    __shared__ double a[SB * LS];
    __shared__ double x[SB * LS];     // L^-1
    __shared__ double z[SB * LS];     // U'^-1 of the unit upper U' = diag(U)^-1 U
    const int t = (int)threadIdx.x;
    double *D = Dall + (size_t)K * SB2;
    for (int e = t; e < SB2; e += kThreads) {
        const int r = e / SB, c = e % SB;
        a[r * LS + c] = D[e];
        x[r * LS + c] = (r == c) ? 1.0 : 0.0;
        z[r * LS + c] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < SB; k++) {
        double piv = a[k * LS + k];
        if (piv == 0.0) {
            if (t == 0) atomicMin(zpiv, K * SB + k);
            piv = 1.0;
        }
        if (t > k && t < SB) a[t * LS + k] /= piv;
        __syncthreads();
        // rank-1 update of the trailing (SB-1-k)^2 entries
        const int m = SB - 1 - k;
        for (int e = t; e < m * m; e += kThreads) {
            const int r = k + 1 + e / m, c = k + 1 + e % m;
            a[r * LS + c] -= a[r * LS + k] * a[k * LS + c];
        }
        __syncthreads();
    }
    // a zero pivot continues as 1 (see above): the inverses use the same value
    // L^-1: row operations X[i][:] -= L[i][k] X[k][:] for i > k (X[k] is
    // final after step k-1); U'^-1: Z[i][:] -= (U[i][k]/U[i][i]) Z[k][:] for
    // i < k, k from the bottom.  The two walks share the barriers.
    for (int s = 0; s < SB; s++) {
        const int kl = s, ku = SB - 1 - s;
        for (int e = t; e < SB * SB; e += kThreads) {
            const int i = e / SB, c = e % SB;
            if (i > kl && c <= kl) x[i * LS + c] -= a[i * LS + kl] * x[kl * LS + c];
            if (i < ku && c >= ku) {
                double d = a[i * LS + i];
                if (d == 0.0) d = 1.0;
                z[i * LS + c] -= (a[i * LS + ku] / d) * z[ku * LS + c];
            }
        }
        __syncthreads();
    }
    double *Li = Liall + (size_t)K * SB2, *Ui = Uiall + (size_t)K * SB2;
    for (int e = t; e < SB2; e += kThreads) {
        const int r = e / SB, c = e % SB;
        double d = a[c * LS + c];
        if (d == 0.0) d = 1.0;
        D[e] = a[r * LS + c];
        Li[e] = x[r * LS + c];
        Ui[e] = z[r * LS + c] / d;             // U^-1 = U'^-1 diag(U)^-1
    }
}

// C = A B (SUB false) or C -= A B (SUB true) for SB x SB row-major blocks;
// C may alias A or B (both are read into LDS before the barrier, C written
// after it).  256 threads, each a 4 x 4 tile of C.
template <bool SUB>
__device__ __forceinline__ void block_gemm(const double *A, const double *B, double *C, double *At, double *Bs)
{
    const int t = (int)threadIdx.x;
    for (int e = t; e < SB2 / 2; e += kThreads) {
        const int r = (2 * e) / SB, c = (2 * e) % SB;
        const double2 va = *reinterpret_cast<const double2 *>(A + 2 * e);
        const double2 vb = *reinterpret_cast<const double2 *>(B + 2 * e);
        At[c * TS + r] = va.x;
        At[(c + 1) * TS + r] = va.y;
        *reinterpret_cast<double2 *>(Bs + r * TS + c) = vb;
    }
    __syncthreads();
    const int r0 = (t / 16) * 4, c0 = (t % 16) * 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
#pragma unroll 8
    for (int k = 0; k < SB; k++) {
        const double2 a01 = *reinterpret_cast<const double2 *>(At + k * TS + r0);
        const double2 a23 = *reinterpret_cast<const double2 *>(At + k * TS + r0 + 2);
        const double2 b01 = *reinterpret_cast<const double2 *>(Bs + k * TS + c0);
        const double2 b23 = *reinterpret_cast<const double2 *>(Bs + k * TS + c0 + 2);
        const double av[4] = {a01.x, a01.y, a23.x, a23.y};
        const double bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double *row = C + (r0 + i) * SB + c0;
        if (SUB) {
            const double2 o01 = *reinterpret_cast<const double2 *>(row);
            const double2 o23 = *reinterpret_cast<const double2 *>(row + 2);
            *reinterpret_cast<double2 *>(row) = make_double2(o01.x - acc[i][0], o01.y - acc[i][1]);
            *reinterpret_cast<double2 *>(row + 2) = make_double2(o23.x - acc[i][2], o23.y - acc[i][3]);
        } else {
            *reinterpret_cast<double2 *>(row) = make_double2(acc[i][0], acc[i][1]);
            *reinterpret_cast<double2 *>(row + 2) = make_double2(acc[i][2], acc[i][3]);
        }
    }
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

// (2) workgroup w < na: L_IK = A_IK U_KK^-1; else U_KJ = L_KK^-1 A_KJ
__global__ __launch_bounds__(256) void k_panel(int K, const int *__restrict__ act, int na, Layout Y,
                                               const double *__restrict__ Li, const double *__restrict__ Ui)
{
    __shared__ __align__(16) double At[SB * TS];
    __shared__ __align__(16) double Bs[SB * TS];
    const int w = (int)blockIdx.x;
    if (w < na) {
        double *c = Y.L(act[w], K);
        block_gemm<false>(c, Ui + (size_t)K * SB2, c, At, Bs);
    } else {
        double *c = Y.U(K, act[w - na]);
        block_gemm<false>(Li + (size_t)K * SB2, c, c, At, Bs);
    }
}

// (3) A_IJ -= L_IK U_KJ for (I, J) in active x active
__global__ __launch_bounds__(256) void k_update(int K, const int *__restrict__ act, int na, Layout Y)
{
    __shared__ __align__(16) double At[SB * TS];
    __shared__ __align__(16) double Bs[SB * TS];
    const int w = (int)blockIdx.x;
    const int I = act[w / na], J = act[w % na];
    double *c = (I > J) ? Y.L(I, J) : (I < J) ? Y.U(I, J) : Y.D(I);
    block_gemm<true>(Y.L(I, K), Y.U(K, J), c, At, Bs);
}

// v = M y_K for one SB x SB block M (row-major) into LDS v[SB]; 4 threads per row
__device__ __forceinline__ void block_gemv(const double *__restrict__ M, const double *__restrict__ yk, double *ys,
                                           double *v)
{
    const int t = (int)threadIdx.x;
    if (t < SB) ys[t] = yk[t];
    __syncthreads();
    const int r = t / 4, q = t % 4;
    double s = 0.0;
    for (int c = q; c < SB; c += 4) s = fma(M[r * SB + c], ys[c], s);
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (q == 0) v[r] = s;
    __syncthreads();
}

// forward step K: z_K = L_KK^-1 y_K; y_I -= L_IK z_K for the active I
__global__ __launch_bounds__(256) void k_fwd(int K, const int *__restrict__ act, int na, Layout Y,
                                             const double *__restrict__ Li, double *y, double *zout)
{
    __shared__ double ys[SB], v[SB], u[SB];
    block_gemv(Li + (size_t)K * SB2, y + (size_t)K * SB, ys, v);
    const int w = (int)blockIdx.x, t = (int)threadIdx.x;
    if (w == 0 && t < SB) zout[(size_t)K * SB + t] = v[t];
    if (w >= na) return;
    const int I = act[w];
    const double *L = Y.L(I, K);
    const int r = t / 4, q = t % 4;
    double s = 0.0;
    for (int c = q; c < SB; c += 4) s = fma(L[r * SB + c], v[c], s);
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (q == 0) u[r] = s;
    __syncthreads();
    if (t < SB) y[(size_t)I * SB + t] -= u[t];
}

// backward step K: x_K = U_KK^-1 z_K; z_I -= U_IK x_K for I in [FB[K], K)
__global__ __launch_bounds__(256) void k_bwd(int K, Layout Y, const double *__restrict__ Ui, double *z,
                                             double *xout)
{
    __shared__ double ys[SB], v[SB], u[SB];
    block_gemv(Ui + (size_t)K * SB2, z + (size_t)K * SB, ys, v);
    const int w = (int)blockIdx.x, t = (int)threadIdx.x;
    if (w == 0 && t < SB) xout[(size_t)K * SB + t] = v[t];
    const int I = Y.fb[K] + w;
    if (I >= K) return;
    const double *U = Y.U(I, K);
    const int r = t / 4, q = t % 4;
    double s = 0.0;
    for (int c = q; c < SB; c += 4) s = fma(U[r * SB + c], v[c], s);
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (q == 0) u[r] = s;
    __syncthreads();
    if (t < SB) z[(size_t)I * SB + t] -= u[t];
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

int symbolic(pxg_ctx *ctx)
{
    SolveState &S = ctx->solve;
    S.ready = 0;
    hipStream_t st = ctx->stream;
    const int n = ctx->n * ctx->csr_ndof;
    const long long nnz = ctx->nnz;
    std::vector<int> rp((size_t)n + 1), cols((size_t)nnz);
    PXG_HIP(hipMemcpyAsync(rp.data(), ctx->rowptr.p, sizeof(int) * rp.size(), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipMemcpyAsync(cols.data(), ctx->cols.p, sizeof(int) * cols.size(), hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
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
    S.hfb = fb;
    S.n = n;
    S.nb = nb;
    S.ns = ns;
    S.amax = amax;
    S.nnz = nnz;
    S.gen = ctx->csr_gen;
    S.factored = 0;
    S.ready = 1;
    return PXG_OK;
}

int ensure_symbolic(pxg_ctx *ctx)
{
    PXG_CHECK(ctx && ctx->nnz > 0 && ctx->csr_ndof > 0, PXG_ERR_STATE, "pxg_solve: call pxg_build_csr first");
    const SolveState &S = ctx->solve;
    if (S.ready && S.gen == ctx->csr_gen && S.nnz == ctx->nnz && S.n == ctx->n * ctx->csr_ndof) return PXG_OK;
    return symbolic(ctx);
}

Layout layout(SolveState &S)
{
    return Layout{S.F.p, S.ns, S.fb.p, S.lp.p};
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
    Layout Y = layout(S);
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(work, 256)), dim3(256), 0, st, S.nnz, kvals_dev, S.loc.p, S.F.p,
                       S.n, (int)npad, S.F.p + (size_t)(2 * S.ns) * SB2);
    PXG_HIP(hipGetLastError());
    double *D = S.F.p + (size_t)(2 * S.ns) * SB2;
    for (int K = 0; K < S.nb; K++) {
        hipLaunchKernelGGL(k_diag, dim3(1), dim3(256), 0, st, K, D, S.Li.p, S.Ui.p, S.zpiv.p);
        const int na = S.hap[K + 1] - S.hap[K];
        if (na > 0) {
            const int *act = S.al.p + S.hap[K];
            hipLaunchKernelGGL(k_panel, dim3(2 * na), dim3(256), 0, st, K, act, na, Y, S.Li.p, S.Ui.p);
            hipLaunchKernelGGL(k_update, dim3(na * na), dim3(256), 0, st, K, act, na, Y);
        }
    }
    PXG_HIP(hipGetLastError());
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
    PXG_CHECK(ctx->solve.factored && ctx->solve.gen == ctx->csr_gen, PXG_ERR_STATE,
              "pxg_solve_apply_dev: call pxg_solve_factor_dev first");
    PXG_FAULT(ctx, "pxg_solve_apply_dev");
    PXG_HIP(hipSetDevice(ctx->device));
    SolveState &S = ctx->solve;
    hipStream_t st = ctx->stream;
    const int npad = S.nb * SB;
    Layout Y = layout(S);
    hipLaunchKernelGGL(k_perm_in, dim3(grid_for(npad, 256)), dim3(256), 0, st, S.n, npad, S.inv.p, rhs_dev, S.y.p);
    for (int K = 0; K < S.nb; K++) {
        const int na = S.hap[K + 1] - S.hap[K];
        hipLaunchKernelGGL(k_fwd, dim3(std::max(1, na)), dim3(256), 0, st, K, S.al.p + S.hap[K], na, Y, S.Li.p,
                           S.y.p, S.z.p);
    }
    for (int K = S.nb - 1; K >= 0; K--) {
        const int nu = K - S.hfb[K];
        hipLaunchKernelGGL(k_bwd, dim3(std::max(1, nu)), dim3(256), 0, st, K, Y, S.Ui.p, S.z.p, S.w.p);
    }
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
    PXG_CHECK(ctx->nnz > 0 && ctx->csr_ndof > 0, PXG_ERR_STATE, "pxg_solve: call pxg_build_csr first");
    PXG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t rows = (size_t)ctx->n * ctx->csr_ndof, nnz = (size_t)ctx->nnz;
    PXG_HIP(ctx->sk.alloc(nnz));
    PXG_HIP(ctx->sr.alloc(rows));
    PXG_HIP(ctx->su.alloc(rows));
    PXG_HIP(hipMemcpyAsync(ctx->sk.p, kvals, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(ctx->sr.p, rhs, sizeof(double) * rows, hipMemcpyHostToDevice, st));
    int rc = pxg_solve_dev(ctx, ctx->sk.p, ctx->sr.p, ctx->su.p);
    if (rc) return rc;
    PXG_HIP(hipMemcpyAsync(x, ctx->su.p, sizeof(double) * rows, hipMemcpyDeviceToHost, st));
    PXG_HIP(hipStreamSynchronize(st));
    return PXG_OK;
}

extern "C" int pxg_solve_options(pxg_ctx *ctx, int use_rcm, double max_profile_gib)
{
    PXG_CHECK(ctx, PXG_ERR_ARG, "pxg_solve_options: bad arguments");
    SolveState &S = ctx->solve;
    S.use_rcm = use_rcm != 0;
    S.max_gib = max_profile_gib;
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

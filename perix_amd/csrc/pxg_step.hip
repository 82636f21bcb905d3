// pxg_step.hip -- the matrix-free explicit PDDO fracture step (one launch).
//
// Reference: ExplicitPDDOFracElement::preprocessIteration, the per-step
// damage + Navier force loop (ExplicitPDDOFracElement.cpp:237-299), and the
// central-difference update of TimeStepping::solveExplicitDynamics
// (TimeStepping.cpp:417-599).
//
// Work layout.  A wavefront owns 64/S consecutive nodes and gives each node S
// adjacent lanes.  A node's bonds [off[i], off[i+1]) are contiguous; the lanes
// walk them in groups of G consecutive bonds aligned to G, so one G-wide
// vector load fetches a group's neighbour ids, its health bytes and each
// operator component (all bond arrays are allocated with zeroed slack and the
// operator components are strided by a multiple of 64, so every such load is
// aligned and in bounds).  A group that straddles two families is loaded by
// both nodes and masked by each.  Wavefronts never wait for each other: the
// kernel has no block-level barrier.
//
// Damage.  The reference breaks bond (i,j) when s = (|y| - |xi|)/|xi| > sc
// (tension-only) or |s| > sc.  The kernel first decides from
// q = |y|^2 - |xi|^2 (1+sc)^2 (and the (1-sc)^2 side), both in FP64 with no
// square root or division; only when |q| <= 1e-9 |xi|^2, far inside any
// rounding difference between the two forms, does it evaluate the
// reference's own expression.  The broken-bond decisions are therefore those
// of the reference formula, bit for bit.
//
// Family sum.  FP64: every lane parks its bond forces in LDS (broken bonds as
// -0.0, the additive identity), then one lane per component adds the node's
// bonds in family order, which is the reference's exact summation order.
// FP32C: forces are formed in FP32 from FP32 operators and FP32 displacement
// differences and accumulated per lane with TwoSum (an FP32 sum/compensation
// pair); the S lanes of a node combine their pairs with a lane butterfly
// (DPP within 16 lanes), and the FP64 update consumes sum + compensation.
#include "pxg_internal.h"
#include <cfloat>
#include <climits>
#include <cmath>

namespace pxg {

struct StepArgs {
    int n;
    int nblk;                  // logical blocks (grid is rounded up to 8*)
    long long ps;              // component stride of the operator arrays
    const int *off;
    const int *nbr;
    const double *xp;
    const double *P64;
    const float *P32;
    unsigned char *health;
    int *break_step;
    int *anybroke;
    int step;
    const double *u;           // u^n
    const double *uold;        // u^{n-1} (MODE_CD/VAR) or v0 (MODE_INIT)
    double *out;               // u^{n+1} / accel / u^{-1}
    double sc, chi, clo;       // critical stretch, (1+sc)^2, (1-sc)^2 or -1
    int tension_only, damage;
    double b0, b1, b2, rho;
    double c1, c2;             // MODE_CD: dt2; MODE_VAR: rho_, w; MODE_INIT: dt, half
    const unsigned char *mask;
    unsigned long long *gmax;
    const int *gstate;
    int frozen;                // >0: health frozen from break_step (no damage)
    int frozen_k;              // bond t is intact iff break_step[t] > frozen_k
};

template <int DIM, bool F32> struct StepCfg;
// S lanes per node (S divides 64), G bonds per lane group.
template <> struct StepCfg<2, true> { static constexpr int S = 8, G = 4; };
template <> struct StepCfg<2, false> { static constexpr int S = 8, G = 4; };
template <> struct StepCfg<3, true> { static constexpr int S = 16, G = 4; };
template <> struct StepCfg<3, false> { static constexpr int S = 16, G = 4; };

template <int B> struct VecT;
template <> struct VecT<2> { using T = unsigned short; };
template <> struct VecT<4> { using T = unsigned; };
template <> struct VecT<8> { using T = uint2; };
template <> struct VecT<16> { using T = uint4; };

// G consecutive elements starting at an address aligned to their total size
template <typename T, int G>
__device__ __forceinline__ void vload(const T *p, T (&o)[G])
{
    constexpr int B = (int)sizeof(T) * G;
    if constexpr (B <= 16) {
        const auto v = *reinterpret_cast<const typename VecT<B>::T *>(p);
        __builtin_memcpy(o, &v, B);
    } else {
        static_assert(B % 16 == 0, "vector width");
#pragma unroll
        for (int h = 0; h < B / 16; h++) {
            const uint4 v = reinterpret_cast<const uint4 *>(p)[h];
            __builtin_memcpy(reinterpret_cast<char *>(o) + 16 * h, &v, 16);
        }
    }
}

// TwoSum (Knuth): a + b = s + e exactly, FP32
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// value held by the partner lane at butterfly level M.  Levels run M = 1, 2,
// 4, ...; before level M every lane of an aligned M-lane group holds the same
// value, so any pairing of the two halves of a 2M-lane group is a valid
// butterfly: quad permutations for M = 1, 2, the half-row and row mirrors
// for M = 4, 8, and a cross-row exchange beyond.
template <int M>
__device__ __forceinline__ float partner(float v)
{
    const int x = __builtin_bit_cast(int, v);
    if constexpr (M == 1) return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    else if constexpr (M == 2) return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    else if constexpr (M == 4) return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));
    else if constexpr (M == 8) return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false));
    else return __shfl_xor(v, M);
}

template <int M, int S, int DIM>
__device__ __forceinline__ void butterfly(float (&s)[DIM], float (&c)[DIM])
{
    if constexpr (M < S) {
#pragma unroll
        for (int k = 0; k < DIM; k++) {
            const float os = partner<M>(s[k]);
            const float oc = partner<M>(c[k]);
            float sn, e;
            two_sum(s[k], os, sn, e);
            s[k] = sn;
            c[k] = c[k] + oc + e;
        }
        butterfly<2 * M, S, DIM>(s, c);
    }
}

// LDS writes of this wavefront become visible to its other lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2; the
// remap gives every XCD one contiguous range of logical blocks, so that the
// neighbour rows a block gathers were mostly fetched into that XCD's L2 by
// the blocks just before it.  Returns -1 for the padding blocks of the grid.
__device__ __forceinline__ int xcd_block(int b, int nblk)
{
    const int per = (nblk + 7) >> 3;
    const int l = (b & 7) * per + (b >> 3);
    return l < nblk ? l : -1;
}

template <int DIM>
__device__ __forceinline__ double pick(const double (&v)[DIM], int c)
{
    return (DIM == 3 && c == 2) ? v[DIM - 1] : (c == 1 ? v[1] : v[0]);
}

template <int DIM, bool F32, int MODE>
__global__ void __launch_bounds__(256) k_frac_step(StepArgs a)
{
    constexpr int S = StepCfg<DIM, F32>::S;
    constexpr int G = StepCfg<DIM, F32>::G;
    constexpr int NPW = 64 / S;               // nodes per wavefront
    constexpr int W = G * S;                  // bonds per node and round
    constexpr int NC = (DIM == 3) ? 6 : 3;
    constexpr int ROW = W * DIM + 1;          // LDS row per node (odd: no bank aliasing)
    __shared__ double s_f[F32 ? 1 : 4 * NPW * ROW];

    if (a.gstate && a.gstate[0]) return;      // blow-up guard tripped earlier
    const int blk = xcd_block(blockIdx.x, a.nblk);
    if (blk < 0) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ln = lane / S, sl = lane % S;
    const int i = (blk * 4 + wv) * NPW + ln;
    const bool nvalid = i < a.n;
    int b0 = 0, b1 = 0;
    if (nvalid) { b0 = a.off[i]; b1 = a.off[i + 1]; }
    const bool live_dmg = a.damage && !a.frozen;
    double ui[DIM], xi[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        ui[c] = nvalid ? a.u[(size_t)i * DIM + c] : 0.0;
        xi[c] = (nvalid && live_dmg) ? a.xp[(size_t)i * DIM + c] : 0.0;
    }
    const int A = b0 & ~(G - 1);
    int broke = 0;

    // one group of G bonds: loads, damage, forces (dead bonds give f = 0)
    // bond-array bases; offsets below are unsigned 32-bit element indices
    const float *P32c[F32 ? NC : 1];
    const double *P64c[F32 ? 1 : NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if constexpr (F32) P32c[c] = a.P32 + c * a.ps;
        else P64c[c] = a.P64 + c * a.ps;
    }
    using F64G = double[F32 ? 1 : G][DIM];
    using F32G = float[F32 ? G : 1][DIM];
    auto group = [&](unsigned t0, F64G &f64, F32G &f32) {
        int jv[G];
        vload(a.nbr + t0, jv);
        bool ok[G];
        if (a.frozen) {
            int bs[G];
            vload(a.break_step + t0, bs);
#pragma unroll
            for (int k = 0; k < G; k++) ok[k] = bs[k] > a.frozen_k;
        } else {
            unsigned char hv[G];
            vload(a.health + t0, hv);
#pragma unroll
            for (int k = 0; k < G; k++) ok[k] = hv[k] != 0;
        }
        float p32[F32 ? NC : 1][G];
        double p64[F32 ? 1 : NC][G];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if constexpr (F32) vload(P32c[c] + t0, p32[c]);
            else vload(P64c[c] + t0, p64[c]);
        }
#pragma unroll
        for (int k = 0; k < G; k++) {
            const int t = (int)t0 + k;
            bool live = ok[k] && t >= b0 && t < b1;
            double du[DIM];
#pragma unroll
            for (int c = 0; c < DIM; c++) du[c] = 0.0;
            if (live) {
                const unsigned jo = (unsigned)(jv[k] - 1) * (unsigned)(DIM * sizeof(double));
                const double *uj = reinterpret_cast<const double *>(reinterpret_cast<const char *>(a.u) + jo);
#pragma unroll
                for (int c = 0; c < DIM; c++) du[c] = uj[c] - ui[c];
                if (live_dmg) {
                    // stretch of ExplicitPDDOFracElement.cpp:249-266
                    const double *xj = reinterpret_cast<const double *>(reinterpret_cast<const char *>(a.xp) + jo);
                    double xs[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; c++) xs[c] = xj[c] - xi[c];
                    double rr = xs[0] * xs[0] + xs[1] * xs[1];
                    if (DIM == 3) rr = rr + xs[DIM - 1] * xs[DIM - 1];
                    if (rr > 0.0) {
                        const double yx = xs[0] + du[0], yy = xs[1] + du[1];
                        double yy2 = yx * yx + yy * yy;
                        if (DIM == 3) { const double yz = xs[DIM - 1] + du[DIM - 1]; yy2 = yy2 + yz * yz; }
                        const double band = 1.0e-9 * rr;
                        const double qh = yy2 - rr * a.chi;
                        bool brk = qh > 0.0;
                        bool unsure = fabs(qh) <= band;
                        if (!a.tension_only) {
                            const double ql = yy2 - rr * a.clo;
                            brk = brk || ql < 0.0;
                            unsure = unsure || fabs(ql) <= band;
                        }
                        if (unsure) {
                            const double rl = sqrt(rr);
                            const double s = (sqrt(yy2) - rl) / rl;
                            brk = a.tension_only ? (s > a.sc) : (fabs(s) > a.sc);
                        }
                        if (brk) {
                            a.health[t] = 0;
                            if (a.break_step) a.break_step[t] = a.step;
                            broke = 1;
                            live = false;
                        }
                    }
                }
            }
            if constexpr (F32) {
                float d[DIM];
#pragma unroll
                for (int c = 0; c < DIM; c++) d[c] = live ? (float)du[c] : 0.0f;
                if (DIM == 2) {
                    f32[k][0] = fmaf(p32[0][k], d[0], p32[2][k] * d[1]);
                    f32[k][1] = fmaf(p32[2][k], d[0], p32[1][k] * d[1]);
                } else {
                    f32[k][0] = fmaf(p32[0][k], d[0], fmaf(p32[3][k], d[1], p32[4][k] * d[DIM - 1]));
                    f32[k][1] = fmaf(p32[3][k], d[0], fmaf(p32[1][k], d[1], p32[5][k] * d[DIM - 1]));
                    f32[k][DIM - 1] = fmaf(p32[4][k], d[0], fmaf(p32[5][k], d[1], p32[2][k] * d[DIM - 1]));
                }
            } else {
                // health-gated Navier force (ExplicitPDDOFracElement.cpp:268-295)
                if (live) {
                    if (DIM == 2) {
                        f64[k][0] = p64[0][k] * du[0] + p64[2][k] * du[1];
                        f64[k][1] = p64[2][k] * du[0] + p64[1][k] * du[1];
                    } else {
                        f64[k][0] = p64[0][k] * du[0] + p64[3][k] * du[1] + p64[4][k] * du[DIM - 1];
                        f64[k][1] = p64[3][k] * du[0] + p64[1][k] * du[1] + p64[5][k] * du[DIM - 1];
                        f64[k][DIM - 1] = p64[4][k] * du[0] + p64[5][k] * du[1] + p64[2][k] * du[DIM - 1];
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < DIM; c++) f64[k][c] = -0.0;
                }
            }
        }
    };

    double Lown = 0.0;     // family sum of component sl (lanes sl < DIM)
    if constexpr (F32) {
        float Ls[DIM], Lc[DIM];
#pragma unroll
        for (int c = 0; c < DIM; c++) { Ls[c] = 0.0f; Lc[c] = 0.0f; }
        for (int t0 = A + G * sl; t0 < b1; t0 += W) {
            F64G f64;
            F32G f32;
            group((unsigned)t0, f64, f32);
#pragma unroll
            for (int k = 0; k < G; k++)
#pragma unroll
                for (int c = 0; c < DIM; c++) {
                    float sn, e;
                    two_sum(Ls[c], f32[k][c], sn, e);
                    Ls[c] = sn;
                    Lc[c] += e;
                }
        }
        butterfly<1, S, DIM>(Ls, Lc);
        double L[DIM];
#pragma unroll
        for (int c = 0; c < DIM; c++) L[c] = (double)Ls[c] + (double)Lc[c];
        Lown = pick(L, sl);
    } else {
        int nr = nvalid ? (b1 - A + W - 1) / W : 0;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) nr = max(nr, __shfl_xor(nr, m));
        double *row = s_f + (wv * NPW + ln) * ROW;
        for (int r = 0; r < nr; r++) {
            const int base = A + r * W;
            const int t0 = base + G * sl;
            F64G f64;
            if (t0 < b1) {
                F32G f32;
                group((unsigned)t0, f64, f32);
            } else {
#pragma unroll
                for (int k = 0; k < G; k++)
#pragma unroll
                    for (int c = 0; c < DIM; c++) f64[k][c] = -0.0;
            }
#pragma unroll
            for (int k = 0; k < G; k++)
#pragma unroll
                for (int c = 0; c < DIM; c++) row[(G * sl + k) * DIM + c] = f64[k][c];
            wave_lds_sync();
            if (nvalid && sl < DIM) {
                // the reference's family-order sum (ExplicitPDDOFracElement.cpp:268-295)
                const int lo = max(b0, base) - base, hi = min(b1, base + W) - base;
                for (int q = lo; q < hi; q++) Lown += row[q * DIM + sl];
            }
            wave_lds_sync();
        }
    }
    if (__any(broke) && lane == 0 && a.anybroke) atomicAdd(a.anybroke, 1);

    // central-difference update, one lane per component
    unsigned long long mymax = 0;
    if (nvalid && sl < DIM) {
        const int c = sl;
        const double bb = (c == 0) ? a.b0 : (c == 1 ? a.b1 : a.b2);
        const double acc = (Lown + bb) / a.rho;
        const double uu = pick(ui, c);
        const size_t idx = (size_t)i * DIM + c;
        double v;
        if (MODE == MODE_CD) v = 2.0 * uu - a.uold[idx] + a.c1 * acc;
        else if (MODE == MODE_VAR) v = uu + a.c1 * (uu - a.uold[idx]) + a.c2 * acc;
        else if (MODE == MODE_ACCEL) v = acc;
        else v = uu - a.c1 * a.uold[idx] + a.c2 * acc;
        a.out[idx] = v;
        if ((MODE == MODE_CD || MODE == MODE_VAR) && !a.mask[i])
            mymax = (unsigned long long)__double_as_longlong(fabs(v));
    }
    if (MODE == MODE_CD || MODE == MODE_VAR) {
        // blow-up guard maximum: |v| bit patterns order like the values
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const unsigned long long o = __shfl_xor(mymax, m);
            mymax = o > mymax ? o : mymax;
        }
        if (lane == 0 && mymax) {
            const unsigned long long cur = __hip_atomic_load(&a.gmax[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (mymax > cur) atomicMax(&a.gmax[0], mymax);
        }
    }
}

static unsigned gridn(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// launches one k_frac_step for the context's dimension / precision / mode
int launch_step(pxg_ctx *ctx, int mode, const double *u, const double *uold, double *out,
                double c1, double c2, int step, bool guard, int frozen_k)
{
    FracState &F = ctx->frac;
    StepArgs a;
    a.n = ctx->n; a.ps = F.ps; a.off = ctx->off.p; a.nbr = ctx->nbr.p; a.xp = ctx->xp.p;
    a.P64 = F.P64.p; a.P32 = F.P32.p; a.health = F.health.p; a.break_step = F.break_step.p;
    a.anybroke = F.anybroke.p; a.step = step; a.u = u; a.uold = uold; a.out = out;
    a.sc = F.sc; a.tension_only = F.tension_only; a.damage = (F.damage_on && F.sc > 0.0) ? 1 : 0;
    a.chi = (1.0 + F.sc) * (1.0 + F.sc);
    a.clo = (F.sc < 1.0) ? (1.0 - F.sc) * (1.0 - F.sc) : -1.0;
    a.b0 = F.b[0]; a.b1 = F.b[1]; a.b2 = F.b[2]; a.rho = F.rho; a.c1 = c1; a.c2 = c2;
    a.mask = F.preset_mask.p; a.gmax = F.gmax.p; a.gstate = guard ? F.gstate.p : nullptr;
    a.frozen = (frozen_k > INT_MIN) ? 1 : 0; a.frozen_k = frozen_k;
    const bool f32 = (F.precision == PXG_FP32C) && !a.frozen;
    hipStream_t st = ctx->stream;
#define PXG_LAUNCH(D, FF, M)                                                                        \
    do {                                                                                            \
        constexpr int npb = 4 * (64 / StepCfg<D, FF>::S);                                           \
        a.nblk = (int)gridn(ctx->n, npb);                                                           \
        hipLaunchKernelGGL((k_frac_step<D, FF, M>), dim3(8 * gridn(a.nblk, 8)), dim3(256), 0, st, a); \
    } while (0)
#define PXG_MODES(D, FF)                                                                            \
    switch (mode) {                                                                                 \
    case MODE_CD: PXG_LAUNCH(D, FF, MODE_CD); break;                                                \
    case MODE_VAR: PXG_LAUNCH(D, FF, MODE_VAR); break;                                              \
    case MODE_ACCEL: PXG_LAUNCH(D, FF, MODE_ACCEL); break;                                          \
    default: PXG_LAUNCH(D, FF, MODE_INIT); break;                                                   \
    }
    if (ctx->dim == 2) { if (f32) { PXG_MODES(2, true) } else { PXG_MODES(2, false) } }
    else { if (f32) { PXG_MODES(3, true) } else { PXG_MODES(3, false) } }
#undef PXG_MODES
#undef PXG_LAUNCH
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

} // namespace pxg

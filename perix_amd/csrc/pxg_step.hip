// pxg_step.hip -- the matrix-free explicit PDDO fracture step (one launch).
//
// Reference: ExplicitPDDOFracElement::preprocessIteration, the per-step
// damage + Navier force loop (ExplicitPDDOFracElement.cpp:237-299), and the
// central-difference update of TimeStepping::solveExplicitDynamics
// (TimeStepping.cpp:417-599).
//
// Work layout.  A 256-thread block owns 4*64/S nodes; each wavefront owns
// 64/S of them and gives each node S adjacent lanes.
//  * Neighbour rows are staged in LDS first.  Blocks own compact tiles of
//    nodes (2 x 4 x 4 cells on a structured grid, 2 x 4 in 2D, via the tile
//    order perm), whose horizon union is about 14 rows per node in 3D and 4
//    in 2D; the block copies those rows {x_j, u_j} into LDS with coalesced
//    loads, and a bond reads its neighbour from the LDS slot recorded for it
//    at setup.  Gathering straight from global memory costs one cache-line
//    lookup per lane and load, which saturates the texture units long
//    before HBM (the kernel keeps that path for blocks whose rows exceed
//    the LDS budget).
//  * Dependent loads come in two rounds before the family loop.  Round one
//    reads what needs no other load: the node record of the lane's slot
//    (node id, bond range, shared-run offset), its BC record, the block's
//    row ids and the guard slots.  Round two reads the node's own u, x,
//    u^{n-1} and BC values and the staged rows.  When every load a step
//    needs hung off index loads (perm, then off/bsh, then the bond groups;
//    the BC program after the family sum), the latency of that chain set
//    the time of the latency-bound 2D step.
//  * The per-bond streams use the kernel's own bond layout (bsh: line-aligned
//    node runs in tile order), where a node's bonds are contiguous; the lanes walk them in
//    groups of G consecutive bonds aligned to G, so one G-wide vector load
//    fetches a group's LDS slots (or ids), its health bytes and each operator
//    component (all bond arrays are allocated with zeroed slack and the
//    operator components are strided by a multiple of 64, so every such load
//    is aligned and in bounds).  A group that straddles two families is
//    loaded by both nodes and masked by each.
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
// FP32C: the operators are stored in FP32 (the bytes the step streams) and
// every bond force is formed and summed in FP64: an FP64 fma on an FP32
// operator costs the same issue slot as an FP32 fma on CDNA4 (packed FP32 is
// not used), and it makes the bond force exact up to the operators' own
// FP32 rounding -- the result an FP32-pair (error-free product + TwoSum)
// evaluation reaches at ten times the instructions.  The error against FP64
// is then a fixed perturbation of the operator, with no rounding noise added
// per step.  The S lanes of a node combine their FP64 sums with a lane
// butterfly (DPP within 16 lanes).
#include "pxg_internal.h"
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <thread>

namespace pxg {

struct StepArgs {
    int n;
    int nblk;                  // logical blocks (grid is rounded up to 8*)
    long long pst;             // component stride of the operator arrays
    const int *bsh;            // tiled bond layout: the per-bond streams (operators,
                               // slots, health, break steps) of CSR bond t of node i
                               // sit at t + bsh[i]
    const int *nbr;
    const int4 *nrec;          // node slot q of the tiling: {node, first and end bond
                               // position, FP32C shared-run offset}
    const int4 *brec;          // its BC program record (fused path)
    const int *rcnt;           // rows of block b: ulist[b*gcap .. + rcnt[b]), -1 beyond
    const int *ulist;
    const unsigned char *bovf; // block gathers from global memory (no LDS rows)
    const unsigned short *gslot;
    int gcap;                  // LDS slots per block
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
    int *gstate;
    int frozen;                // >0: health frozen from break_step (no damage)
    int frozen_k;              // bond t is intact iff break_step[t] > frozen_k
    // fused BC presets (merged program, see BcNodeProgram) and guard history
    const int *bnode, *bmode, *inv_q;
    const double *bval0, *bvel;
    double tpre;               // time of the presets (t^{n+1})
    double *ghist64;
    int guard_on;              // fused path: blow-up guard active
};

// blow-up guard of TimeStepping.cpp:550-573 for step s (fused path): trips
// on a non-finite max|u^s| or on growth beyond 1e3 * max(1, max|u| of step
// s-50).  The launch of step s folds max|u^s| into kGuardSlots slots (a
// wave of logical block b into slot b % kGuardSlots, 256 bytes apart), so
// that no single address takes an atomic from every wave; a later launch
// reads the slots with one load per lane and reduces them.  Three slot sets
// rotate with s % 3.
__device__ __forceinline__ int guard_slot(int s, int q) { return kGuardBase + ((s % 3) * kGuardSlots + q) * kGuardStride; }

// max|u^s| in two parts: each lane loads one slot (written by an earlier
// launch, so a plain load sees the final value), and the wave reduces the
// loaded values where the decision is taken, so that the load's latency
// overlaps the work issued between the two (all 64 lanes call the reduction)
__device__ __forceinline__ unsigned long long guard_slot_load(int s, const unsigned long long *gmax)
{
    const int l = threadIdx.x & 63;
    return l < kGuardSlots ? gmax[guard_slot(s, l)] : 0ULL;
}
__device__ __forceinline__ unsigned long long guard_reduce(unsigned long long m)
{
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
        const unsigned long long o = __shfl_xor(m, k);
        m = o > m ? o : m;
    }
    return m;
}
__device__ __forceinline__ double guard_past(int s, const double *h64)
{
    return (s > 50) ? h64[(s - 50) & 63] : -1.0;
}
__device__ __forceinline__ bool guard_trips(unsigned long long m, double past, double &umax)
{
    const bool finite = m <= (unsigned long long)__double_as_longlong(DBL_MAX);
    umax = __longlong_as_double((long long)m);
    const double cap = (past >= 0.0) ? 1.0e3 * fmax(1.0, past) : INFINITY;
    return !finite || umax > cap;
}

// records the decision on step s (one thread)
__device__ __forceinline__ void guard_record(int s, bool trip, double umax, double *h64, int *gstate)
{
    if (trip) { gstate[0] = 1; gstate[1] = s; }
    else h64[s & 63] = umax;
    gstate[2] = s;
}

// Doubles per staged neighbour record {x_j, u_j} in LDS: one more than the
// record's 2*DIM values.  Lanes read the records of random slots; at 32
// bytes (2D) apart they fall on 4 distinct bank groups of the 32 four-byte
// LDS banks, at an odd number of doubles on 16 (2D plate step: 526 bank
// conflict cycles per wave against 62 cycles of LDS instruction issue).
__host__ __device__ constexpr int rec_stride(int dim) { return 2 * dim + 1; }

template <int DIM, bool F32> struct StepCfg;
// S lanes per node (S divides 64) and G bonds per lane group; both
// precisions of a dimension share the node tiling, so S is common to them.
// 2D: 4 lanes per node, and with G = 8 one round of the FP32C loop covers
// the 28-bond family of the 400x200 plate (81,200 nodes, 16 per wave: one
// round of resident waves on the GPU); 15.2 us per step against 16.4 us at
// G = 4, 21.7 us at S = 2 and 20.9 us gathering from global memory at S = 4.
// 2D FP64 keeps G = 4 (its LDS force rows grow with G: 18.4 us, 26.2 us at
// G = 8).  3D: 8 lanes per node; with 4 the 64-node tiles overflow the LDS
// row budget.
template <> struct StepCfg<2, true> { static constexpr int S = 4, G = 8; };
template <> struct StepCfg<2, false> { static constexpr int S = 4, G = 4; };
template <> struct StepCfg<3, true> { static constexpr int S = 8, G = 4; };
template <> struct StepCfg<3, false> { static constexpr int S = 8, G = 2; };

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <int B> struct VecT;
template <> struct VecT<2> { using T = unsigned short; };
template <> struct VecT<4> { using T = unsigned; };
template <> struct VecT<8> { using T = u32x2; };
template <> struct VecT<16> { using T = u32x4; };

template <bool NT, typename V>
__device__ __forceinline__ V load_hint(const V *q)
{
    if constexpr (NT) return __builtin_nontemporal_load(q);
    else return *q;
}

// G consecutive elements starting at an address aligned to their total size
// (NT: a non-temporal load, for a stream read once per step)
template <bool NT = false, typename T, int G>
__device__ __forceinline__ void vload(const T *p, T (&o)[G])
{
    constexpr int B = (int)sizeof(T) * G;
    if constexpr (B <= 16) {
        const auto v = load_hint<NT>(reinterpret_cast<const typename VecT<B>::T *>(p));
        __builtin_memcpy(o, &v, B);
    } else {
        static_assert(B % 16 == 0, "vector width");
#pragma unroll
        for (int h = 0; h < B / 16; h++) {
            const u32x4 v = load_hint<NT>(reinterpret_cast<const u32x4 *>(p) + h);
            __builtin_memcpy(reinterpret_cast<char *>(o) + 16 * h, &v, 16);
        }
    }
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

// FP64 value held by the partner lane at butterfly level M (two 32-bit moves)
template <int M>
__device__ __forceinline__ double partner64(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const float lo = __builtin_bit_cast(float, (int)(b & 0xffffffffLL));
    const float hi = __builtin_bit_cast(float, (int)(b >> 32));
    const long long rl = (long long)(unsigned)__builtin_bit_cast(int, partner<M>(lo));
    const long long rh = (long long)__builtin_bit_cast(int, partner<M>(hi));
    return __builtin_bit_cast(double, (rh << 32) | rl);
}

// sums the FP64 family sums of the S lanes of a node (every lane ends with the total)
template <int M, int S, int DIM>
__device__ __forceinline__ void butterfly64(double (&L)[DIM])
{
    if constexpr (M < S) {
#pragma unroll
        for (int k = 0; k < DIM; k++) L[k] = L[k] + partner64<M>(L[k]);
        butterfly64<2 * M, S, DIM>(L);
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

// o = P v for a symmetric per-bond operator given by its 3 (2D) or 6 (3D)
// components
template <int DIM>
__device__ __forceinline__ void pmul_sym(const double (&p)[DIM == 3 ? 6 : 3], const double (&v)[DIM], double (&o)[DIM])
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

template <int DIM>
__device__ __forceinline__ double pick(const double (&v)[DIM], int c)
{
    return (DIM == 3 && c == 2) ? v[DIM - 1] : (c == 1 ? v[1] : v[0]);
}

// stream data of one group of G consecutive bonds (one vector load per array)
template <int DIM, bool F32, int G>
struct GroupData {
    static constexpr int NC = (DIM == 3) ? 6 : 3;
    unsigned short sv[G];                  // LDS slots of the neighbours
    bool ok[G];                            // bond intact
    float p32[F32 ? NC : 1][G];
    double p64[F32 ? 1 : NC][G];
};

// NTO: the operator streams are read with non-temporal loads (each node
// reads its own operators once per step, so keeping them in the caches only
// evicts the staged rows and the other streams)
template <int DIM, bool F32, int MODE, bool FUSED, bool NTO = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_frac_step(StepArgs a)
{
    constexpr int S = StepCfg<DIM, F32>::S;
    constexpr int G = StepCfg<DIM, F32>::G;
    constexpr int NPW = 64 / S;               // nodes per wavefront
    constexpr int W = G * S;                  // bonds per node and round
    constexpr int NC = (DIM == 3) ? 6 : 3;
    constexpr int ROW = W * DIM + 1;          // LDS row per node (odd: no bank aliasing)
    constexpr int FROWS = F32 ? 0 : 4 * NPW * ROW;
    constexpr int RS = rec_stride(DIM);      // doubles per staged row record
    using GD = GroupData<DIM, F32, G>;
    // dynamic LDS: [FP64 force rows][gcap neighbour records {x, u}]
    extern __shared__ double s_dyn[];
    __shared__ unsigned long long s_gmax;     // guard maximum of the block
    double *s_f = s_dyn;
    double *s_rec = s_dyn + FROWS;
    double *s_park = s_rec + (size_t)a.gcap * RS;   // 2D FP32C: 4 x 256 parked update inputs

    if constexpr (!FUSED) {
        if (a.gstate && a.gstate[0]) return;  // blow-up guard tripped earlier
    }
    const int blk = xcd_block(blockIdx.x, a.nblk);
    if (blk < 0) return;
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const bool live_dmg = a.damage && !a.frozen;
    const int ln = lane / S, sl = lane % S;
    const int gi = (blk * 4 + wv) * NPW + ln;
    const bool nvalid = gi < a.n;
    const bool upd = nvalid && sl < DIM;      // lane sl < DIM updates component sl

    // round one: loads that depend on no other load -- the fused guard's
    // inputs (wave 0 reduces them for the block), the node and BC records of
    // the lane's slot, and the block's first 512 row ids
    int g_tripped = 0;
    unsigned long long g_max = 0;
    double g_past = -1.0;
    if constexpr (FUSED) {
        if (a.guard_on) {
            g_tripped = a.gstate[0];
            if (a.step >= 2) { if (wv == 0) g_max = guard_slot_load(a.step - 1, a.gmax); g_past = guard_past(a.step - 1, a.ghist64); }
        }
    }
    // The loads of both rounds take clamped indices rather than branches: a
    // load under a branch makes the compiler consume its value inside the
    // branch, which waits for it before any later load can issue and splits
    // the round in two.  A lane past the last node reads the last slot's
    // records and is given an empty bond range.
    const int gq = nvalid ? gi : a.n - 1;
    const int4 nr = a.nrec[gq];
    int4 br = make_int4(-1, 0, 0, -1);
    if constexpr (FUSED) br = a.brec[gq];
    const int *rows = a.ulist + (size_t)blk * a.gcap;
    // the second 256 row slots exist only where some block has more rows
    // (3D): the 2D plate's blocks stage under 256 rows in one load per thread
    const bool two = a.gcap > 256;
    const int ra0 = rows[min(tid, a.gcap - 1)];
    int rb0 = -1;
    if (two) rb0 = rows[min(tid + 256, a.gcap - 1)];
    const int nrows = a.rcnt[blk];
    const bool govf = a.bovf[blk];
    const int ida0 = tid < a.gcap ? ra0 : -1, idb0 = tid + 256 < a.gcap ? rb0 : -1;

    // [b0, b1): the node's bonds in the stream layout; po: FP32C shared run
    const int i = nr.x, b0 = nr.y, b1 = nvalid ? nr.z : nr.y;
    const int po = F32 ? nr.w : 0;
    // BC program of the node (fused path): own entry bq and its mode m_own;
    // the bcnt entries bp0.. whose mirror source it is, the first of them
    // (entry qq0, mode m_inv0, target node tn0) resolved here
    const int bq = br.x, bp0 = br.y, bcnt = br.z & 0xffff, qq0 = br.w;
    const int bmodes = (int)((unsigned)br.z >> 16);
    const int m_own = (FUSED && upd && bq >= 0) ? (bmodes >> (2 * sl)) & 3 : 0;
    const int m_inv0 = (FUSED && upd && bcnt > 0) ? (bmodes >> (8 + 2 * sl)) & 3 : 0;

    // round two: the node's own state and BC values, then the staged rows
    double ui[DIM], xi[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        ui[c] = a.u[(size_t)i * DIM + c];
        xi[c] = a.xp[(size_t)i * DIM + c];
    }
    // The update's own inputs: u^{n-1} of the lane's component (when out
    // aliases uold the lane reads its own entry before writing it; preset
    // targets written by other lanes do not use theirs), and g(0), dg/dt of
    // the one program entry whose value the lane needs first (its own direct
    // value, else the first reflection it writes).  2D grids are one round of
    // resident waves, bound by the latency of each wave, so 2D issues these
    // loads with the staging loads (after them, so that the BC loads, which
    // stay under branches, cannot hold the staging back), and FP32C, whose
    // staging phase also holds its first bond group (below), parks them in
    // LDS once the staging writes have waited for them: held through the
    // family loop they would cost a resident wave per SIMD.  3D loads them
    // at the update: its grids run many rounds, and holding them would cost
    // a resident wave per SIMD (100 VGPRs against 92).  The early form reads
    // u^{n-1} on every lane (lanes sl >= DIM read the node's last component
    // and drop it).
    constexpr bool PRE = F32 && DIM == 2;     // first bond group before the barrier
    constexpr bool EARLY = DIM == 2;
    double uo_pre = 0.0, pv0 = 0.0, pvel = 0.0;
    int tn0 = -1;
    auto load_update = [&]() {
        if constexpr (MODE != MODE_ACCEL) {
            if (EARLY || upd) uo_pre = a.uold[(size_t)i * DIM + min(sl, DIM - 1)];
        }
        if constexpr (FUSED) {
            const int qv = (m_own == 3) ? bq : (m_inv0 == 2 ? qq0 : -1);
            if (qv >= 0) { pv0 = a.bval0[qv * 3 + sl]; pvel = a.bvel[qv * 3 + sl]; }
            // read on every lane (entry 0 stands in for none): under a branch
            // the compiler consumes it there and waits for it
            tn0 = a.bnode[max(qq0, 0)];
        }
    };
    const int A = b0 & ~(G - 1);
    int broke = 0;

    // bond-array bases; offsets below are unsigned 32-bit element indices
    const float *P32c[F32 ? NC : 1];
    const double *P64c[F32 ? 1 : NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if constexpr (F32) P32c[c] = a.P32 + c * a.pst;
        else P64c[c] = a.P64 + c * a.pst;
    }
    // The compiler tests the bond states inside the frozen/live branch,
    // which waits for every load issued so far.  FP32C reads them after its
    // operators (its first group goes out with the staging loads, and the
    // wait then covers them all); 2D FP64 reads them first (measured on the
    // 400x200 plate: 10.75 us against 11.10 us for FP32C, 17.27 us against
    // 18.17 us for FP64), 3D FP64 after (160x80x40 block: 592 us against
    // 611 us).
    auto load_group = [&](unsigned t0, GD &d) {
        auto load_state = [&]() {
            if (a.frozen) {
                int bs[G];
                vload(a.break_step + t0, bs);
#pragma unroll
                for (int k = 0; k < G; k++) d.ok[k] = bs[k] > a.frozen_k;
            } else {
                unsigned char hv[G];
                vload(a.health + t0, hv);
#pragma unroll
                for (int k = 0; k < G; k++) d.ok[k] = hv[k] != 0;
            }
        };
        vload(a.gslot + t0, d.sv);
        if constexpr (!F32 && DIM == 2) load_state();
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if constexpr (F32) vload<NTO>(P32c[c] + (unsigned)((int)t0 + po), d.p32[c]);
            else vload<NTO>(P64c[c] + t0, d.p64[c]);
        }
        if constexpr (F32 || DIM == 3) load_state();
    };
    GD d;
    const int tfirst = A + G * sl;
    // FP32C reads its first bond group on every lane, outside the family
    // loop: it lies inside the node's 32-aligned run (or, for the last run,
    // the zeroed slack after it), and out of the loop its loads go out
    // together, where the loop's first load would wait for register reuse
    // of its later rounds.  2D issues them with the staging, before the
    // barrier (one memory latency fewer in its latency-bound step).
    // nbr keeps CSR order: a block without LDS rows reads bond t at t - sh
    const int sh = govf ? a.bsh[i] : 0;

    // stage the block's neighbour rows as LDS records {x_j, u_j}: two rows
    // per thread and round, every load of a round issued before its LDS
    // writes (the row ids are sorted node ids, so neighbouring lanes read
    // neighbouring nodes); a slot past the block's rows (id -1) reads the
    // lane's own node (already on its way to the same cache) and writes
    // nothing.  Rounds beyond the first read their ids first.
    double ra[2 * DIM], rb[2 * DIM];
    auto stage_load = [&](int ida, int idb) {
        const size_t ja = (size_t)(ida >= 0 ? ida : i) * DIM, jb = (size_t)(idb >= 0 ? idb : i) * DIM;
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            ra[c] = a.xp[ja + c];
            ra[DIM + c] = a.u[ja + c];
        }
        if (two) {
#pragma unroll
            for (int c = 0; c < DIM; c++) {
                rb[c] = a.xp[jb + c];
                rb[DIM + c] = a.u[jb + c];
            }
        }
    };
    auto stage_store = [&](int qa, int ida, int idb) {
        if (ida >= 0) {
#pragma unroll
            for (int c = 0; c < 2 * DIM; c++) s_rec[qa * RS + c] = ra[c];
        }
        if (two && idb >= 0) {
#pragma unroll
            for (int c = 0; c < 2 * DIM; c++) s_rec[(qa + 256) * RS + c] = rb[c];
        }
    };
    stage_load(ida0, idb0);
    if constexpr (EARLY) load_update();
    if constexpr (PRE) load_group((unsigned)tfirst, d);
    stage_store(tid, ida0, idb0);
    if constexpr (PRE) {
        s_park[tid] = uo_pre;
        s_park[256 + tid] = pv0;
        s_park[512 + tid] = pvel;
        reinterpret_cast<int *>(s_park + 768)[tid] = tn0;
    }
    for (int q0 = 512; q0 < nrows; q0 += 512) {
        const int qa = q0 + tid, qb = qa + 256;
        const int ida = qa < a.gcap ? rows[qa] : -1, idb = qb < a.gcap ? rows[qb] : -1;
        stage_load(ida, idb);
        stage_store(qa, ida, idb);
    }
    if constexpr (FUSED) {
        // wave 0 reduces the guard slots for the block, overlapped with the
        // staging loads (which the barrier waits for anyway)
        if (a.guard_on && a.step >= 2 && wv == 0) {
            const unsigned long long m = guard_reduce(g_max);
            if (lane == 0) s_gmax = m;
        }
    }
    __syncthreads();
    if constexpr (FUSED) {
        if (g_tripped) return;                // the guard tripped earlier
        if (a.guard_on && a.step >= 2) {
            // every block takes the guard decision on the previous step from
            // the same data; block 0 records it and clears the maximum slot
            // of step+1 (last used by step-2, already decided)
            double umax;
            const bool trip = guard_trips(s_gmax, g_past, umax);
            if (blockIdx.x == 0 && threadIdx.x == 0) guard_record(a.step - 1, trip, umax, a.ghist64, a.gstate);
            if (blockIdx.x == 0 && threadIdx.x < kGuardSlots) a.gmax[guard_slot(a.step + 1, threadIdx.x)] = 0;
            if (trip) return;
        }
    }

    // damage and forces of one loaded group (dead bonds give f = 0 / -0.0)
    using F64G = double[F32 ? 1 : G][DIM];
    using F32G = double[DIM];                  // FP32C: the lane's FP64 family sum
    auto compute_group = [&](unsigned t0, const GD &dd, F64G &f64, F32G &f32) {
#pragma unroll
        for (int k = 0; k < G; k++) {
            const int t = (int)t0 + k;
            // a group lies inside the node's run and the dead padding after it
            // (runs start on 32-bond lines), so its state bytes alone decide
            bool live = dd.ok[k];
            double du[DIM];
#pragma unroll
            for (int c = 0; c < DIM; c++) du[c] = 0.0;
            if (live) {
                // neighbour row: the block's LDS copy, or global memory in a
                // block whose neighbourhood did not fit the LDS slots
                double uj[DIM], xj[DIM];
                if (!govf) {
                    const double *rec = s_rec + dd.sv[k] * RS;
#pragma unroll
                    for (int c = 0; c < DIM; c++) {
                        xj[c] = rec[c];
                        uj[c] = rec[DIM + c];
                    }
                } else {
                    const size_t q = (size_t)(a.nbr[t - sh] - 1) * DIM;
#pragma unroll
                    for (int c = 0; c < DIM; c++) {
                        uj[c] = a.u[q + c];
                        xj[c] = live_dmg ? a.xp[q + c] : 0.0;
                    }
                }
#pragma unroll
                for (int c = 0; c < DIM; c++) du[c] = uj[c] - ui[c];
                if (live_dmg) {
                    // stretch of ExplicitPDDOFracElement.cpp:249-266
                    double xs[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; c++) xs[c] = xj[c] - xi[c];
                    // 2D FP32C forms q and the band with fused multiply-adds
                    // (plate 400x200: 10.36 us against 10.55 us; the other
                    // forms measured slower with them, the 3D ones by up to
                    // 20 %): they move q by rounding far inside the band, and
                    // inside it the reference's expression, rounded as the
                    // reference rounds it, decides
                    constexpr bool FMAQ = DIM == 2 && F32;
                    double yv[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; c++) yv[c] = xs[c] + du[c];
                    double rr, yy2;
                    if constexpr (FMAQ) {
                        rr = fma(xs[0], xs[0], xs[1] * xs[1]);
                        yy2 = fma(yv[0], yv[0], yv[1] * yv[1]);
                    } else {
                        rr = xs[0] * xs[0] + xs[1] * xs[1];
                        yy2 = yv[0] * yv[0] + yv[1] * yv[1];
                        if (DIM == 3) {
                            rr = rr + xs[DIM - 1] * xs[DIM - 1];
                            yy2 = yy2 + yv[DIM - 1] * yv[DIM - 1];
                        }
                    }
                    if (rr > 0.0) {
                        const double band = 1.0e-9 * rr;
                        const double qh = FMAQ ? fma(-rr, a.chi, yy2) : yy2 - rr * a.chi;
                        bool brk = qh > 0.0;
                        bool unsure = fabs(qh) <= band;
                        if (!a.tension_only) {
                            const double ql = FMAQ ? fma(-rr, a.clo, yy2) : yy2 - rr * a.clo;
                            brk = brk || ql < 0.0;
                            unsure = unsure || fabs(ql) <= band;
                        }
                        if (unsure) {
                            // near the threshold: the reference's own expression
                            double re = xs[0] * xs[0] + xs[1] * xs[1];
                            double ye = yv[0] * yv[0] + yv[1] * yv[1];
                            if (DIM == 3) {
                                re = re + xs[DIM - 1] * xs[DIM - 1];
                                ye = ye + yv[DIM - 1] * yv[DIM - 1];
                            }
                            const double rl = sqrt(re);
                            const double s = (sqrt(ye) - rl) / rl;
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
                // FP32 operators, FP64 products and sums (a dead bond adds nothing)
                if (live) {
                    const auto &p = dd.p32;
                    if constexpr (DIM == 2) {
                        const double q0 = p[0][k], q1 = p[1][k], q2 = p[2][k];
                        f32[0] = fma(q2, du[1], fma(q0, du[0], f32[0]));
                        f32[1] = fma(q1, du[1], fma(q2, du[0], f32[1]));
                    } else {
                        const double q0 = p[0][k], q1 = p[1][k], q2 = p[2][k];
                        const double q3 = p[3][k], q4 = p[4][k], q5 = p[5][k];
                        f32[0] = fma(q4, du[DIM - 1], fma(q3, du[1], fma(q0, du[0], f32[0])));
                        f32[1] = fma(q5, du[DIM - 1], fma(q1, du[1], fma(q3, du[0], f32[1])));
                        f32[DIM - 1] = fma(q2, du[DIM - 1], fma(q5, du[1], fma(q4, du[0], f32[DIM - 1])));
                    }
                }
            } else {
                // health-gated Navier force (ExplicitPDDOFracElement.cpp:268-295);
                // a dead bond contributes -0.0, the additive identity
                const auto &p = dd.p64;
                if (live) {
                    if constexpr (DIM == 2) {
                        f64[k][0] = p[0][k] * du[0] + p[2][k] * du[1];
                        f64[k][1] = p[2][k] * du[0] + p[1][k] * du[1];
                    } else {
                        f64[k][0] = p[0][k] * du[0] + p[3][k] * du[1] + p[4][k] * du[DIM - 1];
                        f64[k][1] = p[3][k] * du[0] + p[1][k] * du[1] + p[5][k] * du[DIM - 1];
                        f64[k][DIM - 1] = p[4][k] * du[0] + p[5][k] * du[1] + p[2][k] * du[DIM - 1];
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
        F32G acc;              // the lane's FP64 family sum per component
#pragma unroll
        for (int c = 0; c < DIM; c++) acc[c] = 0.0;
        F64G f64;
        if constexpr (!PRE) load_group((unsigned)tfirst, d);
        if (tfirst < b1) compute_group((unsigned)tfirst, d, f64, acc);
        for (int t0 = tfirst + W; t0 < b1; t0 += W) {
            load_group((unsigned)t0, d);
            compute_group((unsigned)t0, d, f64, acc);
        }
        butterfly64<1, S, DIM>(acc);
        Lown = pick(acc, sl);
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
                load_group((unsigned)t0, d);
                F32G f32;
                compute_group((unsigned)t0, d, f64, f32);
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
    auto fold = [&](double w) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(w));
        mymax = bits > mymax ? bits : mymax;
    };
    if constexpr (!EARLY) load_update();
    if constexpr (PRE) {
        uo_pre = s_park[tid];
        pv0 = s_park[256 + tid];
        pvel = s_park[512 + tid];
        tn0 = reinterpret_cast<const int *>(s_park + 768)[tid];
    }
    if (upd) {
        const int c = sl;
        const size_t idx = (size_t)i * DIM + c;
        // preset mode of this component (fused path): 0 none, 1 copy and
        // 2 reflection (written by the mirror source's lane), 3 direct
        double v = 0.0;
        if (m_own == 0) {
            const double bb = (c == 0) ? a.b0 : (c == 1 ? a.b1 : a.b2);
            const double acc = (Lown + bb) / a.rho;
            const double uu = pick(ui, c);
            if (MODE == MODE_CD) v = 2.0 * uu - uo_pre + a.c1 * acc;
            else if (MODE == MODE_VAR) v = uu + a.c1 * (uu - uo_pre) + a.c2 * acc;
            else if (MODE == MODE_ACCEL) v = acc;
            else v = uu - a.c1 * uo_pre + a.c2 * acc;
            a.out[idx] = v;
            if (FUSED || ((MODE == MODE_CD || MODE == MODE_VAR) && !a.mask[i])) fold(v);
        } else if (m_own == 3) {
            // direct Dirichlet value (DirichletBC.cpp:107-147)
            v = pv0 + pvel * a.tpre;
            a.out[idx] = v;
            fold(v);
        }
        if constexpr (FUSED) {
            // this node as the mirror source of preset targets (never itself
            // a target: the program has no chains), BCSystem.cpp:114-129
            for (int k = 0; k < bcnt; k++) {
                int qq = qq0, mm = m_inv0, tn = tn0;
                if (k > 0) {
                    qq = a.inv_q[bp0 + k];
                    mm = a.bmode[qq * 3 + c];
                    tn = (mm == 1 || mm == 2) ? a.bnode[qq] : -1;
                }
                if (mm != 1 && mm != 2) continue;
                double w = v;
                if (mm == 2) {
                    const double g = (k == 0 && m_own != 3) ? pv0 + pvel * a.tpre
                                                            : a.bval0[qq * 3 + c] + a.bvel[qq * 3 + c] * a.tpre;
                    w = 2.0 * g - v;
                }
                a.out[(size_t)tn * DIM + c] = w;
                fold(w);
            }
        }
    }
    if (MODE == MODE_CD || MODE == MODE_VAR) {
        // blow-up guard maximum: |u| bit patterns order like the values
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const unsigned long long o = __shfl_xor(mymax, m);
            mymax = o > mymax ? o : mymax;
        }
        unsigned long long *slot = FUSED ? &a.gmax[guard_slot(a.step, blk & (kGuardSlots - 1))] : &a.gmax[0];
        if (lane == 0 && mymax && (!FUSED || a.guard_on)) {
            // the workgroup-scope load is a cached hint that can only lag the
            // slot's value (it only grows during the launch), so it skips
            // atomics that cannot raise it and never skips one that can
            const unsigned long long cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (mymax > cur) atomicMax(slot, mymax);
        }
    }
}

// guard decision on the last step of a fused run
__global__ void k_guard_final(int s, unsigned long long *gmax, double *h64, int *gstate, int *mirror)
{
    if (!gstate[0] && s >= 1) {                // uniform over the one wave
        const unsigned long long m = guard_reduce(guard_slot_load(s, gmax));
        if (threadIdx.x == 0) {
            double umax;
            const bool trip = guard_trips(m, guard_past(s, h64), umax);
            guard_record(s, trip, umax, h64, gstate);
        }
    }
    // the run's outcome, straight into pinned host memory: the host reads it
    // after the stream synchronisation without a copy
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) __hip_atomic_store(&mirror[c], gstate[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

static unsigned gridn(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// LDS doubles of the FP64 family-sum rows (zero for FP32C)
template <int D, bool FF>
constexpr size_t step_force_rows()
{
    return FF ? 0 : (size_t)4 * (64 / StepCfg<D, FF>::S) * (StepCfg<D, FF>::G * StepCfg<D, FF>::S * D + 1);
}

// LDS doubles where the 2D FP32C step parks its update inputs (4 per thread)
template <int D, bool FF>
constexpr size_t step_park()
{
    return (D == 2 && FF) ? 4 * 256 : 0;
}

// nodes per step block; both precisions tile nodes alike, so one LDS
// neighbour table serves either
template <int D>
constexpr int step_npb()
{
    static_assert(StepCfg<D, true>::S == StepCfg<D, false>::S, "precisions must share the node tiling");
    return 4 * (64 / StepCfg<D, true>::S);
}

// nodes per block of one launched configuration
static int launch_npb(int dim, bool f32)
{
    if (dim == 3) return f32 ? 4 * (64 / StepCfg<3, true>::S) : 4 * (64 / StepCfg<3, false>::S);
    return f32 ? 4 * (64 / StepCfg<2, true>::S) : 4 * (64 / StepCfg<2, false>::S);
}


// launches one k_frac_step for the context's dimension / precision / mode
static int launch_any(pxg_ctx *ctx, int mode, const double *u, const double *uold, double *out, double c1,
                      double c2, int step, bool guard, int frozen_k, bool fused, double tpre)
{
    FracState &F = ctx->frac;
    StepArgs a;
    a.n = ctx->n; a.pst = F.pst; a.bsh = F.bsh.p; a.nbr = ctx->nbr.p; a.xp = ctx->xp.p;
    a.nrec = F.nrec.p; a.brec = F.brec.p; a.rcnt = F.rcnt.p; a.ulist = F.ulist.p; a.bovf = F.bovf.p;
    a.gslot = F.gslot.p; a.gcap = F.gcap;
    a.P64 = F.P64.p; a.P32 = F.P32.p; a.health = F.health.p; a.break_step = F.break_step.p;
    a.anybroke = F.anybroke.p; a.step = step; a.u = u; a.uold = uold; a.out = out;
    a.sc = F.sc; a.tension_only = F.tension_only; a.damage = (F.damage_on && F.sc > 0.0) ? 1 : 0;
    a.chi = (1.0 + F.sc) * (1.0 + F.sc);
    a.clo = (F.sc < 1.0) ? (1.0 - F.sc) * (1.0 - F.sc) : -1.0;
    a.b0 = F.b[0]; a.b1 = F.b[1]; a.b2 = F.b[2]; a.rho = F.rho; a.c1 = c1; a.c2 = c2;
    a.mask = F.preset_mask.p; a.gmax = F.gmax.p; a.gstate = guard ? F.gstate.p : nullptr;
    a.frozen = (frozen_k > INT_MIN) ? 1 : 0; a.frozen_k = frozen_k;
    a.bnode = F.bc.node.p; a.bmode = F.bc.mode.p;
    a.inv_q = F.bc.inv_q.p; a.bval0 = F.bc.val0.p; a.bvel = F.bc.vel.p;
    a.tpre = tpre; a.ghist64 = F.ghist64.p; a.guard_on = guard ? 1 : 0;
    const bool f32 = (F.precision == PXG_FP32C) && !a.frozen;
    PXG_CHECK(F.npb == (ctx->dim == 3 ? step_npb<3>() : step_npb<2>()) && F.rcnt.p && F.ulist.p && F.nrec.p &&
              F.brec.p && F.nrec.n == (size_t)ctx->n, PXG_ERR_STATE,
              "launch_step: neighbour staging tables missing (pxg_frac_setup)");
    PXG_CHECK(!fused || mode == MODE_CD || mode == MODE_VAR, PXG_ERR_STATE,
              "launch_step_fused: CD/VAR steps only");
    hipStream_t st = ctx->stream;
    a.nblk = (int)gridn(ctx->n, launch_npb(ctx->dim, f32));
    const dim3 grid(8 * gridn(a.nblk, 8));
    // Non-temporal operator loads for the stepping loop when every node reads
    // its own operators (no shared FP32C runs, which are re-read from L2).
    // FP64 operators that fit the 256 MB Infinity Cache stay cached from step
    // to step instead (400x200 plate: 18.4 us cached against 19.0 us
    // non-temporal; 160x80x40 block FP64 706 -> 656 us, FP32C with own
    // operators 425 -> 393 us, 2D FP32C 15.2 -> 14.5 us).
    const size_t opbytes = (size_t)F.pst * (ctx->dim == 3 ? 6 : 3) * (f32 ? sizeof(float) : sizeof(double));
    const bool nto = fused && !(f32 && F.pofs.p) && (f32 || opbytes >= ((size_t)128 << 20));
#define PXG_LAUNCH(D, FF, M, FU, NT)                                                                \
    hipLaunchKernelGGL((k_frac_step<D, FF, M, FU, NT>), grid, dim3(256),                              \
                       sizeof(double) * (step_force_rows<D, FF>() + (size_t)F.gcap * rec_stride(D) + step_park<D, FF>()), st, a)
#define PXG_FUSED(D, FF, M)                                                                         \
    if (!fused) PXG_LAUNCH(D, FF, M, false, false);                                                 \
    else if (nto) PXG_LAUNCH(D, FF, M, true, true);                                                 \
    else PXG_LAUNCH(D, FF, M, true, false);
#define PXG_MODES(D, FF)                                                                            \
    switch (mode) {                                                                                 \
    case MODE_CD: PXG_FUSED(D, FF, MODE_CD) break;                                                  \
    case MODE_VAR: PXG_FUSED(D, FF, MODE_VAR) break;                                                \
    case MODE_ACCEL: PXG_LAUNCH(D, FF, MODE_ACCEL, false, false); break;                            \
    default: PXG_LAUNCH(D, FF, MODE_INIT, false, false); break;                                     \
    }
    if (ctx->dim == 2) { if (f32) { PXG_MODES(2, true) } else { PXG_MODES(2, false) } }
    else { if (f32) { PXG_MODES(3, true) } else { PXG_MODES(3, false) } }
#undef PXG_MODES
#undef PXG_FUSED
#undef PXG_LAUNCH
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

// launches one k_frac_step for the context's dimension / precision / mode
int launch_step(pxg_ctx *ctx, int mode, const double *u, const double *uold, double *out,
                double c1, double c2, int step, bool guard, int frozen_k)
{
    return launch_any(ctx, mode, u, uold, out, c1, c2, step, guard, frozen_k, false, 0.0);
}

int launch_step_fused(pxg_ctx *ctx, int mode, const double *u, const double *uold, double *out,
                      double c1, double c2, int step, double tpre, bool guard, int frozen_k)
{
    return launch_any(ctx, mode, u, uold, out, c1, c2, step, guard, frozen_k, true, tpre);
}

int launch_guard_final(pxg_ctx *ctx, int step)
{
    FracState &F = ctx->frac;
    hipLaunchKernelGGL(k_guard_final, dim3(1), dim3(64), 0, ctx->stream, step, F.gmax.p, F.ghist64.p, F.gstate.p,
                       F.gstate_host);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

// ---------------------------------------------------------------------------
// One reverse step of the adjoint with the E-gradient (the derivative of the
// frozen-health step above; the recursion is stated in pxg_adjoint.hip):
//
//   wbar = S^T lam_k,  ew = E .* wbar
//   Ebar_i += c (Khat_k u^{k-1})_i . wbar_i
//   lam_{k-1,i} = lamp_i + 2 wbar_i + c ((Khat_k^T ew)_i),   lamp_i <- -wbar_i
//
// on the step kernel's tiling: the block stages its neighbour rows
// {u^{k-1}_j, ew_j} in LDS, forming ew_j = E_j (S^T lam_k)_j while it stages
// (so the transposed BC preset needs no launch of its own), and walks each
// family in aligned bond groups of the tiled streams (slots, break steps,
// Phat and its pre-gathered transpose PhatT).  lam_k is read from lam_in and
// lam_{k-1} written to lam_out: a block may stage a row after another block
// has finished its node.
struct AdjStepArgs {
    int n, nblk, k;            // bond t is intact iff bs[t] > k
    long long pst;
    const int4 *nrec;          // node records of the step tiling
    const int *rcnt, *ulist;
    const unsigned char *bovf;
    const unsigned short *gslot;
    int gcap;
    const int *nbr, *bsh;      // global gathers of blocks without LDS rows
    const int *bs;             // break steps, stream layout
    const double *Phs, *PhTs;  // Phat and PhatT, stream layout
    const double *u;           // u^{k-1}
    const double *E;
    const int4 *trec;          // per node: its S^T program (adj_st_lam)
    const int *inv_q, *bmode, *bnode;
    const double *lam_in;
    double *lam_out, *lamp, *Ebar;
    double *bpart;             // dJ/db: per-block sums of wbar added to bpart[blk*3 + c], or null
    double cf;
};

// (S^T lam)_m, component c, for the merged BC program (k_preset_T's rule):
// a preset component of m takes none of its own lam, and m as the mirror
// source of a copied or reflected target g adds +lam_g or -lam_g, in program
// order.  tr = {own modes | first inverse entry's modes << 8 | entries << 16,
// target node of the first entry, first entry}.
template <int DIM>
__device__ __forceinline__ double adj_st_lam(int m, int c, int4 tr, const double *lam, const int *inv_q,
                                             const int *bmode, const int *bnode)
{
    double v = ((tr.x >> (2 * c)) & 3) ? 0.0 : lam[(size_t)m * DIM + c];
    const int cnt = (int)((unsigned)tr.x >> 16);
    for (int p = 0; p < cnt; p++) {
        int md, g;
        if (p == 0) {
            md = (tr.x >> (8 + 2 * c)) & 3;
            g = tr.y;
        } else {
            const int qq = inv_q[tr.z + p];
            md = bmode[qq * 3 + c];
            g = bnode[qq];
        }
        if (md == 1) v += lam[(size_t)g * DIM + c];
        else if (md == 2) v += -lam[(size_t)g * DIM + c];
    }
    return v;
}

template <int DIM> struct AdjStepCfg { static constexpr int S = StepCfg<DIM, false>::S, G = 2; };

template <int DIM>
__global__ void __launch_bounds__(256) k_adj_step(AdjStepArgs a)
{
    constexpr int S = AdjStepCfg<DIM>::S, G = AdjStepCfg<DIM>::G;
    constexpr int NPW = 64 / S, W = G * S;
    constexpr int NC = (DIM == 3) ? 6 : 3;
    constexpr int RS = rec_stride(DIM);
    extern __shared__ double s_dyn[];
    double *s_rec = s_dyn;
    const int blk = xcd_block(blockIdx.x, a.nblk);
    if (blk < 0) return;
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int ln = lane / S, sl = lane % S;
    const int gi = (blk * 4 + wv) * NPW + ln;
    const bool nvalid = gi < a.n;
    const bool upd = nvalid && sl < DIM;

    // round one: the node record and the block's row ids
    int4 nr = make_int4(0, 0, 0, 0);
    if (nvalid) nr = a.nrec[gi];
    const int *rows = a.ulist + (size_t)blk * a.gcap;
    const int ida0 = tid < a.gcap ? rows[tid] : -1, idb0 = tid + 256 < a.gcap ? rows[tid + 256] : -1;
    const int nrows = a.rcnt[blk];
    const bool govf = a.bovf[blk];
    const int i = nr.x, b0 = nr.y, b1 = nr.z;

    // round two: the node's own u, E, S^T program and lam, then the rows
    double ui[DIM];
    double wb = 0.0, lp = 0.0;      // the lane's component of wbar_i, lamp_i
    double wi[DIM], ewi[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) ui[c] = nvalid ? a.u[(size_t)i * DIM + c] : 0.0;
    {
        const int4 tr = nvalid ? a.trec[i] : make_int4(0, 0, 0, 0);
        const double Ei = nvalid ? a.E[i] : 0.0;
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            wi[c] = nvalid ? adj_st_lam<DIM>(i, c, tr, a.lam_in, a.inv_q, a.bmode, a.bnode) : 0.0;
            ewi[c] = Ei * wi[c];
        }
    }
    if (upd) lp = a.lamp[(size_t)i * DIM + sl];
    wb = pick(wi, sl);
    const int sh = (govf && nvalid) ? a.bsh[i] : 0;
    // stage {u_j, ew_j}: rounds beyond the first read their ids first
    auto row_of = [&](int j, double (&r)[2 * DIM]) {
        const int4 tr = a.trec[j];
        const double Ej = a.E[j];
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            r[c] = a.u[(size_t)j * DIM + c];
            r[DIM + c] = Ej * adj_st_lam<DIM>(j, c, tr, a.lam_in, a.inv_q, a.bmode, a.bnode);
        }
    };
    for (int q0 = 0; q0 == 0 || q0 < nrows; q0 += 512) {
        const int qa = q0 + tid, qb = qa + 256;
        const int ida = q0 == 0 ? ida0 : (qa < a.gcap ? rows[qa] : -1);
        const int idb = q0 == 0 ? idb0 : (qb < a.gcap ? rows[qb] : -1);
        double ra[2 * DIM], rb[2 * DIM];
        if (ida >= 0) row_of(ida, ra);
        if (idb >= 0) row_of(idb, rb);
        if (ida >= 0) {
#pragma unroll
            for (int c = 0; c < 2 * DIM; c++) s_rec[qa * RS + c] = ra[c];
        }
        if (idb >= 0) {
#pragma unroll
            for (int c = 0; c < 2 * DIM; c++) s_rec[qb * RS + c] = rb[c];
        }
    }
    __syncthreads();

    // family sums over the intact bonds: ku = sum Phat (u_j - u_i),
    // zo = sum PhatT ew_j, sp = sum Phat (per-lane partial sums, then a lane
    // butterfly: a fixed order)
    double ku[DIM], zo[DIM], sp[NC];
#pragma unroll
    for (int c = 0; c < DIM; c++) { ku[c] = 0.0; zo[c] = 0.0; }
#pragma unroll
    for (int c = 0; c < NC; c++) sp[c] = 0.0;
    for (int t0 = (b0 & ~(G - 1)) + G * sl; t0 < b1; t0 += W) {
        unsigned short sv[G];
        int bv[G];
        double p[NC][G], pt[NC][G];
        vload(a.gslot + t0, sv);
        vload(a.bs + t0, bv);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            vload(a.Phs + c * a.pst + t0, p[c]);
            vload(a.PhTs + c * a.pst + t0, pt[c]);
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int t = t0 + g;
            if (t < b0 || t >= b1 || bv[g] <= a.k) continue;
            double uj[DIM], ej[DIM];
            if (!govf) {
                const double *rec = s_rec + sv[g] * RS;
#pragma unroll
                for (int c = 0; c < DIM; c++) { uj[c] = rec[c]; ej[c] = rec[DIM + c]; }
            } else {
                double r[2 * DIM];
                row_of(a.nbr[t - sh] - 1, r);
#pragma unroll
                for (int c = 0; c < DIM; c++) { uj[c] = r[c]; ej[c] = r[DIM + c]; }
            }
            double pc[NC], ptc[NC], du[DIM], o[DIM];
#pragma unroll
            for (int c = 0; c < NC; c++) { pc[c] = p[c][g]; ptc[c] = pt[c][g]; sp[c] += pc[c]; }
#pragma unroll
            for (int c = 0; c < DIM; c++) du[c] = uj[c] - ui[c];
            pmul_sym<DIM>(pc, du, o);
#pragma unroll
            for (int c = 0; c < DIM; c++) ku[c] += o[c];
            pmul_sym<DIM>(ptc, ej, o);
#pragma unroll
            for (int c = 0; c < DIM; c++) zo[c] += o[c];
        }
    }
    butterfly64<1, S, DIM>(ku);
    butterfly64<1, S, DIM>(zo);
    butterfly64<1, S, NC>(sp);
    if (a.bpart) {
        // the body force enters every central-difference value: dJ/db +=
        // c S^T lam, summed per block in a fixed order (reproducible)
        __shared__ double red[4][3];
        double v[3] = {0.0, 0.0, 0.0};
        if (nvalid && sl == 0)
#pragma unroll
            for (int c = 0; c < DIM; c++) v[c] = wi[c];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] += __shfl_xor(v[c], o, 64);
        if (lane == 0)
            for (int c = 0; c < 3; c++) red[wv][c] = v[c];
        __syncthreads();
        if (tid == 0)
            for (int c = 0; c < DIM; c++) a.bpart[blk * 3 + c] += ((red[0][c] + red[1][c]) + (red[2][c] + red[3][c]));
    }
    if (!upd) return;
    if (sl == 0) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < DIM; c++) s += ku[c] * wi[c];
        a.Ebar[i] += a.cf * s;
    }
    double d[DIM];
    pmul_sym<DIM>(sp, ewi, d);
    const size_t q = (size_t)i * DIM + sl;
    const double z = pick(zo, sl) - pick(d, sl);
    a.lam_out[q] = lp + 2.0 * wb + a.cf * z;
    a.lamp[q] = -wb;
}

// the step tiling's staging tables with the adjoint's streams: S^T program
// records per node (trec) and the operators in the stream layout
__global__ void k_adj_trec(int n, const int *prog_of, const int *bmode, const int *inv_off, const int *inv_q,
                           const int *bnode, int4 *trec)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int q = prog_of[m], p0 = inv_off[m], cnt = inv_off[m + 1] - p0;
    int x = 0, g0 = -1;
    if (q >= 0)
        for (int c = 0; c < 3; c++) x |= (bmode[q * 3 + c] & 3) << (2 * c);
    if (cnt > 0) {
        const int qq = inv_q[p0];
        for (int c = 0; c < 3; c++) x |= (bmode[qq * 3 + c] & 3) << (8 + 2 * c);
        g0 = bnode[qq];
    }
    trec[m] = make_int4(x | (cnt << 16), g0, p0, 0);
}

__global__ void k_to_stream(int n, int nc, const int *off, const int *bsh, long long ps, long long pst,
                            const double *src, double *dst)
{
    const int i = blockIdx.x * kConvNodes + threadIdx.x / kConvLanes, l = threadIdx.x % kConvLanes;
    if (i >= n) return;
    for (int t = off[i] + l; t < off[i + 1]; t += kConvLanes)
        for (int c = 0; c < nc; c++) dst[(size_t)c * pst + t + bsh[i]] = src[(size_t)c * ps + t];
}

int adj_step_setup(pxg_ctx *ctx)
{
    FracState &F = ctx->frac;
    AdjWork &A = F.adj;
    hipStream_t st = ctx->stream;
    const int n = ctx->n, nc = (ctx->dim == 3) ? 6 : 3;
    PXG_CHECK(F.nrec.p && F.bsh.p && F.Phat.p && F.PhatT.p, PXG_ERR_STATE, "adj_step_setup: step tiling missing");
    PXG_HIP(A.trec.alloc(n));
    hipLaunchKernelGGL(k_adj_trec, dim3(gridn(n, 256)), dim3(256), 0, st, n, F.bc.prog_of.p, F.bc.mode.p,
                       F.bc.inv_off.p, F.bc.inv_q.p, F.bc.node.p, A.trec.p);
    PXG_HIP(A.phs.alloc((size_t)F.pst * nc));
    PXG_HIP(A.phts.alloc((size_t)F.pst * nc));
    PXG_HIP(hipMemsetAsync(A.phs.p, 0, sizeof(double) * (size_t)F.pst * nc, st));
    PXG_HIP(hipMemsetAsync(A.phts.p, 0, sizeof(double) * (size_t)F.pst * nc, st));
    hipLaunchKernelGGL(k_to_stream, dim3(gridn(n, kConvNodes)), dim3(kConvBlock), 0, st, n, nc, ctx->off.p, F.bsh.p,
                       F.ps, F.pst, F.Phat.p, A.phs.p);
    hipLaunchKernelGGL(k_to_stream, dim3(gridn(n, kConvNodes)), dim3(kConvBlock), 0, st, n, nc, ctx->off.p, F.bsh.p,
                       F.ps, F.pst, F.PhatT.p, A.phts.p);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

int adj_step_blocks(pxg_ctx *ctx)
{
    return (int)gridn(ctx->n, 4 * (64 / (ctx->dim == 3 ? AdjStepCfg<3>::S : AdjStepCfg<2>::S)));
}

int launch_adj_step(pxg_ctx *ctx, int k, const double *u, const double *lam_in, double *lam_out, double *lamp,
                    double *Ebar, double *bpart, double cf)
{
    FracState &F = ctx->frac;
    AdjWork &A = F.adj;
    AdjStepArgs a;
    a.n = ctx->n; a.k = k; a.pst = F.pst;
    a.nrec = F.nrec.p; a.rcnt = F.rcnt.p; a.ulist = F.ulist.p; a.bovf = F.bovf.p; a.gslot = F.gslot.p;
    a.gcap = F.gcap; a.nbr = ctx->nbr.p; a.bsh = F.bsh.p; a.bs = F.break_step.p;
    a.Phs = A.phs.p; a.PhTs = A.phts.p; a.u = u; a.E = F.nodeE.p; a.trec = A.trec.p;
    a.inv_q = F.bc.inv_q.p; a.bmode = F.bc.mode.p; a.bnode = F.bc.node.p;
    a.lam_in = lam_in; a.lam_out = lam_out; a.lamp = lamp; a.Ebar = Ebar; a.bpart = bpart; a.cf = cf;
    PXG_CHECK(a.nrec && a.Phs && a.trec && a.E, PXG_ERR_STATE, "launch_adj_step: adjoint step tables missing");
    const int npb = 4 * (64 / (ctx->dim == 3 ? AdjStepCfg<3>::S : AdjStepCfg<2>::S));
    PXG_CHECK(npb == F.npb, PXG_ERR_STATE, "launch_adj_step: tiling mismatch");
    a.nblk = (int)gridn(ctx->n, npb);
    const dim3 grid(8 * gridn(a.nblk, 8));
    const size_t lds = sizeof(double) * (size_t)F.gcap * rec_stride(ctx->dim);
    if (ctx->dim == 3) hipLaunchKernelGGL(k_adj_step<3>, grid, dim3(256), lds, ctx->stream, a);
    else hipLaunchKernelGGL(k_adj_step<2>, grid, dim3(256), lds, ctx->stream, a);
    PXG_HIP(hipGetLastError());
    return PXG_OK;
}

// Builds the LDS staging tables of the step kernel.  Nodes are ordered by
// tiles of TX x TY x TZ cells (cell = floor(x / dx_char); the tile order is
// lexicographic z, y, x, and cells within a tile follow z, y, x), and step
// block b owns the npb nodes perm[b*npb ...]: on a structured grid every
// block is one compact 2 x 4 x 4 tile, whose horizon union holds about 14
// rows per node where a 32-node row would need 31.  Per block the sorted
// distinct neighbours form the row list ulist[b*gcap ..] (rcnt[b] ids, -1
// beyond), and every bond records the LDS slot of its neighbour.  Blocks
// whose neighbourhood exceeds the slot budget gather from global memory
// instead (bovf).  The per-bond operator and slot streams are stored in the
// same tile order (bsh), so each block reads one contiguous range of them.
// node runs of the per-bond streams start on a 128-byte line of the FP32
// operator arrays (32 bonds)
constexpr int kBondAlign = 32;

int build_step_gather(pxg_ctx *ctx)
{
    FracState &F = ctx->frac;
    const int n = ctx->n, dim = ctx->dim;
    const long long nb = ctx->nb;
    const int npb = (dim == 3) ? step_npb<3>() : step_npb<2>();
    F.npb = npb;
    F.nrec.release(); F.brec.release();   // rebuilt by build_step_records
    constexpr int TX = 2, TY = 4, TZ = 4;
    const int capmax = 640;   // rows per block: 640 x 56 B = 35 KB of LDS in 3D
    const int nblk = (int)gridn(n, npb);
    std::vector<int> off((size_t)n + 1), nbr((size_t)(nb > 0 ? nb : 1));
    PXG_HIP(hipMemcpyAsync(off.data(), ctx->off.p, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (nb > 0)
        PXG_HIP(hipMemcpyAsync(nbr.data(), ctx->nbr.p, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, ctx->stream));
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    // tile order of the nodes
    const double *X = ctx->hcoords.data();
    double lo[3] = {X[0], X[1], X[2]};
    for (int i = 1; i < n; i++)
        for (int a = 0; a < 3; a++) lo[a] = std::min(lo[a], X[3 * (size_t)i + a]);
    const double h = std::max(ctx->dxyz[0], std::max(ctx->dxyz[1], ctx->dxyz[2]));
    PXG_CHECK(h > 0.0, PXG_ERR_STATE, "build_step_gather: zero grid spacing");
    struct Key { long long k0, k1; int id; };
    std::vector<Key> keys((size_t)n);
    for (int i = 0; i < n; i++) {
        long long cc[3];
        for (int a = 0; a < 3; a++) cc[a] = (long long)std::floor((X[3 * (size_t)i + a] - lo[a]) / h + 0.5);
        const long long t[3] = {cc[0] / TX, cc[1] / TY, cc[2] / TZ};
        keys[(size_t)i].k0 = (t[2] << 42) | (t[1] << 21) | t[0];
        keys[(size_t)i].k1 = ((cc[2] % TZ) << 42) | ((cc[1] % TY) << 21) | (cc[0] % TX);
        keys[(size_t)i].id = i;
    }
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        return a.k0 != b.k0 ? a.k0 < b.k0 : (a.k1 != b.k1 ? a.k1 < b.k1 : a.id < b.id);
    });
    std::vector<int> perm((size_t)n);
    for (int q = 0; q < n; q++) perm[(size_t)q] = keys[(size_t)q].id;
    // tiled bond layout of the per-bond streams: the nodes' bond runs in
    // perm order, so a block's streams are one contiguous range, and every
    // run starts on a 128-byte line of the FP32 operator arrays (32 bonds;
    // one round of the node's 8 lanes reads exactly one line per array).
    // Runs that shared a line with the next node made the line be fetched
    // twice, once per node, because the many streams in flight evict it
    // from L2 in between.
    std::vector<int> bsh((size_t)n);
    long long pos = 0;
    for (int q = 0; q < n; q++) {
        const int i = perm[(size_t)q];
        pos = (pos + kBondAlign - 1) / kBondAlign * kBondAlign;
        bsh[(size_t)i] = (int)(pos - off[(size_t)i]);
        pos += off[(size_t)i + 1] - off[(size_t)i];
    }
    PXG_CHECK(pos < INT_MAX - 64, PXG_ERR_ARG, "build_step_gather: bond layout exceeds 2^31 entries");
    F.pst = (pos + 63) / 64 * 64;
    std::vector<unsigned short> slot((size_t)std::max(F.pst, 1LL), 0);
    std::vector<unsigned char> ovf((size_t)nblk, 0);
    std::vector<std::vector<int>> brows((size_t)nblk);
    auto work = [&](int first, int step) {
        std::vector<int> ids;
        for (int b = first; b < nblk; b += step) {
            const int q1 = std::min(n, (b + 1) * npb);
            ids.clear();
            for (int q = b * npb; q < q1; q++) {
                const int i = perm[(size_t)q];
                ids.insert(ids.end(), nbr.begin() + off[(size_t)i], nbr.begin() + off[(size_t)i + 1]);
            }
            std::sort(ids.begin(), ids.end());
            ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
            if ((int)ids.size() > capmax) { ovf[(size_t)b] = 1; continue; }   // global gathers
            for (int q = b * npb; q < q1; q++) {
                const int i = perm[(size_t)q];
                for (int t = off[(size_t)i]; t < off[(size_t)i + 1]; t++)
                    slot[(size_t)(t + bsh[(size_t)i])] =
                        (unsigned short)(std::lower_bound(ids.begin(), ids.end(), nbr[(size_t)t]) - ids.begin());
            }
            std::vector<int> &rr = brows[(size_t)b];
            rr.resize(ids.size());
            for (size_t q = 0; q < ids.size(); q++) rr[q] = ids[q] - 1;
        }
    };
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int k = 0; k < nth; k++) pool.emplace_back(work, k, nth);
    for (auto &th : pool) th.join();
    int gcap = 1;
    std::vector<int> rcnt((size_t)nblk, 0);
    for (int b = 0; b < nblk; b++) {
        rcnt[(size_t)b] = (int)brows[(size_t)b].size();
        gcap = std::max(gcap, rcnt[(size_t)b]);
    }
    PXG_CHECK((long long)nblk * gcap < INT_MAX, PXG_ERR_ARG, "build_step_gather: row table exceeds 2^31 entries");
    std::vector<int> ulist((size_t)nblk * gcap, -1);
    for (int b = 0; b < nblk; b++)
        std::copy(brows[(size_t)b].begin(), brows[(size_t)b].end(), ulist.begin() + (size_t)b * gcap);
    PXG_HIP(F.perm.alloc(perm.size()));
    PXG_HIP(F.rcnt.alloc(rcnt.size()));
    PXG_HIP(F.ulist.alloc(ulist.size()));
    PXG_HIP(F.gslot.alloc(slot.size()));
    PXG_HIP(F.bovf.alloc(ovf.size()));
    PXG_HIP(F.bsh.alloc(bsh.size()));
    hipStream_t st = ctx->stream;
    PXG_HIP(hipMemcpyAsync(F.bsh.p, bsh.data(), sizeof(int) * bsh.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(F.perm.p, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(F.rcnt.p, rcnt.data(), sizeof(int) * rcnt.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(F.ulist.p, ulist.data(), sizeof(int) * ulist.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(F.gslot.p, slot.data(), sizeof(unsigned short) * slot.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipMemcpyAsync(F.bovf.p, ovf.data(), ovf.size(), hipMemcpyHostToDevice, st));
    PXG_HIP(hipStreamSynchronize(st));
    F.gcap = gcap;
    return PXG_OK;
}

// the node and BC records of the step kernel (StepArgs::nrec, brec), one
// per node slot q of the tiling: set-up data read with one load per lane
__global__ void k_step_records(int n, const int *perm, const int *off, const int *bsh, const int *pofs,
                               const int *prog_of, const int *inv_off, const int *inv_q, const int *bmode,
                               int4 *nrec, int4 *brec)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = perm[q];
    nrec[q] = make_int4(i, off[i] + bsh[i], off[i + 1] + bsh[i], pofs ? pofs[i] : 0);
    const int bq = prog_of[i], p0 = inv_off[i], cnt = inv_off[i + 1] - p0;
    int modes = 0, qq0 = -1;
    if (bq >= 0)
        for (int c = 0; c < 3; c++) modes |= (bmode[bq * 3 + c] & 3) << (2 * c);
    if (cnt > 0) {
        qq0 = inv_q[p0];
        for (int c = 0; c < 3; c++) modes |= (bmode[qq0 * 3 + c] & 3) << (8 + 2 * c);
    }
    brec[q] = make_int4(bq, p0, cnt | (modes << 16), qq0);
}

// (re)builds the step records from the tiling, the shared runs and the BC
// program; called whenever one of them changes (pxg_frac_setup,
// pxg_frac_set_bcs).  A node mirrors at most a few ghosts, far below the
// 16-bit count of the record.
int build_step_records(pxg_ctx *ctx)
{
    FracState &F = ctx->frac;
    const int n = ctx->n;
    PXG_CHECK(F.perm.p && F.bsh.p && F.bc.prog_of.p && F.bc.inv_off.p, PXG_ERR_STATE,
              "build_step_records: tiling or BC program missing");
    PXG_HIP(F.nrec.alloc(n));
    PXG_HIP(F.brec.alloc(n));
    hipLaunchKernelGGL(k_step_records, dim3(gridn(n, 256)), dim3(256), 0, ctx->stream, n, F.perm.p, ctx->off.p,
                       F.bsh.p, F.pofs.p, F.bc.prog_of.p, F.bc.inv_off.p, F.bc.inv_q.p, F.bc.mode.p, F.nrec.p,
                       F.brec.p);
    PXG_HIP(hipGetLastError());
    PXG_HIP(hipStreamSynchronize(ctx->stream));
    return PXG_OK;
}

} // namespace pxg

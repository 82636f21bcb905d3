// pxg_internal.h -- shared state and helpers of the PDDO hot-path library.
//
// One pxg_ctx owns the device copy of one PD problem on one GPU: the point
// cloud, the horizon families, the PDDO operators, the implicit CSR pattern
// and the explicit-fracture state.  All device memory belongs to the context
// and is released by pxg_destroy.  One host thread per context.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>
#include <string>
#include <vector>
#include "pxg.h"

namespace pxg {

constexpr int kMaxNop = 9;        // Order-2 slots: 5 (2D), 9 (3D)
constexpr int kWave = 64;

// thread-local last error message (pxg_last_error)
void set_error(const std::string &msg);

#define PXG_HIP(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ::pxg::set_error(std::string("HIP error ") + hipGetErrorString(e_) +      \
                             " at " + __FILE__ + ":" + std::to_string(__LINE__) +     \
                             " (" #expr ")");                                         \
            return PXG_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

#define PXG_CHECK(cond, code, msg)                                                    \
    do {                                                                              \
        if (!(cond)) { ::pxg::set_error(msg); return (code); }                        \
    } while (0)

// Host-test fault injection (pxg_set_fault): the first entry point that
// reaches this check after the fault was set returns PXG_ERR_FAULT before
// any launch and consumes the fault, so the context works again afterwards.
#define PXG_FAULT(ctx, name)                                                          \
    do {                                                                              \
        if ((ctx)->fault) {                                                           \
            (ctx)->fault = 0;                                                         \
            ::pxg::set_error(std::string(name) + ": injected fault");                 \
            return PXG_ERR_FAULT;                                                     \
        }                                                                             \
    } while (0)

// Per-node loops over a family whose bonds are independent (conversions,
// set-up): 16 consecutive lanes walk one family, 8 nodes per 128-thread
// block, so a family's loads and stores are coalesced.
// fused-path guard slots: kGuardSlots 256-byte-spaced words per slot set,
// three sets after kGuardBase words of the separate-preset path
constexpr int kGuardBase = 8, kGuardSlots = 64, kGuardStride = 32;
constexpr int kGuardWords = kGuardBase + 3 * kGuardSlots * kGuardStride;
static_assert(kGuardSlots <= 64 && (kGuardSlots & (kGuardSlots - 1)) == 0, "one wave reads every slot");
constexpr int kConvLanes = 16, kConvBlock = 128, kConvNodes = kConvBlock / kConvLanes;

// device buffer with size bookkeeping.  Every allocation carries at least 64
// zeroed elements of slack past `count` (rounded to a multiple of 64), so
// kernels may read aligned vectors that run past the last element.
template <typename T>
struct dbuf {
    T *p = nullptr;
    size_t n = 0;
    static constexpr size_t SLACK = 64;
    static size_t padded(size_t count) { return (count + 2 * SLACK - 1) / SLACK * SLACK; }
    hipError_t alloc(size_t count) {
        if (count == n && p) return hipSuccess;
        release();
        n = count;
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, sizeof(T) * padded(count));
        if (e != hipSuccess) { p = nullptr; n = 0; return e; }
        e = hipMemset(p, 0, sizeof(T) * padded(count));
        // the context streams are non-blocking: finish the fill before any
        // of them can touch the buffer
        return e != hipSuccess ? e : hipDeviceSynchronize();
    }
    void release() {
        if (p) (void)hipFree(p);       // hipFree waits for the device to finish with it
        p = nullptr;
        n = 0;
    }
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
};

struct BcNodeProgram {
    // one entry per preset-target node (all its components handled together)
    dbuf<int> node;          // 0-based target node
    dbuf<int> src;           // 0-based mirror bulk node (-1 none)
    dbuf<int> mode;          // per node*3+comp: 0 untouched, 1 copy, 2 reflect, 3 direct
    dbuf<double> val0;       // per node*3+comp: g(0)
    dbuf<double> vel;        // per node*3+comp: dg/dt
    int count = 0;
    bool merged = true;      // single-stage (parallel) program
    dbuf<int> prog_of;       // N: program entry of a target node, -1 otherwise
    dbuf<int> inv_off;       // N+1: program entries whose mirror source is node m
    dbuf<int> inv_q;
};

// BC row program of the implicit CSR system: one entry per constrained row
// (the last BC entry that reaches the row), sorted by row
struct CsrBcProgram {
    dbuf<int> row;           // constrained dof row
    dbuf<int> dslot;         // CSR index of the row's diagonal
    dbuf<int> pslot;         // CSR index of (row, partner dof), -1 none
    dbuf<int> pdof;          // partner dof (mirror bulk node, same component)
    dbuf<signed char> kind;  // BCROW_DIRECT, BCROW_REFLECT, BCROW_MIRROR
    dbuf<double> val0, vel;  // Dirichlet g(t) = val0 + vel t
    int count = 0;
    int set = 0;             // program built for the current CSR pattern
    int chained = 0;         // some entry's partner dof is itself a constrained row
};
enum { BCROW_DIRECT = 0, BCROW_REFLECT = 1, BCROW_MIRROR = 2 };

// implicit dynamic fracture element state (PDDODynamicFracElement): bond
// health (CSR order), committed velocity and displacement history
// stress-coupled Cahn-Hilliard element: C11, C12, C66, A, Kh, Ch, D, kappa,
// chi, Omega, cref
struct StressChState {
    double k[11] = {0};
    int plane_strain = 1;   // 2D stress state of the projections (s_zz)
    int ready = 0;
};

// fracture stress-coupled Cahn-Hilliard element: k = C11, C12, C66, A, Kh,
// Ch, D, kappa, chi, Omega, cref, rho, residual stiffness
struct FracSchState {
    double k[13] = {0};
    double sc = 0;
    int tension_only = 1, damage_on = 1;
    int plane_strain = 1;         // 2D stress state of the projections (s_zz)
    dbuf<unsigned char> health;   // nb, CSR order
    dbuf<double> lapc;            // nb species Laplacian (0 on ghost bonds)
    dbuf<double> vel, uprev;      // n*dim
    int ready = 0;
};

struct DynFracState {
    int ready = 0;
    double C11 = 0, C12 = 0, C66 = 0, rho = 1, b[3] = {0, 0, 0}, sc = 0;
    int tension_only = 1, damage_on = 1;
    dbuf<unsigned char> health;
    dbuf<double> vel, uprev;
};

// adjoint workspace, kept between calls and reallocated on size changes
struct AdjWork {
    dbuf<double> cks;        // 2 * ncheckpoints * N*dim: (u^{k-1}, u^k) pairs
    dbuf<double> seg;        // (ckpt + 2) * N*dim: one recomputed segment
    dbuf<double> lam, lamp, wbar, ew, Ebar, dJ, eo;
    dbuf<double> bpart, gb;  // body-force gradient: per-block partials, result
    dbuf<double> spg;        // family sums of intact Phat (sweeps without the E-gradient)
    dbuf<int> cur, brk_off, brk_t;
    dbuf<int> ds;            // sensor node ids
    dbuf<double> fu[2];      // frozen forward's own u ping-pong (the run state in FracState::u stays)
    dbuf<double> lam2;       // lam ping-pong of the fused reverse step
    dbuf<double> phs, phts;  // Phat, PhatT in the step kernel's stream layout
    dbuf<int4> trec;         // per node: its S^T preset program (k_adj_step)
};

struct FracState {
    int ready = 0;
    int precision = PXG_FP64;
    int dim = 2;
    double C11 = 0, C12 = 0, C66 = 0, rho = 1, b[3] = {0, 0, 0}, sc = 0;
    int tension_only = 1, damage_on = 1;
    int opts = 0;            // PXG_FRAC_* diagnostic options (pxg_frac_options)
    // per-bond force operators P (CSR bond order): 3 (2D) or 6 (3D) components
    // component-major: component c of bond t at [c*ps + t]
    long long ps = 0;        // CSR-order component stride: nb rounded up to a multiple of 64
    // P64 / P32 (and gslot, health, break_step) use the tiled bond layout of
    // the step blocks: CSR bond t of node i at t + bsh[i], with component
    // stride pst
    long long pst = 0;
    dbuf<int> bsh;
    // FP32C shared operator runs: node i reads its FP32 operators at bond
    // position + pofs[i], the run of a representative node (2D: the same
    // FP32 bits; 3D: FP64 operators that agree to 1e-12 relative); empty
    // when nothing is shared
    dbuf<int> pofs;
    long long nshared = 0, nclasses = 0;
    dbuf<double> P64;
    dbuf<float> P32;
    // LDS neighbour staging of the step kernel: step block b owns the npb
    // nodes perm[b*npb ..] (a compact tile) and copies its rcnt[b] neighbour
    // rows ulist[b*gcap ..] (node ids in slot order, padded with -1 to gcap)
    // into LDS; bond t reads its neighbour from LDS slot gslot[t].  A block
    // whose neighbourhood exceeds the slot budget has bovf set and gathers
    // its neighbours from global memory via nbr.
    int npb = 0, gcap = 0;        // nodes per step block, LDS slots per block
    dbuf<int> perm;
    dbuf<int> rcnt;
    dbuf<int> ulist;
    // per step-block node slot q (perm order), so that one load gives what a
    // lane needs to start: nrec[q] = {i, first and end bond position of i in
    // the stream layout, FP32C shared-run offset}; brec[q] = the node's BC
    // program (fused step) {own entry or -1, first inverse entry, count |
    // own modes << 16 | first inverse entry's modes << 24 (2 bits per
    // component), program entry of the first inverse entry or -1}
    dbuf<int4> nrec, brec;
    dbuf<unsigned short> gslot;
    dbuf<unsigned char> bovf;
    dbuf<unsigned char> health;   // pst, stream layout (see bsh)
    dbuf<int> break_step;         // pst: recorded break step, INT_MAX intact
    dbuf<unsigned char> health_csr;   // nb: CSR-order copies (tiled layout only)
    dbuf<int> bs_csr;
    dbuf<double> u[2];            // ping-pong displacement buffers (N*dim)
    int cur = 0;                  // u[cur] = u^n, u[cur^1] = u^{n-1}
    dbuf<unsigned char> preset_mask;   // N: node is a preset target
    BcNodeProgram bc;
    // blow-up guard state
    // separate-preset path (k_bc_preset): [0] CD part, [1] preset part,
    // [2] block counter; fused path: max|u^{s}| of step s spread over the
    // kGuardSlots slots of set s % 3 (kGuardWords words in all)
    dbuf<unsigned long long> gmax;
    dbuf<double> ghist;                // 50-entry ring of max|u| (separate-preset path)
    dbuf<double> ghist64;              // max|u^{s}| at [s % 64] (fused path)
    dbuf<int> gstate;                  // [0] aborted flag, [1] abort step, [2] steps done
    int *gstate_host = nullptr;        // pinned host mirror of gstate, written by k_guard_final
    // node-E (optional): per-node Young's modulus for the adjoint configuration
    dbuf<double> nodeE;
    double nu = 0.25;
    int plane_stress = 0;
    int has_nodeE = 0;
    dbuf<double> Phat;            // nb*ncomp: operators with C(E=1) (adjoint)
    dbuf<double> PhatT;           // PhatT[c*ps + t] = Phat[c*ps + rev[t]] (adjoint)
    dbuf<int> rev;                // nb: reverse bond index
    AdjWork adj;
    dbuf<double> accel;           // N*dim scratch (drop-in accel / init)
    dbuf<int> anybroke;           // [0] breaks in the last launch
    int health_uploaded = 0;
    // schedule state of the device-resident loop
    double t = 0.0, dt_prev = 0.0;
    int step = 0;
    int cur0 = 0;                 // buffer index at step 0
    // optional timing of the step kernel (pxg_frac_profile): each event pair
    // spans a run of back-to-back step launches
    int prof = 0;
    std::vector<hipEvent_t> ev;   // start/stop pairs
    int nev = 0;                  // closed pairs
    long long nlaunch = 0;        // launches inside closed pairs
    int ev_open = 0, ev_pending = 0;
};

// launch_step modes: u^{n+1} by the fixed-dt central difference, by the
// variable-dt form of the last step, the acceleration alone, or u^{-1}
enum { MODE_CD = 0, MODE_VAR = 1, MODE_ACCEL = 2, MODE_INIT = 3 };
int launch_step(pxg_ctx *ctx, int mode, const double *u, const double *uold, double *out,
                double c1, double c2, int step, bool guard, int frozen_k);
// MODE_CD / MODE_VAR step with the BC presets at time tpre folded in and,
// with `guard`, the blow-up guard (the decision on step `step`-1 is taken at
// the start of the launch; launch_guard_final decides the last step of a
// run).  frozen_k > INT_MIN freezes health from the recorded break steps.
int launch_step_fused(pxg_ctx *ctx, int mode, const double *u, const double *uold, double *out,
                      double c1, double c2, int step, double tpre, bool guard = true,
                      int frozen_k = INT_MIN);
int launch_guard_final(pxg_ctx *ctx, int step);
int build_step_gather(pxg_ctx *ctx);
// node and BC records of the step kernel, after any change of the tiling,
// the shared runs or the BC program
int build_step_records(pxg_ctx *ctx);
// the fused adjoint reverse step on the step tiling (pxg_step.hip): its
// tables, and one launch reading lam_k from lam_in and writing lam_{k-1}
// (bpart: per-block sums of S^T lam for dJ/db, adj_step_blocks() blocks, or null)
int adj_step_setup(pxg_ctx *ctx);
int adj_step_blocks(pxg_ctx *ctx);
int launch_adj_step(pxg_ctx *ctx, int k, const double *u, const double *lam_in, double *lam_out, double *lamp,
                    double *Ebar, double *bpart, double cf);
int launch_preset(pxg_ctx *ctx, double *u, double t, int step);
// device pointers to the bond health / break steps in CSR order (a gathered
// copy when the stream layout is tiled); either output may be NULL
int bond_state_csr(pxg_ctx *ctx, const unsigned char **health, const int **break_step);
int launch_frac_ops(pxg_ctx *ctx, const double *nodeE_dev);
int launch_transpose_ops(pxg_ctx *ctx);
// Cahn-Hilliard bond geometry (lambda-corrected Laplacian, reverse bonds,
// ghost flags) into ctx->lap / rev / isghost; sets ch_ready
int ch_geometry(pxg_ctx *ctx);
// the Cahn-Hilliard assembly on the block-element kernels (pxg_stressch.hip)
int launch_ch_rows(pxg_ctx *ctx, double kappa, double chi, double M0, double dt, const double *u_dev,
                   const double *uold_dev, double *kvals_dev, double *rhs_dev);
// the dynamic fracture assembly on the block-element kernels (pxg_stressch.hip)
int launch_dyn_rows(pxg_ctx *ctx, double dt, const double *u_dev, const double *uold_dev, double *kvals_dev,
                    double *rhs_dev);

// What a rebuild invalidates.  INV_OPS: the point cloud, the families or the
// operators changed, so every element's bond data (explicit, dynamic
// fracture, both stress-coupled Cahn-Hilliard elements) and the cached
// Laplacian are stale and each element must be set up again.  INV_CRACKS: the
// crack seed changed (elements that start from it must be set up again).
enum { INV_OPS = 1, INV_CRACKS = 2 };
void invalidate(pxg_ctx *ctx, int what);

// takes this device's address of the allocation-free entry point's scratch
// array (pxg_assemble.hip); called by pxg_create with the device current
int scratch_init(int device);

// direct solve of the implicit system (pxg_solve.hip): block edge of the
// block-skyline LU
constexpr int kSolveBlock = 64;
struct SolveState {
    int ready = 0;                  // symbolic phase done for pattern `gen`
    int factored = 0;               // numeric factors valid
    long long gen = -1;
    int use_rcm = 1;                // DefaultSolver "use_rcm"
    double max_gib = 16.0;          // DefaultSolver "max_profile_gb" (<= 0: no guard)
    int n = 0, nb = 0, amax = 0;    // equations, blocks, longest active set
    long long ns = 0, nnz = 0;      // strip blocks per triangle, CSR entries
    std::vector<int> hperm;         // new -> old (reference RCM)
    std::vector<long long> hjp;     // reference skyline pointer (n + 1)
    std::vector<int> hap, hfb;      // active-set offsets (nb + 1), first envelope block
    std::vector<int> hal;           // active sets (host copy)
    dbuf<double> F, Li, Ui;         // block storage, inverted diagonal factors
    dbuf<double> y, z, w;           // substitution vectors (nb * kSolveBlock)
    dbuf<int> inv, fb, al, zpiv;
    dbuf<long long> loc, lp;
    dbuf<double> part, norms;       // Newton norms
    int own = 0;                    // factor own_rp/own_cols instead of the context's CSR
    long long own_gen = 0;
    std::vector<int> own_rp, own_cols;
    int use_graphs = 1;             // replay the per-step launch chains as HIP graphs
    hipGraphExec_t gfac = nullptr;  // factor steps of the current pattern
    hipGraphExec_t gsub = nullptr;  // forward and backward sweeps
    void drop_graphs()
    {
        if (gfac) (void)hipGraphExecDestroy(gfac);
        if (gsub) (void)hipGraphExecDestroy(gsub);
        gfac = gsub = nullptr;
    }
};

// which Laplacian ctx->lap holds
enum { LAP_NONE = 0, LAP_POISSON = 1, LAP_CH = 2 };

} // namespace pxg

struct pxg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int fault = 0;                  // host-test fault injection (pxg_set_fault)
    // ---- point cloud ----
    int n = 0, dim = 2, nbulk = 0;
    double dxyz[3] = {0, 0, 0};
    pxg::dbuf<double> coords;       // 3N interleaved (reference layout)
    pxg::dbuf<double> xp;           // dim-packed coordinates (gathers)
    pxg::dbuf<double> vols;
    pxg::dbuf<int> mirror;          // 1-based, 0 = bulk
    std::vector<double> hcoords;    // host copies for BC box filters
    std::vector<int> hmirror;
    // ---- families ----
    double delta = 0;
    long long nb = 0;
    pxg::dbuf<int> off;             // N+1
    pxg::dbuf<int> nbr;             // nb, 1-based
    // ---- cracks ----
    pxg::dbuf<double> crack_seg;    // 4 per segment
    pxg::dbuf<double> crack_quad;   // 12 per quad
    int nseg = 0, nquad = 0;
    pxg::dbuf<unsigned char> seed;  // nb health seed (1 intact, 0 cut)
    // ---- operators ----
    int order = 0, nop = 0;
    pxg::dbuf<double> ops;          // nb*nop (bond-major, reference slot order)
    pxg::dbuf<unsigned char> flags; // N
    int ndegenerate = 0;
    // ---- implicit CSR ----
    int csr_ndof = 0;
    int csr_maxblk = 0;             // most node blocks in one block row
    long long nnz = 0;
    long long csr_gen = 0;          // bumped by every CSR pattern (re)build
    pxg::dbuf<int> rowptr, cols;
    pxg::dbuf<int> slot;            // nb: CSR index of (row i*ndof, col j*ndof)
    pxg::dbuf<int> diag;            // N: CSR index of (i*ndof, i*ndof)
    pxg::dbuf<int> mslot;           // N: CSR index of mirror block (-1 if in family/none)
    pxg::dbuf<double> lap;          // nb: pdLaplacian per bond (Poisson) / lambda-corrected (CH)
    int lap_kind = 0;               // LAP_NONE / LAP_POISSON / LAP_CH
    long long lap_gen = -1;         // ops_gen the Laplacian was built from
    long long ops_gen = 0;          // bumped by every cloud, family or operator rebuild
    pxg::dbuf<int> rev;             // nb
    pxg::dbuf<unsigned char> isghost;
    int ch_ready = 0;
    pxg::dbuf<double> mob;          // n: host-frozen Cahn-Hilliard mobility (pxg_ch_set_mobility)
    int mob_set = 0;
    // element state handed in by the host on every call (pxg_assemble_*_state)
    pxg::dbuf<int> xrev;
    pxg::dbuf<double> xlap, xmob, xvel;
    pxg::dbuf<unsigned char> xghost, xhealth;
    pxg::CsrBcProgram csrbc;        // BC rows of the implicit system
    pxg::SolveState solve;          // direct solve of the implicit system
    pxg::DynFracState dyn;          // implicit dynamic fracture element
    pxg::StressChState sch;         // stress-coupled Cahn-Hilliard element
    pxg::FracSchState fsch;         // fracture stress-coupled Cahn-Hilliard element
    // scratch for host-pointer entry points
    pxg::dbuf<double> su, suo, sk, sr, sf;
    // ---- explicit fracture ----
    pxg::FracState frac;
};

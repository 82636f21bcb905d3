/*
 * pxg.h -- C ABI of the MI355X PDDO hot-path library (libpxg.so).
 *
 * Plain C: pointers, sizes and int status codes only.  It is the device seam
 * that the reference's host code reaches through perix_cuda::* in
 * include/PDSystem/CudaAssemble.h:22-259 (POD structs so the device compiler
 * never parses framework headers), widened to the pieces the reference keeps
 * on the CPU: family construction, the operator build and the whole
 * central-difference loop.
 *
 * Conventions (same as the reference):
 *   - node ids inside neighbour lists are 1-based; CSR indices are 0-based;
 *   - coordinates are interleaved x,y,z per node (z = 0 in 2D);
 *   - solution vectors are node-major, index i*ndof + a;
 *   - every host pointer is borrowed for the duration of the call.
 * Errors: every entry point returns 0 on success or a negative PXG_ERR_*
 * code; pxg_last_error() gives the message.  As in the reference
 * (CudaAssembleKernels.cu:534-540), a failing call leaves the caller free to
 * take its CPU path.  A context is used by one host thread at a time.
 */
#ifndef PXG_H
#define PXG_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PXG_OK 0
#define PXG_ERR_ARG (-1)      /* invalid argument / size */
#define PXG_ERR_HIP (-2)      /* HIP runtime failure */
#define PXG_ERR_STATE (-3)    /* call order (e.g. operators before families) */
#define PXG_ERR_FAULT (-4)    /* injected fault (pxg_set_fault) */
#define PXG_ERR_DEGENERATE (-5) /* degenerate moment matrix */
#define PXG_ERR_BLOWUP (-6)   /* explicit blow-up guard tripped */
#define PXG_ERR_SINGULAR (-7) /* zero pivot in the direct solve */

#define PXG_FP64 0            /* plain FP64 (bit-exact with the reference) */
#define PXG_FP32C 1           /* FP32 operators; FP64 bond products and family sums */

typedef struct pxg_ctx pxg_ctx;

/* ---- lifetime ------------------------------------------------------------
 * Replaces the static device contexts g_ctx/g_epfctx/g_chctx and
 * perix_cuda::shutdown (CudaAssemble.h:257; CudaAssembleKernels.cu:505-560,
 * 1068-1107) with an explicit handle. */
int pxg_create(pxg_ctx **out, int device);
void pxg_destroy(pxg_ctx *ctx);
const char *pxg_last_error(void);
/* Host-test fault injection: the next entry point that would launch work
 * returns PXG_ERR_FAULT before any launch and consumes the fault, so the
 * context's device state is unchanged and later calls work (code 0 clears a
 * pending fault).  Mirrors the `return false` seam failure of
 * CudaAssembleKernels.cu:1112,1193; the latch that keeps a failed explicit
 * run on the CPU (ExplicitPDDOFracElement.cpp:225-234) is the caller's. */
int pxg_set_fault(pxg_ctx *ctx, int code);
/* Blocks until all queued work of the context has finished. */
int pxg_sync(pxg_ctx *ctx);

/* ---- a1: structured point cloud (host) ----------------------------------
 * Restates the quad4/hex8 background generator (Lagrange2DQuad4MeshGenerator
 * .cpp:40-153, Lagrange3DHex8MeshGenerator.cpp:74-130) and the structured
 * branch of PDMesh::createPDMesh (CreatePDMesh.cpp:451-799): cell-centre
 * points in element order, then one ghost per boundary face in group order
 * left,right,bottom,top[,back,front].  wall[i] = -1 for bulk, else the wall
 * index in that order (the "<wall>nodes_ghost" groups). */
typedef struct {
    int dim, nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
} pxg_grid;
int pxg_structured_sizes(const pxg_grid *g, int *n, int *nbulk);
int pxg_structured_cloud(const pxg_grid *g, double *coords, double *vols, int *mirror,
                         int *wall, double *dxyz);
/* HorizonRadius = factor * |xmax-xmin|/nx * (1+1e-6) (ReadPDMeshBlock.cpp:17-23,
 * PDMesh.h:37-41) */
double pxg_horizon_radius(const pxg_grid *g, double factor);
/* a21: RandomIC::apply (RandomIC.cpp:13-42), std::mt19937_64 top 53 bits;
 * dofs are 1-based. */
int pxg_random_ic(uint64_t seed, double lo, double hi, int nnodes, const int *nodes,
                  int ndof, int ndofs_set, const int *dofs, double *values);
/* Step schedule of TimeStepping::solveExplicitDynamics (TimeStepping.cpp:
 * 497-510); returns the step count, fills dts/ts (tNew) up to cap entries. */
int pxg_schedule(double dt, double total, int cap, double *dts, double *ts);

/* ---- a1: point cloud upload ----------------------------------------------
 * Uploads PDMeshData::NodeCoords / NodeVolumes / GhostMirrorBulkID
 * (PDMeshData.h; built by PDMesh::createPDMesh, CreatePDMesh.cpp:451-799)
 * and DX/DY/DZ.  coords[3n], vols[n], mirror[n] (1-based, 0 = bulk). */
int pxg_set_points(pxg_ctx *ctx, int n, int dim, const double *coords,
                   const double *vols, const int *mirror, const double *dxyz);

/* ---- a2: horizon families on the device ---------------------------------
 * Replaces PDMesh::createNeighborNodes (CreatePDMesh.cpp:910-1130).  The
 * lists come out in the reference's insertion order.  r is HorizonRadius
 * (PDMesh.h:37-41, margin included).  nb_out receives the directed bond count. */
int pxg_build_families(pxg_ctx *ctx, double r, long long *nb_out);
/* Copies the flat families (ExplicitPDDOFracElement.cpp:115-153 layout):
 * off[n+1], nbr[nb] (1-based). */
int pxg_get_families(pxg_ctx *ctx, int *off, int *nbr);

/* The host's own families instead of pxg_build_families: off[n+1] (off[0] =
 * 0) and nbr[off[n]] with 1-based ids, in the host's order (the flat layout
 * FormResidualAndJacobianCUDA.cpp:187-205 and ExplicitPDDOFracElement.cpp:
 * 115-153 marshal into AssembleMesh::neigh_off / neigh_id, CudaAssemble.h:
 * 66-67).  Everything built on the old families is dropped. */
int pxg_set_families(pxg_ctx *ctx, const int *off, const int *nbr);

/* ---- a3: force-only crack seeds ------------------------------------------
 * Replaces the force-only health seeding of ExplicitPDDOFracElement.cpp:
 * 134-155 through initialCrackCutsBond (PDMesh.cpp:104-159): segments
 * {x1,y1,x2,y2} (through-thickness in 3D) and planar quads (12 doubles). */
int pxg_set_cracks(pxg_ctx *ctx, int nseg, const double *seg, int nquad,
                   const double *quad);
/* health seed per directed bond (1 intact, 0 cut), family order */
int pxg_get_seed(pxg_ctx *ctx, unsigned char *seed);

/* ---- a4: CSR pattern -----------------------------------------------------
 * Replaces SparseMatrix::initFromPDMesh + setup (SparseMatrix.cpp:73-185).
 * rowptr[n*ndof+1], cols[nnz] (0-based, sorted); 1 <= ndof <= 8 (the
 * elements use 1, dim, 2 and dim+2). */
int pxg_build_csr(pxg_ctx *ctx, int ndof, long long *nnz_out);
int pxg_get_csr(pxg_ctx *ctx, int *rowptr, int *cols);
/* The host's own CSR pattern (SparseMatrix::getCSRRowsIndexPtr /
 * getCSRColsIndexPtr, passed as AssembleMesh::csr_off / csr_col,
 * CudaAssemble.h:64-65) instead of pxg_build_csr: rowptr[n*ndof+1],
 * cols[rowptr[n*ndof]], 0-based, sorted whole ndof x ndof blocks, the rows of
 * one node sharing one pattern that holds the node's diagonal block, every
 * family member's block and at most one other block (the ghost's mirror,
 * SparseMatrix.cpp:90-115).  PXG_ERR_ARG names the first violation.
 * Needs the families. */
int pxg_set_csr(pxg_ctx *ctx, int ndof, const int *rowptr, const int *cols);

/* ---- a5-a8: PDDO moment matrix and g-weights (FP64) ---------------------
 * Replaces PDOperators::calcAMatrixCompute/calcOperatorsCompute
 * (CalcPDOperators.cpp:108-354) and the device twin opsBuildKernel
 * (CudaAssembleKernels.cu:208-243).  dx_char = max(DX,DY,DZ) or 0 to disable
 * the volume correction.  ndeg_out receives the degenerate-node count
 * (usable bonds < nop, or inconsistent moment system). */
int pxg_build_operators(pxg_ctx *ctx, int order, double delta, double dx_char,
                        int *ndeg_out);
/* ops[nb*nop] in the reference slot order (PDOperators.cpp:51-75), flags[n]
 * (bit0 degenerate, bit1 usable<nop); either pointer may be NULL. */
int pxg_get_operators(pxg_ctx *ctx, double *ops, unsigned char *flags);

/* ---- a10-a11: Poisson CSR assembly --------------------------------------
 * Replaces perix_cuda::assemble for ELMT_POISSON (CudaAssemble.h:93-95) and
 * the serial/OpenMP assemblers (FormResidualAndJacobian*.cpp) with
 * PoissonElement (PoissonElement.cpp:26-58).  K = -dF/dU, rhs = +F.
 * f_node (optional, n) generalises the reference's scalar source f_const.
 * Host pointers: u[n], kvals[nnz], rhs[n]. */
int pxg_assemble_poisson(pxg_ctx *ctx, double sigma, const double *f_node,
                         double f_const, const double *u, double *kvals,
                         double *rhs);
/* Same on device pointers (inputs already resident in HBM). */
int pxg_assemble_poisson_dev(pxg_ctx *ctx, double sigma, const double *f_node_dev,
                             double f_const, const double *u_dev,
                             double *kvals_dev, double *rhs_dev);

/* Allocation-free variant of pxg_assemble_poisson_dev.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same bits as pxg_assemble_poisson_dev;
 * it calls no hipMalloc, hipFree or allocating routine and keeps its one
 * scratch buffer (the per-bond Laplacian) in a 16 MiB device array of the
 * library, so the context must have at most 2,097,152 directed bonds
 * (PXG_ERR_ARG otherwise).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams or two host threads) race. */
int pxg_assemble_poisson_noalloc(pxg_ctx *ctx, double sigma, const double *f_node_dev,
                                 double f_const, const double *u_dev, double *kvals_dev,
                                 double *rhs_dev);

/* ---- remaining elements: diffusion ---------------------------------------
 * DiffusionElement (DiffusionElement.cpp:28-62) through the same assembler:
 * backward Euler, K = -dF/dU with bond coeff = -D lap_ij and nodal
 * R = (u - uold)/dt - f, K = 1/dt.  f_node (optional, n) generalises the
 * scalar f_const.  Host pointers: u, uold [n], kvals[nnz], rhs[n]; _dev:
 * device pointers. */
int pxg_assemble_diffusion(pxg_ctx *ctx, double D, const double *f_node, double f_const,
                           double dt, const double *u, const double *uold, double *kvals,
                           double *rhs);
int pxg_assemble_diffusion_dev(pxg_ctx *ctx, double D, const double *f_node_dev,
                               double f_const, double dt, const double *u_dev,
                               const double *uold_dev, double *kvals_dev, double *rhs_dev);

/* ---- a12-a13: Cahn-Hilliard ---------------------------------------------
 * Replaces perix_cuda::assembleCahnHilliard (CudaAssemble.h:139-142) and the
 * host-built geometry of CahnHilliardElement::preprocessIteration
 * (CahnHilliardElement.cpp:93-153), which here is built on the device.
 * Requires pxg_build_csr(ctx, 2, ...) and the operators. */
int pxg_ch_setup(pxg_ctx *ctx);
/* lambda-corrected g2 per bond, reverse-bond index, ghost flags (optional) */
int pxg_ch_get_geometry(pxg_ctx *ctx, double *lapbond, int *rev,
                        unsigned char *isghost);
/* The host element's own geometry instead of pxg_ch_setup: the arrays of
 * perix_cuda::CahnHilliardMesh (lapbond, rev, isghost; CudaAssemble.h:
 * 116-125), built by CahnHilliardElement::preprocessIteration
 * (CahnHilliardElement.cpp:93-153) in the family order of the context.
 * isghost may be NULL (no ghosts). */
int pxg_ch_set_geometry(pxg_ctx *ctx, const double *lapbond, const int *rev,
                        const unsigned char *isghost);
/* The host's Picard-frozen mobility M_i (CahnHilliardMesh::mnode,
 * CudaAssemble.h:124; CahnHilliardElement.cpp:155-175): the following
 * Cahn-Hilliard assemblies use it in place of M0 c (1 - c) from u (their M0
 * argument is then unused).  NULL returns to M0 c (1 - c). */
int pxg_ch_set_mobility(pxg_ctx *ctx, const double *mnode);
/* Picard-frozen mobility M_i = M0 c_i(1-c_i) (CahnHilliardElement.cpp:
 * 155-175) then assembly (:178-292).  u, uold: 2n (c, mu); kvals[nnz];
 * rhs[2n]. */
int pxg_assemble_ch(pxg_ctx *ctx, double kappa, double chi, double M0,
                    double dt, const double *u, const double *uold,
                    double *kvals, double *rhs);
int pxg_assemble_ch_dev(pxg_ctx *ctx, double kappa, double chi, double M0,
                        double dt, const double *u_dev, const double *uold_dev,
                        double *kvals_dev, double *rhs_dev);

/* ---- a15-a20: explicit PDDO fracture --------------------------------------
 * Material / damage parameters (ExplicitPDDOFracParams, CudaAssemble.h:
 * 225-235).  sc is the critical stretch itself (FractureCriterion.h:30-43
 * evaluated by the caller); sc <= 0 or damage_on = 0 disables bond failure. */
typedef struct {
    int dim;
    double C11, C12, C66, rho;
    double bx, by, bz;
    double sc;
    int tension_only, damage_on;
} pxg_frac_params;

/* Builds the per-bond Navier operators from the g-weights (the one-off
 * op cache of ExplicitPDDOFracElement.cpp:106-172) and seeds health from
 * the cracks.  precision = PXG_FP64 or PXG_FP32C.  node_E (optional, n, host)
 * gives a per-node Young's modulus with Poisson ratio nu (then node i's
 * family uses C(E_i); C11/C12/C66 of params are ignored). */
int pxg_frac_setup(pxg_ctx *ctx, const pxg_frac_params *p, int precision,
                   const double *node_E, double nu, int plane_stress);

/* Diagnostic options of the explicit set-up (no reference counterpart; they
 * serve the FP32C accuracy study).  The bits take effect at the next
 * pxg_frac_setup of the context:
 *   PXG_FRAC_NO_SHARED_RUNS  every node reads its own FP32 operators (the 3D
 *                            FP32C step otherwise lets nodes whose FP64
 *                            operators agree to 1e-12 share one run);
 *   PXG_FRAC_FP32_OPERATORS  an FP64 run on operators rounded to FP32 and
 *                            back: the operator rounding of FP32C without
 *                            its FP32 family sums. */
#define PXG_FRAC_NO_SHARED_RUNS 1
#define PXG_FRAC_FP32_OPERATORS 2
int pxg_frac_options(pxg_ctx *ctx, int flags);

/* BC presets run after every central-difference update
 * (BCSystem::presetSolution, BCSystem.cpp:114-129 -> DirichletBC.cpp:107-147,
 * NeumannBC.cpp:88-105), in list order, last writer wins. */
typedef struct {
    int type;          /* 0 neumann (copy mirror bulk), 1 dirichlet */
    int direct;        /* dirichlet: 1 writes g(t), 0 reflects 2g - u_bulk */
    int ncomp;         /* number of constrained components */
    int comp[3];       /* 0-based components */
    double value[3];   /* g(0) per component */
    double velocity[3];/* dg/dt per component */
    int has_box;       /* inclusive coordinate box filter (DirichletBC.cpp:89-98) */
    double box[6];     /* xlo,xhi,ylo,yhi,zlo,zhi; lo > hi leaves the axis open */
    int nnodes;
    const int *nodes;  /* 1-based node ids */
} pxg_bc;
int pxg_frac_set_bcs(pxg_ctx *ctx, int nbc, const pxg_bc *bcs);

/* ---- BC row elimination of the implicit CSR system -----------------------
 * Replaces BCSystem::applyBCs (BCSystem.cpp:212-233) -> DirichletBC::
 * applyAtTime (DirichletBC.cpp:176-230), NeumannBC::apply (NeumannBC.cpp:
 * 127-151) with bcRowScale (BCBase.h:33-39), which the reference runs on the
 * host after every assembly.  pxg_csr_set_bcs resolves the list (same
 * pxg_bc entries as the explicit presets; components index the CSR dofs;
 * the box applies to Dirichlet entries only, as in the reference) against the
 * current CSR pattern; call it again after pxg_build_csr.  Errors
 * (PXG_ERR_ARG): a Neumann node without a mirror bulk node, a constrained row
 * without a stored diagonal, a mirror column outside the pattern.
 * pxg_csr_apply_bcs then rewrites the constrained rows of kvals and rhs at
 * time t with the state u (n*ndof): host arrays, or device arrays with _dev. */
int pxg_csr_set_bcs(pxg_ctx *ctx, int nbc, const pxg_bc *bcs);
int pxg_csr_bc_rows(pxg_ctx *ctx, int *count);
int pxg_csr_apply_bcs(pxg_ctx *ctx, double t, const double *u, double *kvals, double *rhs);
int pxg_csr_apply_bcs_dev(pxg_ctx *ctx, double t, const double *u_dev, double *kvals_dev,
                          double *rhs_dev);

/* ---- direct solve of the implicit system on the device ---------------------
 * Replaces LinearSolver::solve (LinearSolver.cpp:97-121) with its default
 * back-end, DefaultSolver::solveLinearSystem (DefaultSolver.cpp:464-485):
 * the same reverse Cuthill-McKee ordering (:52-150) and skyline envelope
 * (:164-194), factored A = L U without pivoting as 64 x 64 dense blocks of
 * that envelope, then substitution (:372-424).  The symbolic phase runs on
 * the host once per CSR pattern (again after pxg_build_csr / pxg_set_csr).
 * pxg_solve_options mirrors the solver's "use_rcm" and "max_profile_gb"
 * parameters (defaults 1 and 16; a profile above the limit is PXG_ERR_ARG, as
 * the reference refuses it, :196-238; <= 0 disables the guard); flags
 * PXG_SOLVE_NO_GRAPHS launches the per-block kernel chains one by one instead
 * of replaying them as HIP graphs captured once per pattern.
 * pxg_solve_factor_dev factors the device CSR values kvals (the context's
 * pattern, BC rows already applied); an exactly zero pivot is
 * PXG_ERR_SINGULAR naming the equation (:362-367).  pxg_solve_apply_dev
 * solves with the current factors; pxg_solve_dev does both; pxg_solve takes
 * host arrays.  Every _dev call is asynchronous on the context stream except
 * the factorisation's pivot check (one host wait).
 * pxg_solve_info: out[8] = {equations, blocks, block edge, reference profile
 * size, off-diagonal blocks per triangle, largest active set, bytes of block
 * storage, block products per factorisation}.  pxg_solve_ordering copies the
 * ordering (perm[n]: new -> old) and the reference skyline pointer
 * (jp[n+1]); either may be NULL. */
/* pxg_solve_set_pattern gives the solver a square CSR pattern of its own
 * (n rows, sorted distinct columns, every diagonal stored): the reference's
 * SparseMatrix CSR as DefaultSolver::initSolver takes it (:428-462), for
 * systems assembled elsewhere.  Every later pxg_solve* call factors that
 * pattern (kvals index it) until pxg_build_csr or pxg_set_csr gives the
 * context a new implicit CSR. */
int pxg_solve_set_pattern(pxg_ctx *ctx, int n, const int *rowptr, const int *cols);
#define PXG_SOLVE_NO_GRAPHS 1
int pxg_solve_options(pxg_ctx *ctx, int use_rcm, double max_profile_gib, int flags);
int pxg_solve_factor_dev(pxg_ctx *ctx, const double *kvals_dev);
int pxg_solve_apply_dev(pxg_ctx *ctx, const double *rhs_dev, double *x_dev);
int pxg_solve_dev(pxg_ctx *ctx, const double *kvals_dev, const double *rhs_dev, double *x_dev);
int pxg_solve(pxg_ctx *ctx, const double *kvals, const double *rhs, double *x);
int pxg_solve_info(pxg_ctx *ctx, long long *out);
int pxg_solve_ordering(pxg_ctx *ctx, int *perm, long long *jp);

/* Newton iterate on the device (NonlinearSolver.cpp:90-125 and
 * TimeStepping.cpp:167).  pxg_csr_preset_dev applies the BC presets of the
 * resolved pxg_csr_set_bcs list to u (n*ndof) at time t
 * (BCSystem::presetSolution, BCSystem.cpp:114-129).  pxg_newton_update_dev
 * computes norms[0] = |rhs| and norms[1] = |du| and, when |du| is finite,
 * applies u += du and the presets at t (the reference refuses a non-finite
 * update and leaves u unchanged, NonlinearSolver.cpp:100-110); it waits for
 * the stream to read the two norms.  Both refuse (PXG_ERR_STATE) a BC list in
 * which a preset reads a dof that another entry presets. */
int pxg_csr_preset_dev(pxg_ctx *ctx, double t, double *u_dev);
int pxg_newton_update_dev(pxg_ctx *ctx, double t, double *u_dev, const double *du_dev,
                          const double *rhs_dev, double *norms);

/* ---- remaining elements: implicit dynamic fracture ------------------------
 * PDDODynamicFracElement (pddo_dynamic_frac): backward-Euler elastodynamics
 * on the strong-form PDDO operators with critical-stretch bond failure.
 * Requires order-2 operators and pxg_build_csr(ctx, dim, ...); p gives the
 * elastic constants, rho, the body force and the critical stretch
 * (FractureCriterion.h:30-43 evaluated by the caller), v0 the initial
 * velocity (dim values, may be NULL) and uold0 (n*dim host) the initial
 * displacement history; health starts from the crack seed
 * (PDDODynamicFracElement.cpp:112-147).
 * pxg_dynfrac_commit: start of a new time step on the converged u_old
 * (PDDODynamicFracElement.cpp:148-204): vel = (uold - uprev)/dt_prev,
 * uprev = uold, bond failure from u_old.
 * pxg_assemble_dynfrac: K = -dF/dU and the residual
 * (PDDODynamicFracElement.cpp:207-297) for (u, uold) at step size dt.
 * Host pointers, or device pointers with _dev. */
int pxg_dynfrac_setup(pxg_ctx *ctx, const pxg_frac_params *p, const double *v0,
                      const double *uold0);
int pxg_dynfrac_commit(pxg_ctx *ctx, const double *uold, double dt_prev);
int pxg_dynfrac_commit_dev(pxg_ctx *ctx, const double *uold_dev, double dt_prev);
int pxg_assemble_dynfrac(pxg_ctx *ctx, double dt, const double *u, const double *uold,
                         double *kvals, double *rhs);
int pxg_assemble_dynfrac_dev(pxg_ctx *ctx, double dt, const double *u_dev,
                             const double *uold_dev, double *kvals_dev, double *rhs_dev);
/* bond health (nb bytes, CSR order) and committed velocity (n*dim); either may be NULL */
int pxg_dynfrac_get_state(pxg_ctx *ctx, unsigned char *health, double *vel);
/* The same assembly on the host element's own state, as
 * perix_cuda::assembleCached passes it (CudaAssemble.h:196-219, called from
 * FormResidualAndJacobianCUDA.cpp:237-266): the constants of CachedParams,
 * the committed velocity (n*dim; the caller unpacks the cache's 3 slots per
 * node) and the committed bond health (nb doubles, > 0 intact; NULL: all
 * intact).  Needs order-2 operators and a CSR pattern with dim dofs per
 * node; no pxg_dynfrac_setup.  Host arrays. */
typedef struct pxg_dynfrac_state {
    double C11, C12, C66, rho, bx, by, bz;
    const double *vel;
    const double *health;
} pxg_dynfrac_state;
int pxg_assemble_dynfrac_state(pxg_ctx *ctx, const pxg_dynfrac_state *s, double dt, const double *u,
                               const double *uold, double *kvals, double *rhs);

/* ---- remaining elements: stress-coupled Cahn-Hilliard ---------------------
 * StressCahnHilliardElement (stress_cahnhilliard): species c, chemical
 * potential mu and displacement u per node (dofs c, mu, ux, uy[, uz]);
 * the element parameters of the input deck (ReadElmtSystemBlock.cpp:88-103).  plane_stress selects the 2D
 * constants (ignored in 3D). */
typedef struct pxg_stressch_params {
    double E, nu;       /* Young's modulus, Poisson ratio */
    double D;           /* mobility prefactor: M(c) = D c (1 - c) */
    double Omega;       /* partial molar volume (chemo-mechanical coupling) */
    double cref;        /* stress-free reference concentration */
    double chi;         /* regular-solution interaction parameter */
    double kappa;       /* gradient-energy coefficient */
    int plane_stress;   /* 2D: 1 plane stress, 0 plane strain */
} pxg_stressch_params;
/* Requires order-2 operators and pxg_build_csr(ctx, dim + 2, ...); computes
 * the elastic constants (recalcElasticConstants,
 * StressCahnHilliardElement.cpp:80-128; PXG_ERR_ARG for a Poisson ratio the
 * reference rejects) and the bond geometry of preprocessIteration
 * (StressCahnHilliardElement.cpp:166-250). */
int pxg_stressch_setup(pxg_ctx *ctx, const pxg_stressch_params *p);
/* out[11] = C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref */
int pxg_stressch_constants(pxg_ctx *ctx, double *out);
/* K = -dF/dU and the residual of the element through the assembler
 * (StressCahnHilliardElement.cpp:253-467, FormResidualAndJacobian.cpp:25-132)
 * for (u, uold) at step size dt > 0; u, uold n*(dim+2) interleaved per node.
 * Host pointers, or device pointers with _dev. */
int pxg_assemble_stressch(pxg_ctx *ctx, double dt, const double *u, const double *uold,
                          double *kvals, double *rhs);
int pxg_assemble_stressch_dev(pxg_ctx *ctx, double dt, const double *u_dev,
                              const double *uold_dev, double *kvals_dev, double *rhs_dev);

/* ---- remaining elements: fracture stress-coupled Cahn-Hilliard -------------
 * FracStressCahnHilliardElement (frac_stress_cahnhilliard): the stress-coupled
 * Cahn-Hilliard system plus irreversible bond-stretch failure that gates the
 * mechanical bonds (deck parameters ReadElmtSystemBlock.cpp:105-131). */
typedef struct pxg_fsch_params {
    pxg_stressch_params base;   /* E, nu, D, Omega, cref, chi, kappa, plane_stress */
    double rho;                 /* density; <= 0: quasi-static mechanics */
    double sc;                  /* critical stretch (criticalStretchFromGc,
                                   FractureCriterion.h:30-43, evaluated by the
                                   caller); <= 0: no failure */
    double residual_stiffness;  /* stiffness floor of a broken bond (the
                                   reference's default is 1e-8) */
    int tension_only, damage_on;
} pxg_fsch_params;
/* Requires order-2 operators and pxg_build_csr(ctx, dim + 2, ...); bond health
 * starts from the crack seed, the velocity at zero and u_prev at the
 * displacement dofs of uold0 (n*(dim+2) host; only read when rho > 0)
 * (FracStressCahnHilliardElement.cpp:160-301). */
int pxg_fsch_setup(pxg_ctx *ctx, const pxg_fsch_params *p, const double *uold0);
/* The fracture stress-coupled Cahn-Hilliard assembly on the host element's
 * own state, as perix_cuda::assembleStressCahnHilliard passes it
 * (CudaAssemble.h:150-191, called from FormResidualAndJacobianCUDA.cpp:
 * 268-297): the element constants of StressCahnHilliardParams, the reverse
 * bonds, the species Laplacian that drives both the flux and the kappa term,
 * the ghost flags (NULL: none), the Picard-frozen mobility, the committed
 * bond health (nb doubles, > 0 intact; NULL: all intact; a broken bond's
 * mechanical terms are scaled by kres) and, with rho > 0, the committed
 * velocity (n*dim) of the backward-Euler inertia.  Needs order-2 operators
 * and a CSR pattern with dim + 2 dofs per node.  Host arrays; the state is
 * uploaded on every call. */
typedef struct pxg_sch_state {
    double C11, C12, C66, A, Kh, Ch, Omega, cref, chi, kappa, rho, kres;
    const int *rev;
    const double *lap;
    const unsigned char *isghost;
    const double *mnode;
    const double *health;
    const double *vel;
} pxg_sch_state;
int pxg_assemble_sch_state(pxg_ctx *ctx, const pxg_sch_state *s, double dt, const double *u,
                           const double *uold, double *kvals, double *rhs);

/* out[14] = C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref, rho,
 * residual stiffness, critical stretch */
int pxg_fsch_constants(pxg_ctx *ctx, double *out);
/* start of a new time step on the accepted u_old
 * (FracStressCahnHilliardElement.cpp:303-398): with rho > 0 the velocity
 * (u_old - u_prev)/dt_prev and u_prev = u_old; bond failure on the
 * swelling-corrected stretch of u_old. */
int pxg_fsch_commit(pxg_ctx *ctx, const double *uold, double dt_prev);
int pxg_fsch_commit_dev(pxg_ctx *ctx, const double *uold_dev, double dt_prev);
/* K = -dF/dU and the residual (FracStressCahnHilliardElement.cpp:400-561
 * through FormResidualAndJacobian.cpp:25-132) for (u, uold) at step dt > 0. */
int pxg_assemble_fsch(pxg_ctx *ctx, double dt, const double *u, const double *uold,
                      double *kvals, double *rhs);
int pxg_assemble_fsch_dev(pxg_ctx *ctx, double dt, const double *u_dev,
                          const double *uold_dev, double *kvals_dev, double *rhs_dev);
/* bond health (nb bytes, CSR order) and committed velocity (n*dim); either may be NULL */
int pxg_fsch_get_state(pxg_ctx *ctx, unsigned char *health, double *vel);

/* Output projections of the chemo-mechanical elements
 * (StressCahnHilliardElement::computeNodalProjection,
 * StressCahnHilliardElement.cpp:481-658; FracStressCahnHilliardElement::
 * computeNodalProjection, FracStressCahnHilliardElement.cpp:47-110, 575-672)
 * for the state u (n*(dim+2)).  stress_cahnhilliard lists strain .. gradC,
 * frac_stress_cahnhilliard strain .. sigmaH, gradC, damage, DamageI (its
 * current bond health).  out holds pxg_sch_components(which) values per node
 * (tensors row-major 3x3, 2D zero-padded as the reference pads them).
 * Host pointers, or device pointers with _dev. */
#define PXG_SCH_STRAIN 0
#define PXG_SCH_STRESS 1
#define PXG_SCH_VONMISES 2
#define PXG_SCH_SIGMAH 3
#define PXG_SCH_FPRIME 4
#define PXG_SCH_FDOUBLEPRIME 5
#define PXG_SCH_GRADC 6
#define PXG_SCH_DAMAGE 7
#define PXG_SCH_DAMAGEI 8
int pxg_sch_components(int which);
int pxg_stressch_project(pxg_ctx *ctx, int which, const double *u, double *out);
int pxg_stressch_project_dev(pxg_ctx *ctx, int which, const double *u_dev, double *out_dev);
int pxg_fsch_project(pxg_ctx *ctx, int which, const double *u, double *out);
int pxg_fsch_project_dev(pxg_ctx *ctx, int which, const double *u_dev, double *out_dev);

/* ---- f3: result output ----------------------------------------------------
 * Replaces PDProblem::savePDResults (SaveResults.cpp:53-204): the same VTU
 * content and order (N points and VTK_VERTEX cells; point data ElmtID = the
 * parent cell id for the first nbulk nodes and -1 for ghosts, one Float64
 * scalar per dof named dof_names[a] from u[n*ndof], then nfields projections
 * field_data[f] of field_comps[f] = 1, 3 or 9 values per node; the
 * coordinates coords[3n]), written as raw little-endian binary appended data
 * (UInt64 headers) instead of 16-digit ASCII.  Host arrays only; needs no
 * context or GPU.  The reference names the file <deck>-PD-results-<step:05>.vtu. */
int pxg_write_vtu(const char *path, int n, int nbulk, const double *coords, int ndof,
                  const char *const *dof_names, const double *u, int nfields,
                  const char *const *field_names, const int *field_comps,
                  const double *const *field_data);

/* Per-step drop-in of perix_cuda::assembleExplicitPDDOFracAccel
 * (CudaAssemble.h:251-254; CPU loop ExplicitPDDOFracElement.cpp:237-299):
 * damage with u, then health-gated Navier force, accel = (L+b)/rho.
 * u[n*dim] host; health_inout[nb] as double 0/1 (uploaded on the first call,
 * kept on the device afterwards, written back); accel_out[n*dim]. */
int pxg_frac_accel(pxg_ctx *ctx, const double *u, double *health_inout,
                   double *accel_out, int *any_broke);

/* Device-resident central-difference loop (TimeStepping::solveExplicitDynamics,
 * TimeStepping.cpp:417-599): the host passes the state once; every step runs
 * damage + force + update + BC preset + blow-up guard on the device.
 * pxg_frac_init: u0 = preset initial state (n*dim), v0 initial velocity (may be
 * NULL); resets bond health to the crack seed, performs the t=0 rate
 * evaluation and u^{-1}.  Steps follow the
 * reference schedule for (dt, total); nsteps_max (>0) stops early after that
 * many steps (for timing).  Returns PXG_ERR_BLOWUP when the guard trips. */
int pxg_frac_init(pxg_ctx *ctx, const double *u0, const double *v0, double dt);
int pxg_frac_run(pxg_ctx *ctx, double dt, double total, int nsteps_max,
                 int *steps_done);
/* Nodal output projections of the current state (ExplicitPDDOFracElement::
 * computeNodalProjection, ExplicitPDDOFracElement.cpp:314-405, called by the
 * reference's output writer): which = 0 damage, 1 von Mises (n values),
 * 2 stress, 3 strain (9 per node, row-major 3x3).  out: host (n or 9n
 * doubles); _dev: a device pointer of that size. */
int pxg_frac_project(pxg_ctx *ctx, int which, double *out);
int pxg_frac_project_dev(pxg_ctx *ctx, int which, double *out_dev);
/* Stream layout of the step kernel (no reference counterpart; reporting):
 * info[0] bond slots per operator component, info[1] 1 when the per-bond
 * streams follow the tile order of the step blocks, info[2] shared FP32
 * operator runs, info[3] nodes reading a shared run (FP32C). */
int pxg_frac_layout(pxg_ctx *ctx, long long *info);
/* Downloads u^n, u^{n-1} (n*dim), health (nb, 0/1 bytes) and break steps
 * (nb ints: -1 seeded, 0 broken at t=0, k during step k, INT32_MAX intact)
 * in CSR bond order; any pointer may be NULL. */
int pxg_frac_get_state(pxg_ctx *ctx, double *u, double *uold,
                       unsigned char *health, int *break_step);
/* Device pointer of the displacement u^n and u^{n-1} (valid until the next
 * pxg_frac_* call). */
int pxg_frac_state_ptrs(pxg_ctx *ctx, double **u_dev, double **uold_dev);

/* Timing of the step kernel inside pxg_frac_run (HIP events on the
 * context's stream, one pair around each run of back-to-back launches
 * between guard polls): enable, run, then read the summed time and the
 * number of step launches it covers. */
int pxg_frac_profile(pxg_ctx *ctx, int enable);
int pxg_frac_kernel_time(pxg_ctx *ctx, double *ms_total, int *launches);

/* ---- a22: adjoint of the explicit step (new; no reference counterpart) ----
 * Gradient of J = sum_s u^K[sens_s*dim+comp]^2 with respect to the per-node
 * Young's modulus given to pxg_frac_setup, with bond health frozen from the
 * recorded forward break steps, uniform steps t_k = t_{k-1} + dt, and
 * displacement checkpoints every `ckpt` steps (recomputed in segments).
 * u0 (n*dim host, preset initial state), v0 (optional).  grad[n] host. */
int pxg_frac_adjoint(pxg_ctx *ctx, const double *u0, const double *v0,
                     double dt, int nsteps, int ckpt, int nsens,
                     const int *sens, int comp, double *J_out, double *grad);
/* The same sweep with the gradient with respect to the uniform body force b
 * of pxg_frac_params as well: grad_b[dim] = dJ/db (b enters every
 * central-difference value and u^{-1} through a = (L + b)/rho).  grad (n)
 * and grad_b may each be NULL. */
int pxg_frac_adjoint_ex(pxg_ctx *ctx, const double *u0, const double *v0,
                        double dt, int nsteps, int ckpt, int nsens,
                        const int *sens, int comp, double *J_out, double *grad,
                        double *grad_b);
/* Replaces the body force b (dim values) of pxg_frac_setup. */
int pxg_frac_set_body(pxg_ctx *ctx, const double *b);
/* Frozen-health forward loss of the same map (finite-difference checks);
 * node_E (optional, host, n) replaces the per-node modulus for this call. */
int pxg_frac_loss(pxg_ctx *ctx, const double *u0, const double *v0, double dt,
                  int nsteps, const double *node_E, int nsens, const int *sens,
                  int comp, double *J_out);

#ifdef __cplusplus
}
#endif
#endif

// perix_cuda_pxg.cpp -- the reference's device seam (namespace perix_cuda,
// include/PDSystem/CudaAssemble.h, compiled here unchanged) implemented over
// libpxg.so.  A PeriX build links this file in place of
// CudaAssembleKernels.cu; its call sites stay as they are:
//   PDSystem::formResidualAndJacobianCUDA   FormResidualAndJacobianCUDA.cpp:34-309
//   ExplicitPDDOFracElement::preprocessIteration   ExplicitPDDOFracElement.cpp:174-235
//
// Contract (CudaAssemble.h:87-95 and the callers): every function returns
// true when the device produced the result and false when it could not; the
// caller then takes its CPU path (the explicit element stays there for the
// rest of the run and calls shutdown()).  Inputs are host arrays borrowed for
// the call; outputs are host arrays the call fills.  One host thread, no
// re-entrancy (one context per seam entry, like the reference's static
// device contexts, rebuilt when the mesh or the pattern changes).
//
// What crosses the seam and how libpxg takes it:
//   - AssembleMesh: points and volumes (pxg_set_points; the seam carries no
//     ghost mirror map and none is needed: the host's own families and CSR
//     pattern are uploaded with pxg_set_families / pxg_set_csr), the PDDO
//     order (nop = 5 in 2D, 9 in 3D: order 2, the only order libpxg builds)
//     and the volume-correction spacing for pxg_build_operators.  Variable
//     horizons (nodeHorizon) are not supported: false, CPU path.
//   - CahnHilliardMesh: families, volumes and the element's own geometry and
//     frozen mobility (pxg_ch_set_geometry / pxg_ch_set_mobility).  It has no
//     coordinates; the Cahn-Hilliard assembly reads none.
//   - the element state of assembleStressCahnHilliard / assembleCached
//     (committed health, velocity, frozen mobility) goes to
//     pxg_assemble_sch_state / pxg_assemble_dynfrac_state on every call.
//   - the explicit element's health is uploaded on the first call and kept on
//     the device; it is copied back only when a bond broke (pxg_frac_accel).
#include "PDSystem/CudaAssemble.h"
#include "pxg.h"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

long long g_ok = 0, g_failed = 0;   // seam calls served by the device / refused
long long g_calls = 0;
long long g_fault_at = 0;           // > 0: inject a fault into that call (tests)

bool fail(const char *where)
{
    std::fprintf(stderr, "[PeriX pxg] %s: %s\n", where, pxg_last_error());
    ++g_failed;
    return false;
}

bool done(int rc, const char *where)
{
    if (rc != PXG_OK) return fail(where);
    ++g_ok;
    return true;
}

// order of the PDDO slot count (PDOperators.cpp:31-111): 2D 2/5/9, 3D 3/9/19
int order_of(int dim, int nop)
{
    const int tab2[] = {2, 5, 9}, tab3[] = {3, 9, 19};
    for (int o = 0; o < 3; o++)
        if ((dim == 2 ? tab2[o] : tab3[o]) == nop) return o + 1;
    return 0;
}

// One seam entry's device context: rebuilt when the mesh (sizes, arrays,
// horizon) or the CSR pattern changes.
struct Slot {
    pxg_ctx *ctx = nullptr;
    int N = -1, nbtot = -1, dim = -1, nop = -1;
    double delta = -1.0, dx_char = -1.0;
    const void *coords = nullptr, *neigh_off = nullptr, *neigh_id = nullptr;
    int ndof = -1, nnz = -1;
    const int *csr_off = nullptr, *csr_col = nullptr;
    bool frac_ready = false;

    void reset()
    {
        if (ctx) pxg_destroy(ctx);
        *this = Slot();
    }

    bool create()
    {
        if (ctx) return true;
        if (pxg_create(&ctx, 0) != PXG_OK) { ctx = nullptr; return fail("pxg_create"); }
        return true;
    }

    // points, the host's families and the order-2 operators of an AssembleMesh
    bool mesh(const perix_cuda::AssembleMesh &m)
    {
        if (m.nodeHorizon || m.nodeSpacing) {
            std::fprintf(stderr, "[PeriX pxg] variable horizons are not supported on the device\n");
            return fail("mesh");
        }
        const int nbtot = (m.N > 0 && m.neigh_off) ? m.neigh_off[m.N] : -1;
        if (ctx && N == m.N && nbtot == this->nbtot && dim == m.dim && nop == m.nop && delta == m.delta &&
            dx_char == m.dx_char && coords == m.coords && neigh_off == m.neigh_off && neigh_id == m.neigh_id)
            return true;
        reset();
        if (order_of(m.dim, m.nop) != 2) {
            std::fprintf(stderr, "[PeriX pxg] PDDO order with %d slots in %dD: only order 2 is built\n", m.nop, m.dim);
            return fail("mesh");
        }
        if (!create()) return false;
        std::vector<int> mirror((size_t)m.N, 0);
        // any positive cell size serves the step kernel's node tiles; the
        // volume correction itself takes dx_char (0: off)
        const double h = m.dx_char > 0.0 ? m.dx_char : m.delta / 3.0;
        const double dxyz[3] = {h, h, m.dim == 3 ? h : 0.0};
        int ndeg = 0;
        if (pxg_set_points(ctx, m.N, m.dim, m.coords, m.vols, mirror.data(), dxyz) != PXG_OK ||
            pxg_set_families(ctx, m.neigh_off, m.neigh_id) != PXG_OK ||
            pxg_build_operators(ctx, 2, m.delta, m.dx_char, &ndeg) != PXG_OK) {
            fail("mesh set-up");
            reset();
            return false;
        }
        N = m.N; this->nbtot = nbtot; dim = m.dim; nop = m.nop; delta = m.delta; dx_char = m.dx_char;
        coords = m.coords; neigh_off = m.neigh_off; neigh_id = m.neigh_id;
        return true;
    }

    bool csr(int nd, const int *off, const int *col, int nz)
    {
        if (ndof == nd && nnz == nz && csr_off == off && csr_col == col) return true;
        if (pxg_set_csr(ctx, nd, off, col) != PXG_OK) return fail("pxg_set_csr");
        ndof = nd; nnz = nz; csr_off = off; csr_col = col;
        return true;
    }
};

Slot g_assemble, g_ch, g_sch, g_cached, g_explicit;

// test hook: the call number that receives an injected fault
bool inject(Slot &s)
{
    ++g_calls;
    if (g_fault_at > 0 && g_calls == g_fault_at && s.ctx) pxg_set_fault(s.ctx, 1);
    return true;
}

} // namespace

namespace perix_cuda {

bool assemble(const AssembleMesh &m, const ElmtParams &p, const OpSlots &, const double *u, const double *uold,
              double *out_Kvals, double *out_rhs)
{
    Slot &s = g_assemble;
    if (p.ndof != 1 || !s.mesh(m) || !s.csr(1, m.csr_off, m.csr_col, m.nnz)) return false;
    inject(s);
    if (p.type == ELMT_POISSON)
        return done(pxg_assemble_poisson(s.ctx, p.sigma, nullptr, p.f, u, out_Kvals, out_rhs), "pxg_assemble_poisson");
    if (p.type == ELMT_DIFFUSION)
        return done(pxg_assemble_diffusion(s.ctx, p.D, nullptr, p.f, p.dt, u, uold, out_Kvals, out_rhs),
                    "pxg_assemble_diffusion");
    return fail("assemble: element type without a device port");
}

bool assembleCahnHilliard(const CahnHilliardMesh &cm, const CahnHilliardParams &p, const int *csr_off,
                          const int *csr_col, int nnz, const double *u, const double *uold, double *out_Kvals,
                          double *out_rhs)
{
    Slot &s = g_ch;
    const int nbtot = cm.nbtot;
    const bool same = s.ctx && s.N == cm.N && s.nbtot == nbtot && s.neigh_off == cm.off && s.neigh_id == cm.nbr;
    if (!same) {
        s.reset();
        if (!s.create()) return false;
        // the Cahn-Hilliard assembly reads volumes, families, the CSR
        // pattern and the element's geometry; no coordinates cross this seam
        std::vector<double> zeros((size_t)3 * cm.N, 0.0);
        std::vector<int> mirror((size_t)cm.N, 0);
        const double dxyz[3] = {1.0, 1.0, 0.0};
        if (pxg_set_points(s.ctx, cm.N, 2, zeros.data(), cm.vol, mirror.data(), dxyz) != PXG_OK ||
            pxg_set_families(s.ctx, cm.off, cm.nbr) != PXG_OK) {
            fail("Cahn-Hilliard set-up");
            s.reset();
            return false;
        }
        s.N = cm.N; s.nbtot = nbtot; s.neigh_off = cm.off; s.neigh_id = cm.nbr;
    }
    if (!s.csr(2, csr_off, csr_col, nnz)) return false;
    inject(s);
    // the element's geometry (fixed after its first preprocessIteration) and
    // this iteration's frozen mobility
    if (pxg_ch_set_geometry(s.ctx, cm.lapbond, cm.rev, reinterpret_cast<const unsigned char *>(cm.isghost)) !=
            PXG_OK ||
        pxg_ch_set_mobility(s.ctx, cm.mnode) != PXG_OK)
        return fail("Cahn-Hilliard state");
    return done(pxg_assemble_ch(s.ctx, p.kappa, p.chi, 0.0, p.dt, u, uold, out_Kvals, out_rhs), "pxg_assemble_ch");
}

bool assembleStressCahnHilliard(const AssembleMesh &m, const StressCahnHilliardParams &p, const OpSlots &,
                                const int *rev, const double *lapbond, const char *isghost, const double *mnode,
                                const double *health, const double *vel, const double *u, const double *uold,
                                double *out_Kvals, double *out_rhs)
{
    Slot &s = g_sch;
    if (p.ndof != p.dim + 2 || !s.mesh(m) || !s.csr(p.ndof, m.csr_off, m.csr_col, m.nnz)) return false;
    inject(s);
    pxg_sch_state st;
    st.C11 = p.C11; st.C12 = p.C12; st.C66 = p.C66; st.A = p.A; st.Kh = p.Kh; st.Ch = p.Ch;
    st.Omega = p.Omega; st.cref = p.cref; st.chi = p.chi; st.kappa = p.kappa; st.rho = p.rho; st.kres = p.kres;
    st.rev = rev; st.lap = lapbond; st.isghost = reinterpret_cast<const unsigned char *>(isghost);
    st.mnode = mnode; st.health = health; st.vel = vel;
    return done(pxg_assemble_sch_state(s.ctx, &st, p.dt, u, uold, out_Kvals, out_rhs), "pxg_assemble_sch_state");
}

bool assembleCached(const AssembleMesh &m, const CachedParams &p, const OpSlots &, const double *nodecache,
                    const double *bondcache, const double *u, const double *uold, double *out_Kvals,
                    double *out_rhs)
{
    Slot &s = g_cached;
    if (p.ndof != p.dim || p.ncache < p.dim || !s.mesh(m) || !s.csr(p.ndof, m.csr_off, m.csr_col, m.nnz))
        return false;
    inject(s);
    // committed velocity: ncache slots per node in the cache, dim in libpxg
    std::vector<double> vel((size_t)m.N * p.dim);
    for (int i = 0; i < m.N; i++)
        for (int d = 0; d < p.dim; d++) vel[(size_t)i * p.dim + d] = nodecache[(size_t)i * p.ncache + d];
    pxg_dynfrac_state st;
    st.C11 = p.C11; st.C12 = p.C12; st.C66 = p.C66; st.rho = p.rho; st.bx = p.bx; st.by = p.by; st.bz = p.bz;
    st.vel = vel.data(); st.health = bondcache;
    return done(pxg_assemble_dynfrac_state(s.ctx, &st, p.dt, u, uold, out_Kvals, out_rhs),
                "pxg_assemble_dynfrac_state");
}

bool assembleExplicitPDDOFracAccel(const AssembleMesh &m, const ExplicitPDDOFracParams &p, const OpSlots &,
                                   const double *u, double *health_inout, double *accel_out, int *health_changed)
{
    Slot &s = g_explicit;
    if (!s.mesh(m)) return false;      // a rebuilt context starts with frac_ready = false
    if (!s.frac_ready) {
        pxg_frac_params fp;
        fp.dim = p.dim; fp.C11 = p.C11; fp.C12 = p.C12; fp.C66 = p.C66; fp.rho = p.rho;
        fp.bx = p.bx; fp.by = p.by; fp.bz = p.bz; fp.sc = p.sc;
        fp.tension_only = p.tension_only; fp.damage_on = p.damage_on;
        // Poisson ratio from the element's constants (ExplicitPDDOFracElement.cpp:
        // 37-52): C12/C11 = nu (plane stress), nu/(1-nu) (plane strain, 3D); only
        // a per-node modulus would use it
        const double nu = (p.C11 > 0.0) ? (p.plane_stress && p.dim == 2 ? p.C12 / p.C11 : p.C12 / (p.C11 + p.C12))
                                        : 0.0;
        if (pxg_frac_setup(s.ctx, &fp, PXG_FP64, nullptr, nu, p.plane_stress) != PXG_OK)
            return fail("pxg_frac_setup");
        s.frac_ready = true;
    }
    inject(s);
    return done(pxg_frac_accel(s.ctx, u, health_inout, accel_out, health_changed), "pxg_frac_accel");
}

void shutdown()
{
    g_assemble.reset();
    g_ch.reset();
    g_sch.reset();
    g_cached.reset();
    g_explicit.reset();
}

} // namespace perix_cuda

// Counters of the seam (for tests and logs): calls the device served and
// calls it refused; and a test hook that injects a fault (pxg_set_fault) into
// the given call number so that the caller's CPU fallback can be exercised.
extern "C" void perix_cuda_pxg_stats(long long *ok, long long *failed)
{
    if (ok) *ok = g_ok;
    if (failed) *failed = g_failed;
}

extern "C" void perix_cuda_pxg_reset_stats(long long fault_at_call)
{
    g_ok = g_failed = g_calls = 0;
    g_fault_at = fault_at_call;
}

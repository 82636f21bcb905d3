// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver over the reference's own C++ classes, compiled
// together with the reference sources where they lie under the read-only
// reference tree (see oracle/ref/Makefile).  It is used only to produce
// golden vectors and to validate the C oracle (oracle/pxo_*.c).  Nothing in
// the product links it.
//
// The harness reads object state the reference keeps private (PDProblem's
// members, the explicit element's health/op caches, the CH element's bond
// Laplacian) by compiling the reference headers with `private` mapped to
// `public`; standard headers are included first so that only the reference
// classes are affected.  Object layout is unchanged by this.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>
#include <fstream>
#include <sstream>
#include <iostream>
#include <chrono>
#include <filesystem>
#include <nlohmann/json.hpp>
#include <Eigen/Dense>
#define private public
#define protected public
#include "PDProblem/PDProblem.h"
#include "ElmtSystem/ExplicitPDDOFracElement.h"
#include "ElmtSystem/CahnHilliardElement.h"
#include "ElmtSystem/PoissonElement.h"
#include "ElmtSystem/PDDODynamicFracElement.h"
#include "ElmtSystem/StressCahnHilliardElement.h"
#include "ElmtSystem/FracStressCahnHilliardElement.h"
#include "LinearSolver/DefaultSolver.h"
#undef private
#undef protected

namespace {
double &last_bc_seconds()
{
    static double s = 0.0;
    return s;
}
struct RefCase {
    PDProblem prob;
    std::string deck;
};
}

extern "C" {

// Runs PDProblem::init on a JSON deck (mesh, families, operators setup).
void *ref_open(const char *deck)
{
    auto *c = new RefCase();
    c->deck = deck;
    std::string a0 = "perix", a1 = "-i", a2 = deck;
    char *argv[3] = {a0.data(), a1.data(), a2.data()};
    c->prob.init(3, argv);
    return c;
}

void ref_close(void *h) { delete static_cast<RefCase *>(h); }

// out: N, nb, dim, ndof, nnz (0 for explicit), nbulk
void ref_info(void *h, long long *out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const auto &md = p.m_PDMesh.getDataConstRef();
    long long nb = 0;
    for (const auto &f : md.NodesNeighNodesID) nb += (long long)f.size();
    out[0] = md.NodesNum;
    out[1] = nb;
    out[2] = p.m_Mesh.getMeshDataRef().MeshDim;
    out[3] = p.m_ElmtSystem.getMaxDofsPerNode();
    out[4] = p.m_ElmtSystem.isExplicit() ? 0 : p.m_Matrix.getNNZNum();
    out[5] = md.BulkElmtsNum;
}

double ref_horizon(void *h)
{
    return static_cast<RefCase *>(h)->prob.m_PDMesh.getDataConstRef().HorizonRadius;
}

void ref_dxyz(void *h, double *d)
{
    const auto &md = static_cast<RefCase *>(h)->prob.m_PDMesh.getDataConstRef();
    d[0] = md.DX; d[1] = md.DY; d[2] = md.DZ;
}

void ref_mesh(void *h, double *coords, double *vols, int *mirror)
{
    const auto &md = static_cast<RefCase *>(h)->prob.m_PDMesh.getDataConstRef();
    std::copy(md.NodeCoords.begin(), md.NodeCoords.end(), coords);
    std::copy(md.NodeVolumes.begin(), md.NodeVolumes.end(), vols);
    for (int i = 0; i < md.NodesNum; i++)
        mirror[i] = (i < (int)md.GhostMirrorBulkID.size()) ? md.GhostMirrorBulkID[i] : 0;
}

void ref_families(void *h, int *off, int *nbr)
{
    const auto &md = static_cast<RefCase *>(h)->prob.m_PDMesh.getDataConstRef();
    long long p = 0;
    off[0] = 0;
    for (int i = 0; i < md.NodesNum; i++) {
        for (int j : md.NodesNeighNodesID[i]) nbr[p++] = j;
        off[i + 1] = (int)p;
    }
}

// returns the group size; fills ids (1-based) when non-null
int ref_group(void *h, const char *name, int *ids)
{
    const auto &v = static_cast<RefCase *>(h)->prob.m_PDMesh.getNodeIDsViaPhyName(name);
    if (ids) std::copy(v.begin(), v.end(), ids);
    return (int)v.size();
}

// full operator vectors (nop per bond, family order); flags[i]=1 if degenerate
int ref_ops(void *h, double *ops, unsigned char *flags)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const auto &md = p.m_PDMesh.getDataConstRef();
    PDOperators op = p.m_Operators;
    op.setAbortOnDegenerate(false);
    const int nop = op.getOperatorsVecSize();
    long long t = 0;
    for (int i = 1; i <= md.NodesNum; i++) {
        op.calcAMatrix(i, md);
        if (flags) flags[i - 1] = op.m_LastFamilyDegenerate ? 1 : 0;
        for (int j : md.NodesNeighNodesID[i - 1]) {
            op.calcOperators(i, j, md);
            for (int k = 1; k <= nop; k++) ops[t * nop + k - 1] = op.m_OperatorsVec(k);
            t++;
        }
    }
    return nop;
}

// CSR pattern of the implicit matrix
void ref_csr(void *h, int *rowptr, int *cols)
{
    auto &m = static_cast<RefCase *>(h)->prob.m_Matrix;
    const auto r = m.getCSRRowsIndexCopy();
    const auto c = m.getCSRColsIndexCopy();
    std::copy(r.begin(), r.end(), rowptr);
    std::copy(c.begin(), c.end(), cols);
}

// Initial state exactly as PDProblem::run: zero, ICs, BC presets (t=0).
void ref_initial_state(void *h, double *u, double *v0)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    p.m_U.setToZeros();
    p.m_InitialVelocity.setToZeros();
    p.m_ICSystem.applyInitialConditions(p.m_PDMesh, ndof, p.m_U, p.m_InitialVelocity);
    p.m_BCSystem.presetSolution(p.m_PDMesh, ndof, p.m_U);
    p.m_Uold = p.m_U;
    const int n = p.m_U.getSize();
    std::copy(p.m_U.getDataPtr(), p.m_U.getDataPtr() + n, u);
    if (v0) std::copy(p.m_InitialVelocity.getDataPtr(), p.m_InitialVelocity.getDataPtr() + n, v0);
}

// BC preset on a caller vector at time t (BCSystem::presetSolution)
void ref_preset(void *h, double *u, double t)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    VectorXd U(p.m_U.getSize());
    std::copy(u, u + U.getSize(), U.getDataPtr());
    p.m_BCSystem.presetSolution(p.m_PDMesh, ndof, U, t);
    std::copy(U.getDataPtr(), U.getDataPtr() + U.getSize(), u);
}

// One formResidualAndJacobian at (u, uold) with Info.Dt = dt: the serial
// path, or with `parallel` the OpenMP one (FormResidualAndJacobianParallel.cpp).
static void assemble_any(void *h, const double *u, const double *uold, double dt,
                         double *kvals, double *rhs, bool parallel)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int n = p.m_U.getSize();
    VectorXd U(n), Uo(n);
    std::copy(u, u + n, U.getDataPtr());
    std::copy(uold, uold + n, Uo.getDataPtr());
    LocalElmtInfo info;
    info.Dt = dt; info.T = dt; info.Timestep = 1;
    if (parallel)
        p.m_PDSystem.formResidualAndJacobianParallel(p.m_PDMesh, p.m_Operators, p.m_ElmtSystem, info,
                                                     U, Uo, p.m_Matrix, p.m_RHS);
    else
        p.m_PDSystem.formResidualAndJacobian(p.m_PDMesh, p.m_Operators, p.m_ElmtSystem, info,
                                             U, Uo, p.m_Matrix, p.m_RHS);
    const int nnz = p.m_Matrix.getNNZNum();
    std::copy(p.m_Matrix.getCSRValuesPtr(), p.m_Matrix.getCSRValuesPtr() + nnz, kvals);
    std::copy(p.m_RHS.getDataPtr(), p.m_RHS.getDataPtr() + n, rhs);
}

void ref_assemble(void *h, const double *u, const double *uold, double dt, double *kvals, double *rhs)
{
    assemble_any(h, u, uold, dt, kvals, rhs, false);
}

void ref_assemble_parallel(void *h, const double *u, const double *uold, double dt, double *kvals,
                           double *rhs)
{
    assemble_any(h, u, uold, dt, kvals, rhs, true);
}

// BCSystem::applyBCs (BCSystem.cpp:212-233) on caller K values / RHS at
// time t with the state u: the row elimination of the implicit solve.
void ref_apply_bcs(void *h, const double *u, double t, double *kvals, double *rhs)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    const int n = p.m_U.getSize();
    const int nnz = p.m_Matrix.getNNZNum();
    VectorXd U(n);
    std::copy(u, u + n, U.getDataPtr());
    std::copy(kvals, kvals + nnz, p.m_Matrix.getCSRValuesPtr());
    std::copy(rhs, rhs + n, p.m_RHS.getDataPtr());
    LocalElmtInfo info;
    info.T = t;
    const auto t0 = std::chrono::steady_clock::now();
    p.m_BCSystem.applyBCs(p.m_PDMesh, p.m_Operators, ndof, U, U, info, p.m_Matrix, p.m_RHS);
    const auto t1 = std::chrono::steady_clock::now();
    last_bc_seconds() = std::chrono::duration<double>(t1 - t0).count();
    std::copy(p.m_Matrix.getCSRValuesPtr(), p.m_Matrix.getCSRValuesPtr() + nnz, kvals);
    std::copy(p.m_RHS.getDataPtr(), p.m_RHS.getDataPtr() + n, rhs);
}

// wall time of the applyBCs call alone in the last ref_apply_bcs
double ref_last_bc_seconds() { return last_bc_seconds(); }

// One serial formResidualAndJacobian as time step `timestep` (Info.Dt = dt,
// Info.T = timestep*dt): time-split elements commit their per-step state
// when the step number changes (PDDODynamicFracElement::preprocessIteration).
void ref_assemble_step(void *h, const double *u, const double *uold, double dt, int timestep, double *kvals,
                       double *rhs)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int n = p.m_U.getSize();
    VectorXd U(n), Uo(n);
    std::copy(u, u + n, U.getDataPtr());
    std::copy(uold, uold + n, Uo.getDataPtr());
    LocalElmtInfo info;
    info.Dt = dt; info.T = dt * timestep; info.Timestep = timestep;
    p.m_PDSystem.formResidualAndJacobian(p.m_PDMesh, p.m_Operators, p.m_ElmtSystem, info, U, Uo, p.m_Matrix,
                                         p.m_RHS);
    const int nnz = p.m_Matrix.getNNZNum();
    std::copy(p.m_Matrix.getCSRValuesPtr(), p.m_Matrix.getCSRValuesPtr() + nnz, kvals);
    std::copy(p.m_RHS.getDataPtr(), p.m_RHS.getDataPtr() + n, rhs);
}

// PDDODynamicFracElement state: bond health (family order) and the committed
// velocity; the element's body force and initial velocity setters.
void ref_dynfrac_state(void *h, double *health, double *vel)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    auto *e = dynamic_cast<PDDODynamicFracElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (!e) return;
    long long q = 0;
    for (const auto &row : e->getBondHealth())
        for (double v : row) health[q++] = v;
    std::copy(e->getVel().begin(), e->getVel().end(), vel);
}

void ref_dynfrac_set(void *h, const double *b, const double *v0)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    auto *e = dynamic_cast<PDDODynamicFracElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (!e) return;
    e->setBodyForce(b[0], b[1], b[2]);
    e->setInitialVelocity(v0[0], v0[1], v0[2]);
}

// StressCahnHilliardElement constants as the element holds them:
// C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref
void ref_stressch_consts(void *h, double *out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    auto *e = dynamic_cast<StressCahnHilliardElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (!e) return;
    const double v[11] = {e->m_C11, e->m_C12, e->m_C66, e->m_A, e->m_Kh, e->m_Ch,
                          e->m_D, e->m_Kappa, e->m_Chi, e->m_Omega, e->m_CRef};
    std::copy(v, v + 11, out);
}

// FracStressCahnHilliardElement state (valid after one ref_assemble_step):
// bond health (CSR order), committed velocity (n*dim, zeros when rho <= 0),
// the species Laplacian per bond, and k = C11, C12, C66, A, Kh, Ch, D, kappa,
// chi, Omega, cref, rho, residual stiffness, critical stretch
void ref_fsch_state(void *h, double *health, double *vel, double *lapc, double *k)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    auto *e = dynamic_cast<FracStressCahnHilliardElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (!e) return;
    long long q = 0;
    for (const auto &row : e->getBondHealth())
        for (double v : row) health[q++] = v;
    std::copy(e->getVel().begin(), e->getVel().end(), vel);
    std::copy(e->getLapC().begin(), e->getLapC().end(), lapc);
    const double v[14] = {e->m_C11, e->m_C12, e->m_C66, e->m_A, e->m_Kh, e->m_Ch, e->m_D, e->m_Kappa,
                          e->m_Chi, e->m_Omega, e->m_CRef, e->m_Rho, e->m_ResidualStiffness,
                          e->criticalStretch(p.m_PDMesh.getDataConstRef().HorizonRadius)};
    std::copy(v, v + 14, k);
}

// CH element geometry (valid after one ref_assemble): lapbond, rev
void ref_ch_geometry(void *h, double *lapbond, int *rev, double *mnode)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    auto *e = dynamic_cast<CahnHilliardElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (!e) return;
    std::copy(e->m_LapBond.begin(), e->m_LapBond.end(), lapbond);
    std::copy(e->m_Rev.begin(), e->m_Rev.end(), rev);
    if (mnode) std::copy(e->m_Mnode.begin(), e->m_Mnode.end(), mnode);
}

// Explicit run from the deck's TimeStepping block.  Returns steps, fills the
// final u, uold, health and the element's op cache (NOps per bond).
int ref_explicit(void *h, double *u, double *uold, double *health, double *opc,
                 double *sc_out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    p.m_U.setToZeros();
    p.m_InitialVelocity.setToZeros();
    p.m_ICSystem.applyInitialConditions(p.m_PDMesh, ndof, p.m_U, p.m_InitialVelocity);
    p.m_BCSystem.presetSolution(p.m_PDMesh, ndof, p.m_U);
    p.m_Uold = p.m_U;
    std::function<void(const int &, const double &)> none;
    const bool ok = p.m_TimeStepping.solveExplicitDynamics(
        p.m_PDMesh, p.m_Operators, p.m_ElmtSystem, p.m_BCSystem, p.m_PDSystem, ndof,
        p.m_U, p.m_Uold, p.m_InitialVelocity, p.m_RHS, none);
    const int n = p.m_U.getSize();
    std::copy(p.m_U.getDataPtr(), p.m_U.getDataPtr() + n, u);
    std::copy(p.m_Uold.getDataPtr(), p.m_Uold.getDataPtr() + n, uold);
    auto *e = dynamic_cast<ExplicitPDDOFracElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (e) {
        std::copy(e->m_HealthFlat.begin(), e->m_HealthFlat.end(), health);
        std::copy(e->m_OpCache.begin(), e->m_OpCache.end(), opc);
        if (sc_out) *sc_out = e->m_Sc;
    }
    return ok ? p.m_TimeStepping.m_LastSteps : -p.m_TimeStepping.m_LastSteps - 1;
}

// ExplicitPDDOFracElement::computeNodalProjection (ExplicitPDDOFracElement.cpp:
// 324-405) of every node at the current state (after ref_explicit): `name` one
// of damage / vonMises (1 value per node) or stress / strain (9 per node).
void ref_projection(void *h, const char *name, double *out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    const int n = p.m_PDMesh.getNodesNum();
    const std::string nm(name);
    const size_t nc = (nm == "stress" || nm == "strain") ? 9 : 1;
    std::vector<double> o(nc);
    auto *e = p.m_ElmtSystem.m_Elements[0].Element.get();
    for (int i = 1; i <= n; i++) {
        std::fill(o.begin(), o.end(), 0.0);
        e->computeNodalProjection(nm, p.m_PDMesh, p.m_Operators, p.m_U, i, ndof, o);
        std::copy(o.begin(), o.end(), out + (size_t)(i - 1) * nc);
    }
}

// computeNodalProjection of every node on the state u (n*ndof), nc
// components per node
void ref_projection_u(void *h, const char *name, const double *u, int nc, double *out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    const int n = p.m_PDMesh.getNodesNum();
    VectorXd U(n * ndof);
    std::copy(u, u + (size_t)n * ndof, U.getDataPtr());
    const std::string nm(name);
    std::vector<double> o((size_t)nc);
    auto *e = p.m_ElmtSystem.m_Elements[0].Element.get();
    for (int i = 1; i <= n; i++) {
        std::fill(o.begin(), o.end(), 0.0);
        e->computeNodalProjection(nm, p.m_PDMesh, p.m_Operators, U, i, ndof, o);
        std::copy(o.begin(), o.end(), out + (size_t)(i - 1) * nc);
    }
}

// PDProblem::savePDResults for the state u (n*ndof): the reference's ASCII
// VTU of the deck's OutputSystem fields, written next to `stem` (stem.json ->
// stem-PD-results-<step>.vtu)
void ref_save_vtu(void *h, const double *u, const char *stem, int step)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    std::copy(u, u + p.m_U.getSize(), p.m_U.getDataPtr());
    p.savePDResults(std::string(stem) + ".json", step, 0.0);
}

// Elastic constants of the explicit element (C11, C12, C66, rho, sc).
void ref_frac_consts(void *h, double *out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    auto *e = dynamic_cast<ExplicitPDDOFracElement *>(p.m_ElmtSystem.m_Elements[0].Element.get());
    if (!e) return;
    out[0] = e->m_C11; out[1] = e->m_C12; out[2] = e->m_C66; out[3] = e->m_Rho;
    out[4] = e->criticalStretch(p.m_PDMesh.getDataConstRef().HorizonRadius);
}

} // extern "C"

extern "C" {
// Runs a deck's whole job through the reference's own driver
// (PDProblem::run, RunPDProblem.cpp:21-: init, initial conditions, BC
// presets, then the static Newton solve or the transient loop, with the
// output the deck asks for) and copies the final U (at most cap values) and,
// for an explicit fracture deck, the element's bond health (at most hcap).
// Returns the dof count; *nhealth gets the bond count (0 for other
// elements) and *seconds the wall time of run().
long long ref_run(const char *deck, double *u_out, long long cap, double *health_out, long long hcap,
                  long long *nhealth, double *seconds)
{
    PDProblem prob;
    std::string a0 = "perix", a1 = "-i", a2 = deck;
    char *argv[3] = {a0.data(), a1.data(), a2.data()};
    const auto t0 = std::chrono::steady_clock::now();
    prob.run(3, argv);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const long long n = prob.m_U.getSize();
    if (u_out) std::copy(prob.m_U.getDataPtr(), prob.m_U.getDataPtr() + std::min(n, cap), u_out);
    long long nh = 0;
    if (prob.m_ElmtSystem.getElementsNum() == 1) {
        if (auto *e = dynamic_cast<const ExplicitPDDOFracElement *>(&prob.m_ElmtSystem.getElementByIndex(0))) {
            nh = (long long)e->m_HealthFlat.size();
            if (health_out) std::copy(e->m_HealthFlat.begin(), e->m_HealthFlat.begin() + std::min(nh, hcap), health_out);
        }
    }
    if (nhealth) *nhealth = nh;
    return n;
}

// Timed explicit run: the element's one-off build (families flattened, op
// cache, crack seeds) is triggered by a first rate evaluation and timed as
// setup; the central-difference loop of the deck's TimeStepping block is
// timed separately.  Returns steps (negative on failure).
int ref_explicit_timed(void *h, double *setup_s, double *loop_s, double *u_out)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int ndof = p.m_ElmtSystem.getMaxDofsPerNode();
    p.m_U.setToZeros();
    p.m_InitialVelocity.setToZeros();
    p.m_ICSystem.applyInitialConditions(p.m_PDMesh, ndof, p.m_U, p.m_InitialVelocity);
    p.m_BCSystem.presetSolution(p.m_PDMesh, ndof, p.m_U);
    p.m_Uold = p.m_U;
    LocalElmtInfo info;
    info.UseParallel = true;
    auto t0 = std::chrono::steady_clock::now();
    p.m_PDSystem.formExplicitDynamicsRate(p.m_PDMesh, p.m_Operators, p.m_ElmtSystem, info,
                                          p.m_U, p.m_Uold, p.m_RHS);
    auto t1 = std::chrono::steady_clock::now();
    std::function<void(const int &, const double &)> none;
    const bool ok = p.m_TimeStepping.solveExplicitDynamics(
        p.m_PDMesh, p.m_Operators, p.m_ElmtSystem, p.m_BCSystem, p.m_PDSystem, ndof,
        p.m_U, p.m_Uold, p.m_InitialVelocity, p.m_RHS, none);
    auto t2 = std::chrono::steady_clock::now();
    *setup_s = std::chrono::duration<double>(t1 - t0).count();
    *loop_s = std::chrono::duration<double>(t2 - t1).count();
    if (u_out) std::copy(p.m_U.getDataPtr(), p.m_U.getDataPtr() + p.m_U.getSize(), u_out);
    return ok ? p.m_TimeStepping.m_LastSteps : -1;
}
}

extern "C" {
// The reference's default linear solver (DefaultSolver::initSolver, then
// solveLinearSystem, DefaultSolver.cpp:428-485: RCM ordering, unified
// skyline envelope, LDU without pivoting, substitution) on the case's CSR
// pattern with caller values kvals and right-hand side rhs.  Writes x; perm
// (new -> old, n ints) and jp (skyline pointer, n+1) when not NULL.  use_rcm
// mirrors the solver's "use_rcm" parameter.  Returns 0, or 1 when the
// factorisation meets a zero pivot.
int ref_default_solve(void *h, const double *kvals, const double *rhs, int use_rcm, double *x, int *perm,
                      long long *jp)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int n = p.m_U.getSize();
    const int nnz = p.m_Matrix.getNNZNum();
    std::copy(kvals, kvals + nnz, p.m_Matrix.getCSRValuesPtr());
    VectorXd b(n), xs(n);
    std::copy(rhs, rhs + n, b.getDataPtr());
    DefaultSolver solver;
    nlohmann::ordered_json prm;
    prm["use_rcm"] = use_rcm != 0;
    solver.initSolver(p.m_Matrix, prm);
    const bool ok = solver.solveLinearSystem(p.m_Matrix, b, xs);
    std::copy(xs.getDataPtr(), xs.getDataPtr() + n, x);
    if (perm) std::copy(solver.m_Perm.begin(), solver.m_Perm.end(), perm);
    if (jp) std::copy(solver.m_JP.begin(), solver.m_JP.end(), jp);
    return ok ? 0 : 1;
}
} // extern "C"

#ifdef PERIX_CUDA_ASSEMBLE
#include "pxg_linear_solver.h"
extern "C" {
// The libpxg back-end of the reference's LinearSolver dispatcher
// (integration/pxg_linear_solver.cpp) driven exactly as LinearSolver::init
// and ::solve drive a back-end: initSolver on the case's SparseMatrix with the
// JSON parameters `params`, then solveLinearSystem with caller values kvals
// and right-hand side rhs.  Returns 1 when the solve succeeded.
int ref_pxg_linear_solve(void *h, const double *kvals, const double *rhs, const char *params, double *x)
{
    auto &p = static_cast<RefCase *>(h)->prob;
    const int n = p.m_U.getSize();
    const int nnz = p.m_Matrix.getNNZNum();
    std::copy(kvals, kvals + nnz, p.m_Matrix.getCSRValuesPtr());
    VectorXd b(n), xs(n);
    std::copy(rhs, rhs + n, b.getDataPtr());
    nlohmann::ordered_json prm = nlohmann::ordered_json::parse(params);
    PxgLinearSolver solver;
    solver.initSolver(p.m_Matrix, prm);
    const bool ok = solver.solveLinearSystem(p.m_Matrix, b, xs);
    std::copy(xs.getDataPtr(), xs.getDataPtr() + n, x);
    return ok ? 1 : 0;
}
} // extern "C"
#endif

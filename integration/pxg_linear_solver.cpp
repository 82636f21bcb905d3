// pxg_linear_solver.cpp -- PxgLinearSolver (see pxg_linear_solver.h).
//
// A maintainer adds it beside the other back-ends: one more LinearSolverType
// value and its name, a member next to m_DefaultSolver, and one branch each
// in LinearSolver::init and LinearSolver::solve (LinearSolver.cpp:72-121)
// calling initSolver / solveLinearSystem, as the cuDSS branch does.
#include "pxg_linear_solver.h"

#include <string>

#include "Utils/MessagePrinter.h"
#include "pxg.h"

namespace {
bool boolParam(const nlohmann::ordered_json &p, const char *key, bool def)
{
    return p.contains(key) ? p.at(key).get<bool>() : def;
}
double doubleParam(const nlohmann::ordered_json &p, const char *key, double def)
{
    return p.contains(key) ? p.at(key).get<double>() : def;
}
bool report(int rc, const char *where)
{
    if (rc == PXG_OK) return true;
    MessagePrinter::printErrorTxt(std::string("PxgLinearSolver: ") + where + ": " + pxg_last_error());
    return false;
}
}  // namespace

PxgLinearSolver::~PxgLinearSolver()
{
    if (m_Ctx) pxg_destroy(m_Ctx);
}

bool PxgLinearSolver::setPattern(SparseMatrix &A)
{
    m_Rows = A.getSize();
    m_NNZ = A.getNNZNum();
    return report(pxg_solve_set_pattern(m_Ctx, m_Rows, A.getCSRRowsIndexPtr(), A.getCSRColsIndexPtr()),
                  "pattern");
}

void PxgLinearSolver::initSolver(SparseMatrix &K, nlohmann::ordered_json &Parameters)
{
    // DefaultSolver::initSolver's parameters (DefaultSolver.cpp:436-441)
    m_UseRCM = boolParam(Parameters, "use_rcm", true);
    m_MaxProfileGB = doubleParam(Parameters, "max_profile_gb", 16.0);
    const int device = Parameters.contains("device") ? Parameters.at("device").get<int>() : 0;
    if (!m_Ctx && !report(pxg_create(&m_Ctx, device), "create")) {
        m_Ctx = nullptr;
        return;
    }
    report(pxg_solve_options(m_Ctx, m_UseRCM ? 1 : 0, m_MaxProfileGB, 0), "options");
    setPattern(K);
}

bool PxgLinearSolver::solveLinearSystem(SparseMatrix &A, VectorXd &b, VectorXd &x)
{
    if (!m_Ctx) return false;
    // the reference's fingerprint of a changed pattern (DefaultSolver.cpp:156-162)
    if ((A.getSize() != m_Rows || A.getNNZNum() != m_NNZ) && !setPattern(A)) return false;
    if (b.getSize() != m_Rows || x.getSize() != m_Rows) {
        MessagePrinter::printErrorTxt("PxgLinearSolver: size mismatch");
        return false;
    }
    return report(pxg_solve(m_Ctx, A.getCSRValuesPtr(), b.getDataPtr(), x.getDataPtr()), "solve");
}
